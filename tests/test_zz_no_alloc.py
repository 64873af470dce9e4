"""lts_bvh_build_static, the LBVH build that allocates nothing (its temporaries live in a static device
buffer of the library), against lts_bvh_build on the same soups: every output array bit for bit, on
two calls in a row, the second with a smaller soup (the buffer keeps the first call's contents),
in both leaf layouts. Inputs and outputs are device buffers the caller owns (allocated here through
the HIP runtime that liblts.so itself links).
"""
import ctypes as C

import numpy as np
import pytest

from lighttransportsimulation_amd import Context, load_library, scenegen

LTS_E_ARG = -1
MAX_PRIMS = 1 << 18


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime liblts.so is linked against."""

    def __init__(self):
        load_library()
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.bufs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(nbytes, 4)) == 0
        self.bufs.append(p)
        return p

    def upload(self, a):
        p = self.alloc(a.nbytes)
        assert self.L.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0  # host to device
        return p

    def download(self, p, a):
        assert self.L.hipMemcpy(a.ctypes.data_as(C.c_void_p), p, a.nbytes, 2) == 0  # device to host
        return a

    def free(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def _static_build(hip, tri: np.ndarray, mode: int):
    n = tri.shape[0]
    soup = hip.upload(np.ascontiguousarray(tri, np.float32).reshape(-1))
    shapes = dict(codes_sorted=(n, np.uint32), perm=(n, np.int32), left=(max(n - 1, 1), np.int32),
                  right=(max(n - 1, 1), np.int32), parent=(2 * n - 1, np.int32), bounds=(6 * (2 * n - 1), np.float32))
    dev = {k: hip.alloc(c * np.dtype(t).itemsize) for k, (c, t) in shapes.items()}
    sb = (C.c_float * 6)()
    rc = load_library().lts_bvh_build_static(soup, n, mode, dev["codes_sorted"], dev["perm"], dev["left"], dev["right"],
                                             dev["parent"], dev["bounds"], sb, None)
    assert rc == 0, rc
    host = {k: hip.download(dev[k], np.zeros(c, t)) for k, (c, t) in shapes.items()}
    host["scene_bounds"] = np.array(list(sb), np.float32)
    return host


@pytest.mark.gpu
def test_static_build_equals_the_allocating_build():
    with Context(0) as ctx:
        hip = _Hip()
        for n, mode, seed in ((200_000, 0, 5), (5_000, 1, 6), (1, 1, 7)):
            print(f"[allocation-free LBVH build: n {n}, mode {mode}]", flush=True)
            tri = scenegen.triangle_soup(n, seed=seed)
            ref, _ = ctx.bvh_build(tri, mode=mode)
            got = _static_build(hip, tri, mode)
            m = n - 1
            assert np.array_equal(got["codes_sorted"], ref["codes_sorted"])
            assert np.array_equal(got["perm"], ref["perm"])
            assert np.array_equal(got["parent"], ref["parent"])
            if m:
                assert np.array_equal(got["left"][:m], ref["left"]) and np.array_equal(got["right"][:m], ref["right"])
            b = got["bounds"].reshape(-1, 6)
            assert np.array_equal(b[:m].view(np.uint32), ref["int_bounds"].view(np.uint32))
            assert np.array_equal(b[m:].view(np.uint32), ref["leaf_bounds"].view(np.uint32))
            leaf_prim = np.arange(n) if mode == 0 else got["perm"]
            assert np.array_equal(leaf_prim, ref["leaf_prim"])
            assert np.array_equal(got["scene_bounds"], ref["scene_bounds"])
        hip.free()


@pytest.mark.gpu
def test_static_build_rejects_sizes_beyond_its_buffer():
    L = load_library()
    nul = C.c_void_p(0)
    one = C.c_void_p(16)  # never dereferenced: the size check comes first
    assert L.lts_bvh_build_static(one, MAX_PRIMS + 1, 0, one, one, one, one, one, one, None, nul) == LTS_E_ARG
    assert L.lts_bvh_build_static(one, 0, 0, one, one, one, one, one, one, None, nul) == LTS_E_ARG
    assert L.lts_bvh_build_static(one, 16, 2, one, one, one, one, one, one, None, nul) == LTS_E_ARG
