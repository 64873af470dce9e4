"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

The oracle is test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg load it, as the checker and as the host-core baseline.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "_build", "liboracle.so")

_lib = None


class BvhOut(C.Structure):
    _fields_ = [("codes_sorted", C.POINTER(C.c_uint32)), ("perm", C.POINTER(C.c_int32)),
                ("left", C.POINTER(C.c_int32)), ("right", C.POINTER(C.c_int32)),
                ("int_bounds", C.POINTER(C.c_float)), ("leaf_prim", C.POINTER(C.c_int32)),
                ("leaf_bounds", C.POINTER(C.c_float)), ("parent", C.POINTER(C.c_int32)),
                ("scene_bounds", C.c_float * 6)]


class RenderParams(C.Structure):
    _fields_ = [("spp", C.c_int32), ("max_bounce", C.c_int32), ("rr_start", C.c_int32), ("seed", C.c_uint32),
                ("stratified", C.c_int32), ("threads", C.c_int32), ("row_begin", C.c_int32), ("row_end", C.c_int32)]


class RenderStats(C.Structure):
    _fields_ = [("seconds", C.c_double), ("samples", C.c_int64), ("node_visits", C.c_int64),
                ("tri_tests", C.c_int64), ("rays", C.c_int64), ("threads", C.c_int32)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        L.oracle_morton3.restype = C.c_uint32
        L.oracle_morton3.argtypes = [C.c_float, C.c_float, C.c_float]
        L.oracle_lbvh_build.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(BvhOut)]
        L.oracle_philox.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.oracle_sample.restype = C.c_float
        L.oracle_sample.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_int32]
        L.oracle_load_scene.restype = C.c_void_p
        L.oracle_load_scene.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int32, C.c_int32]
        L.oracle_free_scene.argtypes = [C.c_void_p]
        L.oracle_last_error.restype = C.c_char_p
        L.oracle_scene_info.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        L.oracle_render.argtypes = [C.c_void_p, C.POINTER(RenderParams), C.c_void_p, C.c_void_p, C.POINTER(RenderStats)]
        L.oracle_intersect.restype = C.c_int32
        L.oracle_intersect.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_float, C.POINTER(C.c_float)]
        L.oracle_occluded.restype = C.c_int32
        L.oracle_occluded.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_float]
        L.oracle_filter_table.argtypes = [C.c_void_p]
        L.oracle_set_texture_filter.argtypes = [C.c_int32]
        L.oracle_mip_lookup.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_float, C.c_float,
                                        C.c_float, C.c_void_p]
        L.oracle_nrc_param_count.restype = C.c_int32
        L.oracle_nrc_encode.argtypes = [C.c_void_p, C.c_void_p]
        L.oracle_nrc_forward.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_void_p]
        L.oracle_nrc_grad.restype = C.c_double
        L.oracle_nrc_grad.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p,
                                      C.c_void_p]
        L.oracle_camera_ray.argtypes = [C.c_void_p, C.c_float, C.c_float, C.c_int32, C.c_void_p]
        L.oracle_primary_hit.restype = C.c_int32
        L.oracle_primary_hit.argtypes = [C.c_void_p, C.c_float, C.c_float, C.c_int32, C.c_void_p]
        L.oracle_infinite_sample.restype = C.c_int32
        L.oracle_infinite_sample.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.oracle_infinite_pdf.restype = C.c_float
        L.oracle_infinite_pdf.argtypes = [C.c_void_p]
        L.oracle_light_info.restype = C.c_int32
        L.oracle_light_info.argtypes = [C.c_void_p, C.c_int32, C.c_void_p]
        L.oracle_sample_light.restype = C.c_int32
        L.oracle_sample_light.argtypes = [C.c_void_p, C.c_float, C.POINTER(C.c_float)]
        L.oracle_film_gather.argtypes = [C.c_int32] * 3 + [C.c_void_p] * 3
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def infinite_pdf(wi) -> float:
    w = np.asarray(wi, np.float32)
    return lib().oracle_infinite_pdf(_p(w))


def set_texture_filter(filtered: bool):
    """Texture pipeline of OracleScene objects created afterwards (False: the reference's)."""
    lib().oracle_set_texture_filter(int(filtered))


def mip_lookup(img: np.ndarray, s: float, t: float, width: float, wrap: int = 1, filtered: bool = True) -> np.ndarray:
    """Pyramid of a linear [h][w][3] image and one lookup (Mipmap::lookUp, Mipmap.cuh:55-72)."""
    img = np.ascontiguousarray(img, np.float32)
    out = np.zeros(3, np.float32)
    lib().oracle_mip_lookup(_p(img), img.shape[1], img.shape[0], wrap, int(filtered), s, t, width, _p(out))
    return out


def lbvh_build(tri: np.ndarray, mode: int = 0) -> dict:
    tri = np.ascontiguousarray(tri, dtype=np.float32).reshape(-1, 9)
    n = tri.shape[0]
    m = max(n - 1, 0)
    out = dict(codes_sorted=np.zeros(n, np.uint32), perm=np.zeros(n, np.int32), left=np.zeros(m, np.int32),
               right=np.zeros(m, np.int32), int_bounds=np.zeros((m, 6), np.float32),
               leaf_prim=np.zeros(n, np.int32), leaf_bounds=np.zeros((n, 6), np.float32),
               parent=np.zeros(2 * n - 1, np.int32))
    s = BvhOut()
    for k, v in out.items():
        setattr(s, k, v.ctypes.data_as(C.POINTER(C.c_uint32 if v.dtype == np.uint32 else
                                                  C.c_int32 if v.dtype == np.int32 else C.c_float)))
    rc = lib().oracle_lbvh_build(_p(tri), n, mode, C.byref(s))
    if rc != 0:
        raise RuntimeError(f"oracle_lbvh_build failed ({rc})")
    out["scene_bounds"] = np.array(list(s.scene_bounds), np.float32)
    return out


def philox(ctr, key) -> np.ndarray:
    c = np.asarray(ctr, np.uint32)
    k = np.asarray(key, np.uint32)
    o = np.zeros(4, np.uint32)
    lib().oracle_philox(_p(c), _p(k), _p(o))
    return o


class OracleScene:
    def __init__(self, project_dir: str, scene: str, subject: str, width: int, height: int):
        d = project_dir if project_dir.endswith("/") else project_dir + "/"
        h = lib().oracle_load_scene(d.encode(), scene.encode(), (subject or "").encode(), width, height)
        if not h:
            raise RuntimeError("oracle_load_scene: " + lib().oracle_last_error().decode())
        self.h = h
        self.width, self.height = width, height

    def info(self):
        a, b, c = C.c_int32(), C.c_int32(), C.c_int32()
        lib().oracle_scene_info(self.h, C.byref(a), C.byref(b), C.byref(c))
        return dict(tris=a.value, lights=b.value, meshes=c.value)

    def render(self, spp, max_bounce, rr_start, seed, stratified=0, threads=0, rows=None, want_samples=False):
        p = RenderParams(spp, max_bounce, rr_start, seed, stratified, threads,
                         rows[0] if rows else 0, rows[1] if rows else 0)
        st = RenderStats()
        rgb = np.zeros((self.height, self.width, 3), np.float32) if rows is None else None
        samples = np.zeros((spp, self.height, self.width, 3), np.float32) if want_samples else None
        rc = lib().oracle_render(self.h, C.byref(p), _p(rgb) if rgb is not None else None,
                                 _p(samples) if samples is not None else None, C.byref(st))
        if rc != 0:
            raise RuntimeError("oracle_render: " + lib().oracle_last_error().decode())
        stats = dict(seconds=st.seconds, samples=st.samples, node_visits=st.node_visits, tri_tests=st.tri_tests,
                     rays=st.rays, threads=st.threads)
        return rgb, samples, stats

    def intersect(self, o, d, tmax=np.inf):
        t = C.c_float()
        oo, dd = np.asarray(o, np.float32), np.asarray(d, np.float32)
        prim = lib().oracle_intersect(self.h, _p(oo), _p(dd), np.float32(tmax), C.byref(t))
        return prim, t.value

    def occluded(self, o, d, tmax):
        oo, dd = np.asarray(o, np.float32), np.asarray(d, np.float32)
        return bool(lib().oracle_occluded(self.h, _p(oo), _p(dd), np.float32(tmax)))

    def camera_ray(self, px, py, spp):
        out = np.zeros(18, np.float32)
        lib().oracle_camera_ray(self.h, np.float32(px), np.float32(py), spp, _p(out))
        return out.reshape(6, 3)

    def primary_hit(self, px, py, spp):
        out = np.zeros(18, np.float32)
        tri = lib().oracle_primary_hit(self.h, np.float32(px), np.float32(py), spp, _p(out))
        return tri, dict(p=out[0:3], n=out[3:6], uv=out[6:8], dpdx=out[8:11], dpdy=out[11:14], dudx=out[14],
                         dvdx=out[15], dudy=out[16], dvdy=out[17])

    def infinite_sample(self, p, u):
        out = np.zeros(4, np.float32)
        pp, uu = np.asarray(p, np.float32), np.asarray(u, np.float32)
        if lib().oracle_infinite_sample(self.h, _p(pp), _p(uu), _p(out)) != 0:
            raise RuntimeError("scene has no infinite light")
        return out[:3], float(out[3])

    def light_info(self, l):
        out = np.zeros(8, np.float32)
        if lib().oracle_light_info(self.h, l, _p(out)) != 0:
            raise IndexError(l)
        return dict(kind=int(out[0]), power=float(out[1]), center=out[2:5].copy(), radius=float(out[5]),
                    area=float(out[6]), tri=int(out[7]))

    def sample_light(self, u):
        pdf = C.c_float()
        k = lib().oracle_sample_light(self.h, np.float32(u), C.byref(pdf))
        return k, pdf.value

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_free_scene(self.h)
            self.h = None


# ---------------------------------------------------------------- learned radiance cache network
NRC_ROUNDING = {"none": 0, "bf16": 1, "fp16": 2}


def nrc_param_count() -> int:
    return lib().oracle_nrc_param_count()


def nrc_encode(raw: np.ndarray) -> np.ndarray:
    raw = np.ascontiguousarray(raw, np.float32).reshape(-1, 16)
    out = np.zeros((raw.shape[0], 64), np.float32)
    for i in range(raw.shape[0]):
        lib().oracle_nrc_encode(_p(raw[i]), _p(out[i]))
    return out


def nrc_forward(params: np.ndarray, raw: np.ndarray, rounding: str = "none") -> np.ndarray:
    params = np.ascontiguousarray(params, np.float32)
    raw = np.ascontiguousarray(raw, np.float32).reshape(-1, 16)
    y = np.zeros((raw.shape[0], 3), np.float64)
    lib().oracle_nrc_forward(_p(params), _p(raw), raw.shape[0], NRC_ROUNDING[rounding], _p(y))
    return y


def nrc_grad(params, raw, target, rounding: str = "none", denom_ref=None, want_grad=True):
    """Returns (loss, grad or None, y) of the relative L2 loss with stop-gradient denominators."""
    params = np.ascontiguousarray(params, np.float32)
    raw = np.ascontiguousarray(raw, np.float32).reshape(-1, 16)
    target = np.ascontiguousarray(target, np.float32).reshape(-1, 3)
    n = raw.shape[0]
    grad = np.zeros(params.size, np.float64) if want_grad else None
    y = np.zeros((n, 3), np.float64)
    d = None if denom_ref is None else np.ascontiguousarray(denom_ref, np.float64)
    loss = lib().oracle_nrc_grad(_p(params), _p(raw), _p(target), n, NRC_ROUNDING[rounding],
                                 None if d is None else _p(d), None if grad is None else _p(grad), _p(y))
    return loss, grad, y
