"""Argument checks of the C-ABI that run before any HIP call, so they hold without a GPU: every
entry point that takes a context returns LTS_E_ARG for a null one, and the allocation-free build
rejects sizes beyond its static buffer. Messages written without a context are kept per thread
and read back with lts_last_error(NULL)."""
import ctypes
import os

import pytest

from lighttransportsimulation_amd import LIB_PATH

LTS_E_ARG = -1


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB_PATH):
        pytest.fail(f"{LIB_PATH} missing: run __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    L.lts_last_error.restype = ctypes.c_char_p
    L.lts_last_error.argtypes = [ctypes.c_void_p]
    return L


N = None  # a null pointer through ctypes
CALLS = {
    "lts_load_scene": (N, b"d/", b"s", b"s", 8, 8),
    "lts_scene_info": (N, N, N, N),
    "lts_set_render_tree": (N, 0),
    "lts_set_texture_filter": (N, 0),
    "lts_trace_rays": (N, N, 4, N, N),
    "lts_render_tree_info": (N, N, N, N),
    "lts_device_copy_benchmark": (N, ctypes.c_uint64(1 << 20), 1, N),
    "lts_bvh_build": (N, N, 16, 0, N, N),
    "lts_bvh_build_device": (N, N, 16, 0, 1, N),
    "lts_mesh_transform": (N, b"", N, 0, N, 0),
    "lts_bvh_export_last": (N, N),
    "lts_render": (N, N, N, N, N),
    "lts_render_batch": (N, 1, N, N, N),
    "lts_nrc_configure": (N, N),
    "lts_nrc_set_weights": (N, N),
    "lts_nrc_get_weights": (N, N),
    "lts_nrc_infer": (N, N, 0, N),
    "lts_nrc_train_step": (N, N, N, 0, N, N),
    "lts_film_save": (N, b"x"),
    "lts_film_load": (N, b"x"),
    "lts_film_resolve": (N, N),
}


@pytest.mark.parametrize("name", sorted(CALLS))
def test_null_context_is_an_argument_error(lib, name):
    assert getattr(lib, name)(*CALLS[name]) == LTS_E_ARG


def test_messages_without_a_context_are_readable(lib):
    assert lib.lts_trace_rays(N, N, 4, N, N) == LTS_E_ARG
    assert lib.lts_last_error(N) == b"bad trace_rays arguments"
    assert lib.lts_nrc_infer(N, N, -1, N) == LTS_E_ARG
    assert lib.lts_last_error(N) == b"bad nrc_infer arguments"


def test_create_rejects_a_null_out_pointer(lib):
    assert lib.lts_create(N, 0) == LTS_E_ARG


@pytest.mark.parametrize("n", [0, -1, (1 << 18) + 1, 1 << 30])
def test_static_build_rejects_sizes_outside_its_buffer_without_a_device(lib, n):
    """The allocation-free build checks n against LTS_BVH_STATIC_MAX_PRIMS (2^18) before it asks
    HIP for the current device; the pointers here are never dereferenced."""
    fake = ctypes.c_void_p(16)
    args = [fake, n, 0] + [fake] * 7 + [N]
    assert lib.lts_bvh_build_static(*args) == LTS_E_ARG


def test_static_build_rejects_an_unknown_mode(lib):
    fake = ctypes.c_void_p(16)
    assert lib.lts_bvh_build_static(fake, 64, 2, *([fake] * 7), N) == LTS_E_ARG
