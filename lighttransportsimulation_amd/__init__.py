"""MI355X-native path tracer with the inputs and stages of tierChampion/LightTransportSimulation.

Python host mirror of the C ABI in include/lts.h (liblts.so, HIP for gfx950). The library is
required: importing works without it, but every call that needs it raises when it is missing.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT_DIR = os.path.dirname(PKG_DIR)
LIB_PATH = os.path.join(PKG_DIR, "liblts.so")

INTEGRATOR_MEGAKERNEL = 0
INTEGRATOR_WAVEFRONT = 1
INTEGRATOR_CACHED = 2
INTEGRATOR_BDPT = 3

# every symbol declared in include/lts.h
EXPORTED_SYMBOLS = (
    "lts_create", "lts_destroy", "lts_last_error", "lts_version", "lts_read_app_params", "lts_load_scene",
    "lts_scene_info", "lts_bvh_build", "lts_bvh_build_device", "lts_bvh_export_last", "lts_render",
    "lts_write_image", "lts_device_copy_benchmark", "lts_render_batch", "lts_nrc_configure", "lts_nrc_set_weights",
    "lts_nrc_get_weights", "lts_nrc_infer", "lts_nrc_train_step", "lts_set_render_tree", "lts_render_tree_info",
    "lts_obj_to_bin", "lts_film_save", "lts_film_load", "lts_film_resolve", "lts_set_texture_filter",
    "lts_load_image", "lts_bvh_build_static", "lts_mesh_transform", "lts_trace_rays",
)

TEXTURE_REFERENCE = 0
TEXTURE_FILTERED = 1

TREE_KARRAS = 0
TREE_PLOC = 1
TREE_COLLAPSE_EVEN = 0x100  # or-ed into the kind: even-depth 4-wide collapse

NRC_PARAMS = 21504
NRC_RAW = 16
PHASES = ("extend", "sort", "shade", "shadow", "mis", "finalize", "cache_infer", "cache_train")


class LtsError(RuntimeError):
    pass


class BvhExport(C.Structure):
    _fields_ = [("codes_sorted", C.c_void_p), ("perm", C.c_void_p), ("left", C.c_void_p), ("right", C.c_void_p),
                ("int_bounds", C.c_void_p), ("leaf_prim", C.c_void_p), ("leaf_bounds", C.c_void_p),
                ("parent", C.c_void_p), ("scene_bounds", C.c_float * 6)]


class BvhTiming(C.Structure):
    _fields_ = [("total_ms", C.c_double), ("sort_ms", C.c_double), ("bounds_ms", C.c_double), ("tree_ms", C.c_double)]


class RenderParams(C.Structure):
    _fields_ = [("spp", C.c_int32), ("max_bounce", C.c_int32), ("rr_start", C.c_int32), ("seed", C.c_uint32),
                ("stratified", C.c_int32), ("integrator", C.c_int32), ("spp_per_pass", C.c_int32),
                ("trace_block", C.c_int32), ("persistent_blocks", C.c_int32), ("sort_materials", C.c_int32),
                ("use_graph", C.c_int32), ("autotune", C.c_int32), ("profile_phases", C.c_int32),
                ("batch_group_pixels", C.c_int32), ("count_traversal", C.c_int32), ("sample_begin", C.c_int32),
                ("sample_end", C.c_int32), ("tile_rows", C.c_int32)]


class NrcConfig(C.Structure):
    _fields_ = [("precision", C.c_int32), ("batch", C.c_int32), ("query_bounce", C.c_int32), ("train_paths", C.c_int32),
                ("warmup_passes", C.c_int32), ("seed", C.c_uint32), ("lr", C.c_float), ("beta1", C.c_float),
                ("beta2", C.c_float), ("eps", C.c_float), ("train_steps", C.c_int32), ("reserved", C.c_int32 * 5)]


class Stats(C.Structure):
    _fields_ = [("render_ms", C.c_double), ("resolve_ms", C.c_double), ("kernel_avg_ms", C.c_double),
                ("launches", C.c_int64), ("samples", C.c_int64), ("msamples_per_s", C.c_double),
                ("node_visits", C.c_int64), ("tri_tests", C.c_int64), ("trace_block", C.c_int32),
                ("persistent_blocks", C.c_int32), ("sort_materials", C.c_int32), ("use_graph", C.c_int32),
                ("autotune_ms", C.c_double), ("nrc_loss", C.c_double), ("nrc_queries", C.c_int64),
                ("nrc_train_samples", C.c_int32), ("nrc_steps", C.c_int32), ("phase_ms", C.c_double * 8),
                ("phase_launches", C.c_int64 * 8), ("batch_groups", C.c_int32), ("pad", C.c_int32),
                ("ray_nodes", C.c_int64 * 3), ("ray_tris", C.c_int64 * 3), ("tile_rows", C.c_int32),
                ("spp_per_pass", C.c_int32)]


RAY_KINDS = ("path", "shadow", "mis")

_lib = None


def load_library():
    """Loads liblts.so; raises LtsError when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LtsError(f"{LIB_PATH} is missing: build it with lighttransportsimulation_amd/build.py")
    L = C.CDLL(LIB_PATH)
    vp, i32, u32 = C.c_void_p, C.c_int32, C.c_uint32
    L.lts_create.argtypes = [C.POINTER(vp), C.c_int]
    L.lts_destroy.argtypes = [vp]
    L.lts_destroy.restype = None
    L.lts_last_error.argtypes = [vp]
    L.lts_last_error.restype = C.c_char_p
    L.lts_version.restype = C.c_char_p
    L.lts_read_app_params.argtypes = [C.c_char_p] + [C.POINTER(i32)] * 6 + [C.c_char_p] * 3 + [i32]
    L.lts_load_scene.argtypes = [vp, C.c_char_p, C.c_char_p, C.c_char_p, i32, i32]
    L.lts_scene_info.argtypes = [vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)]
    L.lts_set_render_tree.argtypes = [vp, i32]
    L.lts_set_texture_filter.argtypes = [vp, i32]
    L.lts_render_tree_info.argtypes = [vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)]
    L.lts_trace_rays.argtypes = [vp, vp, i32, vp, vp]
    L.lts_bvh_build.argtypes = [vp, vp, i32, i32, C.POINTER(BvhExport), C.POINTER(BvhTiming)]
    L.lts_bvh_build_device.argtypes = [vp, vp, i32, i32, i32, C.POINTER(BvhTiming)]
    L.lts_bvh_export_last.argtypes = [vp, C.POINTER(BvhExport)]
    L.lts_render.argtypes = [vp, C.POINTER(RenderParams), vp, vp, C.POINTER(Stats)]
    L.lts_render_batch.argtypes = [C.POINTER(vp), i32, C.POINTER(RenderParams), C.POINTER(vp), C.POINTER(Stats)]
    L.lts_write_image.argtypes = [C.c_char_p, vp, i32, i32, i32]
    L.lts_obj_to_bin.argtypes = [C.c_char_p, C.POINTER(i32)]
    L.lts_load_image.argtypes = [C.c_char_p, C.POINTER(i32), C.POINTER(i32), vp, C.c_int64]
    L.lts_bvh_build_static.argtypes = [vp, i32, i32, vp, vp, vp, vp, vp, vp, vp, vp]
    L.lts_mesh_transform.argtypes = [vp, vp, vp, i32, vp, i32]
    L.lts_film_save.argtypes = [vp, C.c_char_p]
    L.lts_film_load.argtypes = [vp, C.c_char_p]
    L.lts_film_resolve.argtypes = [vp, vp]
    L.lts_device_copy_benchmark.argtypes = [vp, C.c_uint64, i32, C.POINTER(C.c_double)]
    L.lts_nrc_configure.argtypes = [vp, C.POINTER(NrcConfig)]
    L.lts_nrc_set_weights.argtypes = [vp, vp]
    L.lts_nrc_get_weights.argtypes = [vp, vp]
    L.lts_nrc_infer.argtypes = [vp, vp, i32, vp]
    L.lts_nrc_train_step.argtypes = [vp, vp, vp, i32, C.POINTER(C.c_float), vp]
    _lib = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def bvh_arrays(n: int) -> dict:
    m = max(n - 1, 0)
    return dict(codes_sorted=np.zeros(n, np.uint32), perm=np.zeros(n, np.int32), left=np.zeros(m, np.int32),
                right=np.zeros(m, np.int32), int_bounds=np.zeros((m, 6), np.float32), leaf_prim=np.zeros(n, np.int32),
                leaf_bounds=np.zeros((n, 6), np.float32), parent=np.zeros(2 * n - 1, np.int32))


def _export_struct(arrs: dict) -> BvhExport:
    e = BvhExport()
    for k in ("codes_sorted", "perm", "left", "right", "int_bounds", "leaf_prim", "leaf_bounds", "parent"):
        setattr(e, k, _ptr(arrs[k]))
    return e


class Context:
    """One HIP device context (include/lts.h: lts_create)."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = C.c_void_p()
        rc = self.lib.lts_create(C.byref(h), device)
        if rc != 0:
            raise LtsError(f"lts_create({device}) failed with {rc}")
        self.h = h
        self.width = self.height = 0

    def _check(self, rc: int, what: str):
        if rc != 0:
            raise LtsError(f"{what} failed ({rc}): {self.lib.lts_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.lts_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ------------------------------------------------------------------ scene
    def load_scene(self, project_dir: str, scene: str, subject: str | None, width: int, height: int):
        d = project_dir if project_dir.endswith("/") else project_dir + "/"
        self._check(self.lib.lts_load_scene(self.h, d.encode(), scene.encode(), (subject or "").encode(), width, height),
                    "lts_load_scene")
        self.width, self.height = width, height

    def scene_info(self) -> dict:
        a, b, c = C.c_int32(), C.c_int32(), C.c_int32()
        self._check(self.lib.lts_scene_info(self.h, C.byref(a), C.byref(b), C.byref(c)), "lts_scene_info")
        return dict(tris=a.value, lights=b.value, meshes=c.value)

    # ------------------------------------------------------------------ film checkpoint
    def film_save(self, path: str):
        """Writes the accumulated film and its sample count (include/lts.h: lts_film_save)."""
        self._check(self.lib.lts_film_save(self.h, path.encode()), "lts_film_save")

    def film_load(self, path: str):
        self._check(self.lib.lts_film_load(self.h, path.encode()), "lts_film_load")

    def film_resolve(self) -> np.ndarray:
        rgb = np.zeros((self.height, self.width, 3), np.float32)
        self._check(self.lib.lts_film_resolve(self.h, _ptr(rgb)), "lts_film_resolve")
        return rgb

    def set_texture_filter(self, mode: int):
        """Texture pipeline of the scenes loaded from now on: TEXTURE_REFERENCE (default) or TEXTURE_FILTERED."""
        self._check(self.lib.lts_set_texture_filter(self.h, mode), "lts_set_texture_filter")

    def set_render_tree(self, kind: int):
        """Render tree of the scenes loaded from now on: TREE_PLOC (default) or TREE_KARRAS, with
        TREE_COLLAPSE_EVEN or-ed in for the even-depth 4-wide collapse instead of the greedy one."""
        self._check(self.lib.lts_set_render_tree(self.h, kind), "lts_set_render_tree")

    def trace_rays(self, origins: np.ndarray, directions: np.ndarray):
        """Closest hits of rays in the loaded scene (include/lts.h: lts_trace_rays): returns
        (prim [n] int32, -1 for a miss; t [n] float32)."""
        o = np.asarray(origins, np.float32).reshape(-1, 3)
        d = np.asarray(directions, np.float32).reshape(-1, 3)
        rays = np.ascontiguousarray(np.concatenate([o, d], axis=1))
        n = rays.shape[0]
        prim = np.zeros(n, np.int32)
        t = np.zeros(n, np.float32)
        self._check(self.lib.lts_trace_rays(self.h, _ptr(rays), n, _ptr(prim), _ptr(t)), "lts_trace_rays")
        return prim, t

    def render_tree_info(self) -> dict:
        a, b, c = C.c_int32(), C.c_int32(), C.c_int32()
        self._check(self.lib.lts_render_tree_info(self.h, C.byref(a), C.byref(b), C.byref(c)), "lts_render_tree_info")
        return dict(kind="ploc" if (a.value & ~TREE_COLLAPSE_EVEN) == TREE_PLOC else "karras",
                    collapse="even" if a.value & TREE_COLLAPSE_EVEN else "greedy", depth=b.value, nodes4=c.value)

    # ------------------------------------------------------------------ LBVH
    def bvh_build(self, tri: np.ndarray, mode: int = 0, export: bool = True):
        tri = np.ascontiguousarray(tri, dtype=np.float32).reshape(-1, 9)
        n = tri.shape[0]
        arrs = bvh_arrays(n) if export else None
        e = _export_struct(arrs) if export else None
        t = BvhTiming()
        self._check(self.lib.lts_bvh_build(self.h, _ptr(tri), n, mode, C.byref(e) if export else None, C.byref(t)),
                    "lts_bvh_build")
        if export:
            arrs["scene_bounds"] = np.array(list(e.scene_bounds), np.float32)
        return arrs, dict(total_ms=t.total_ms, sort_ms=t.sort_ms, bounds_ms=t.bounds_ms, tree_ms=t.tree_ms)

    def bvh_build_device(self, dev_ptr, n: int, mode: int, repeat: int = 1) -> dict:
        """Builds from a device-resident soup (dev_ptr None: the soup of the last bvh_build)."""
        t = BvhTiming()
        self._check(self.lib.lts_bvh_build_device(self.h, None if dev_ptr is None else C.c_void_p(dev_ptr), n, mode,
                                                  repeat, C.byref(t)),
                    "lts_bvh_build_device")
        return dict(total_ms=t.total_ms, sort_ms=t.sort_ms, bounds_ms=t.bounds_ms, tree_ms=t.tree_ms)

    def mesh_transform(self, subject, points, normals):
        """World transform of a mesh on the GPU (include/lts.h: lts_mesh_transform). `subject` is the
        scene file's 8 subject tokens (translation, axis, angle in radians, scale); returns the
        transformed (points, normals) as new [n][3] float32 arrays."""
        sub = np.ascontiguousarray(subject, np.float32).reshape(8)
        pts = np.array(points, np.float32, copy=True).reshape(-1, 3)
        nrm = np.array(normals, np.float32, copy=True).reshape(-1, 3)
        self._check(self.lib.lts_mesh_transform(self.h, _ptr(sub), _ptr(pts), pts.shape[0], _ptr(nrm), nrm.shape[0]),
                    "lts_mesh_transform")
        return pts, nrm

    def bvh_export_last(self, n: int) -> dict:
        arrs = bvh_arrays(n)
        e = _export_struct(arrs)
        self._check(self.lib.lts_bvh_export_last(self.h, C.byref(e)), "lts_bvh_export_last")
        arrs["scene_bounds"] = np.array(list(e.scene_bounds), np.float32)
        return arrs

    def copy_benchmark(self, nbytes: int, repeat: int = 10) -> float:
        """Mean milliseconds of a plain device copy of nbytes (HIP events)."""
        ms = C.c_double()
        self._check(self.lib.lts_device_copy_benchmark(self.h, nbytes, repeat, C.byref(ms)), "lts_device_copy_benchmark")
        return ms.value

    # ------------------------------------------------------------------ radiance cache
    def nrc_configure(self, precision: str = "bf16", batch: int = 0, query_bounce: int = 0, train_paths: int = 0,
                      warmup_passes: int = 0, seed: int = 0, lr: float = 0.0, beta1: float = 0.0, beta2: float = 0.0,
                      eps: float = 0.0, train_steps: int = 0):
        """(Re)initialises the cache (include/lts.h: lts_nrc_configure); zero arguments take the defaults."""
        cfg = NrcConfig(0 if precision == "bf16" else 1, batch, query_bounce, train_paths, warmup_passes, seed, lr, beta1,
                        beta2, eps, train_steps)
        self._check(self.lib.lts_nrc_configure(self.h, C.byref(cfg)), "lts_nrc_configure")

    def nrc_set_weights(self, params: np.ndarray):
        params = np.ascontiguousarray(params, np.float32)
        if params.size != NRC_PARAMS:
            raise LtsError(f"expected {NRC_PARAMS} parameters")
        self._check(self.lib.lts_nrc_set_weights(self.h, _ptr(params)), "lts_nrc_set_weights")

    def nrc_get_weights(self) -> np.ndarray:
        w = np.zeros(NRC_PARAMS, np.float32)
        self._check(self.lib.lts_nrc_get_weights(self.h, _ptr(w)), "lts_nrc_get_weights")
        return w

    def nrc_infer(self, raw: np.ndarray) -> np.ndarray:
        raw = np.ascontiguousarray(raw, np.float32).reshape(-1, NRC_RAW)
        out = np.zeros((raw.shape[0], 3), np.float32)
        self._check(self.lib.lts_nrc_infer(self.h, _ptr(raw), raw.shape[0], _ptr(out)), "lts_nrc_infer")
        return out

    def nrc_train_step(self, raw: np.ndarray, target: np.ndarray, want_grad: bool = False):
        """One Adam step; returns (loss before the step, gradient or None)."""
        raw = np.ascontiguousarray(raw, np.float32).reshape(-1, NRC_RAW)
        target = np.ascontiguousarray(target, np.float32).reshape(-1, 3)
        loss = C.c_float()
        grad = np.zeros(NRC_PARAMS, np.float32) if want_grad else None
        self._check(self.lib.lts_nrc_train_step(self.h, _ptr(raw), _ptr(target), raw.shape[0], C.byref(loss), _ptr(grad)),
                    "lts_nrc_train_step")
        return loss.value, grad

    # ------------------------------------------------------------------ render
    def render(self, spp: int, max_bounce: int, rr_start: int, seed: int, stratified: int = 0,
               integrator: int = INTEGRATOR_MEGAKERNEL, spp_per_pass: int = 0, want_image: bool = True,
               want_samples: bool = False, trace_block: int = 0, persistent_blocks: int = 0, sort_materials: int = -1,
               use_graph: int = -1, autotune: int = 0, profile_phases: int = 0, count_traversal: int = 0,
               sample_begin: int = 0, sample_end: int = 0, tile_rows: int = 0):
        """Renders the loaded scene (include/lts.h: lts_render). Returns (rgb, samples, stats): rgb is
        [H][W][3] in getRender row order, samples [spp][H][W][3] per-sample radiance when asked for."""
        p = render_params(spp, max_bounce, rr_start, seed, stratified, integrator, spp_per_pass, trace_block,
                          persistent_blocks, sort_materials, use_graph, autotune, profile_phases,
                          count_traversal=count_traversal, sample_begin=sample_begin, sample_end=sample_end,
                          tile_rows=tile_rows)
        rgb = np.zeros((self.height, self.width, 3), np.float32) if want_image else None
        n_smp = (sample_end or spp) - sample_begin
        smp = np.zeros((n_smp, self.height, self.width, 3), np.float32) if want_samples else None
        st = Stats()
        self._check(self.lib.lts_render(self.h, C.byref(p), _ptr(rgb), _ptr(smp), C.byref(st)), "lts_render")
        return rgb, smp, _stats_dict(st)


def render_params(spp, max_bounce, rr_start, seed, stratified=0, integrator=INTEGRATOR_MEGAKERNEL, spp_per_pass=0,
                  trace_block=0, persistent_blocks=0, sort_materials=-1, use_graph=-1, autotune=0,
                  profile_phases=0, batch_group_pixels=0, count_traversal=0, sample_begin=0,
                  sample_end=0, tile_rows=0) -> RenderParams:
    return RenderParams(spp, max_bounce, rr_start, seed, stratified, integrator, spp_per_pass, trace_block,
                        persistent_blocks, sort_materials, use_graph, autotune, profile_phases, batch_group_pixels,
                        count_traversal, sample_begin, sample_end, tile_rows)


def _stats_dict(st: Stats) -> dict:
    return dict(render_ms=st.render_ms, resolve_ms=st.resolve_ms, kernel_avg_ms=st.kernel_avg_ms,
                launches=st.launches, samples=st.samples, msamples_per_s=st.msamples_per_s,
                node_visits=st.node_visits, tri_tests=st.tri_tests, trace_block=st.trace_block,
                persistent_blocks=st.persistent_blocks, sort_materials=st.sort_materials, use_graph=st.use_graph,
                autotune_ms=st.autotune_ms, nrc_loss=st.nrc_loss, nrc_queries=st.nrc_queries,
                nrc_train_samples=st.nrc_train_samples, nrc_steps=st.nrc_steps,
                phase_ms={k: st.phase_ms[i] for i, k in enumerate(PHASES)},
                phase_launches={k: st.phase_launches[i] for i, k in enumerate(PHASES)}, batch_groups=st.batch_groups,
                ray_nodes={k: st.ray_nodes[i] for i, k in enumerate(RAY_KINDS)},
                ray_tris={k: st.ray_tris[i] for i, k in enumerate(RAY_KINDS)}, tile_rows=st.tile_rows,
                spp_per_pass=st.spp_per_pass)


def render_batch(contexts, spp: int, max_bounce: int, rr_start: int, seed: int, stratified: int = 0,
                 spp_per_pass: int = 0, trace_block: int = 0, persistent_blocks: int = 0, sort_materials: int = -1,
                 use_graph: int = -1, autotune: int = 0, want_images: bool = True, group_pixels: int = 0):
    """Renders the scenes of several contexts together with the wavefront integrator
    (include/lts.h: lts_render_batch). Returns ([rgb per scene], stats)."""
    ctxs = list(contexts)
    if not ctxs:
        raise LtsError("render_batch needs at least one context")
    L = ctxs[0].lib
    p = render_params(spp, max_bounce, rr_start, seed, stratified, INTEGRATOR_WAVEFRONT, spp_per_pass, trace_block,
                      persistent_blocks, sort_materials, use_graph, autotune, 0, group_pixels)
    imgs = [np.zeros((c.height, c.width, 3), np.float32) if want_images else None for c in ctxs]
    handles = (C.c_void_p * len(ctxs))(*[c.h.value for c in ctxs])
    outs = (C.c_void_p * len(ctxs))(*[None if a is None else a.ctypes.data for a in imgs])
    st = Stats()
    ctxs[0]._check(L.lts_render_batch(handles, len(ctxs), C.byref(p), outs, C.byref(st)), "lts_render_batch")
    return imgs, _stats_dict(st)


def read_app_params(path: str) -> dict:
    L = load_library()
    v = [C.c_int32() for _ in range(6)]
    bufs = [C.create_string_buffer(1024) for _ in range(3)]
    rc = L.lts_read_app_params(path.encode(), *[C.byref(x) for x in v], *bufs, 1024)
    if rc != 0:
        raise LtsError(f"lts_read_app_params({path}) failed ({rc})")
    w, h, spp, b, rr, fmt = (x.value for x in v)
    return dict(width=w, height=h, spp=spp, max_bounce=b, rr_start=rr, format=fmt, output=bufs[0].value.decode(),
                scene=bufs[1].value.decode(), subject=bufs[2].value.decode())


def obj_to_bin(mesh_stem: str) -> int:
    """Converts <mesh_stem>.obj to <mesh_stem>.bin (include/lts.h: lts_obj_to_bin); returns the
    triangle count."""
    n = C.c_int32()
    L = load_library()
    rc = L.lts_obj_to_bin(mesh_stem.encode(), C.byref(n))
    if rc != 0:
        raise LtsError(f"lts_obj_to_bin({mesh_stem}) failed ({rc}): {L.lts_last_error(None).decode()}")
    return n.value


def write_image(path: str, rgb: np.ndarray, fmt: int) -> None:
    rgb = np.ascontiguousarray(rgb, dtype=np.float32)
    h, w, _ = rgb.shape
    rc = load_library().lts_write_image(path.encode(), _ptr(rgb), w, h, fmt)
    if rc != 0:
        raise LtsError(f"lts_write_image({path}) failed ({rc})")


def load_image(path: str) -> np.ndarray:
    """Decodes an 8-bit image to linear RGB, bottom row first (include/lts.h: lts_load_image,
    the texture loader's decode); returns an [h][w][3] float32 array."""
    L = load_library()
    w, h = C.c_int32(), C.c_int32()
    rc = L.lts_load_image(path.encode(), C.byref(w), C.byref(h), None, 0)
    if rc != 0:
        raise LtsError(f"lts_load_image({path}) failed ({rc}): {L.lts_last_error(None).decode()}")
    out = np.zeros((h.value, w.value, 3), np.float32)
    rc = L.lts_load_image(path.encode(), C.byref(w), C.byref(h), _ptr(out), out.size)
    if rc != 0:
        raise LtsError(f"lts_load_image({path}) failed ({rc})")
    return out
