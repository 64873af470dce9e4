"""Throughput of the path tracer on one MI355X, as one JSON line.

A step is one full frame of the workload (all samples of every pixel) through the wavefront
integrator (hipGraph-replayed passes, autotuned in the warm-up) and the film gather, with the scene
and BVH already resident in HBM. Workloads are generated from fixed seeds in the reference's own
file formats (lighttransportsimulation_amd/scenegen.py).

Extra fields:
  roofline      the traversal kernels against HBM bandwidth: algorithmic bytes are the BVH node
                tests (32 B each) and triangle tests (36 B each) that the reference's traversal
                performs for the same rays, measured per sample by the CPU oracle on the bounded
                sample below, times the samples of a frame, over the traversal kernels' device time
                per frame from HIP events recorded between the kernels of the replayed graphs.
  cpu_baseline  the oracle (the reference integrator restated in C++) on all host cores over a
                bounded band of rows of the same frame.
  quality       (1080p) equal-time relMSE of the cached integrator (learned radiance cache) and the
                plain wavefront integrator against a 4096-spp plain reference, and the cache
                network's MFMA throughput.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Msamples/s & equal-time relMSE: cached vs plain path tracer, 1 MI355X"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    "displaced_1080p": dict(desc="seeded 1M-triangle FBM-displaced sphere on a ground quad, matte/glass/microfacet "
                                 "sectors, FBM/wrinkled/marble/2048^2 image textures, infinite light 0.5 + 2 area "
                                 "lights, 1920x1080, 64 spp, 8 bounces, RR from 3", gen="displaced", width=1920,
                            height=1080, spp=64, bounces=8, rr=3, stratified=0),
    "lbvh_1m": dict(desc="LBVH build only: 1,048,576 fp32 triangles, centroids U[0,1)^3, log-normal(-5, 0.7) edges, "
                         "random orientations; reference leaf layout", gen="soup", n=1 << 20),
    "lbvh_8m": dict(desc="LBVH build only: 8,388,608 fp32 triangles, centroids U[0,1)^3, log-normal(-5, 0.7) edges, "
                         "random orientations; reference leaf layout", gen="soup", n=1 << 23),
    "batch64": dict(desc="64 seeded scenes in one wavefront pass sequence: FBM-displaced spheres and tori, triangles "
                         "log-uniform over 10k-2M, square resolutions from {128, 256, 512, 1024}, 1-4 area lights, "
                         "matte/glass/microfacet with FBM/wrinkled/marble/image textures; 32 spp, 8 bounces, RR from 3",
                    gen="batch", n_scenes=64, spp=32, bounces=8, rr=3),
    "cornell_256": dict(desc="seeded icosphere (20,480 tris) in a Cornell box, 256x256, 16 spp stratified, "
                             "5 bounces, RR from 3", gen="cornell", width=256, height=256, spp=16, bounces=5, rr=3,
                        stratified=1),
}


def make_workload(name: str, d: str) -> dict:
    from lighttransportsimulation_amd import scenegen
    w = WORKLOADS[name]
    if w["gen"] == "displaced":
        cfg = scenegen.gen_displaced_sphere(d, width=w["width"], height=w["height"], spp=w["spp"], bounces=w["bounces"],
                                            rr=w["rr"])
    elif w["gen"] == "cornell":
        cfg = scenegen.gen_cornell_sphere(d, width=w["width"], height=w["height"], spp=w["spp"], bounces=w["bounces"],
                                          rr=w["rr"])
    else:
        raise ValueError(name)
    cfg.update(width=w["width"], height=w["height"], spp=w["spp"], max_bounce=w["bounces"], rr_start=w["rr"],
               stratified=w["stratified"])
    return cfg


def cpu_baseline(d: str, cfg: dict, budget_s: float) -> dict:
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    sc = oracle_lib.OracleScene(d, cfg["scene"], cfg["subject"], cfg["width"], cfg["height"])
    H = cfg["height"]
    r0 = H // 2
    _, _, st = sc.render(cfg["spp"], cfg["max_bounce"], cfg["rr_start"], cfg["seed"], cfg["stratified"],
                         rows=(r0, min(H, r0 + 2)))
    per_row = max(st["seconds"] / 2, 1e-4)
    rows = int(max(2, min(H - r0, budget_s / per_row)))
    rows = min(rows, H)
    r0 = max(0, (H - rows) // 2)
    _, _, st = sc.render(cfg["spp"], cfg["max_bounce"], cfg["rr_start"], cfg["seed"], cfg["stratified"],
                         rows=(r0, r0 + rows))
    return dict(value=st["samples"] / st["seconds"] / 1e6, unit="Msamples/s", cores=st["threads"], kind="port",
                sample=f"rows {r0}-{r0 + rows - 1} of {cfg['width']}x{H} at {cfg['spp']} spp "
                       f"({st['samples']} samples, {st['seconds']:.1f} s)",
                node_visits_per_sample=st["node_visits"] / st["samples"],
                tri_tests_per_sample=st["tri_tests"] / st["samples"])


# Algorithmic HBM bytes of one LBVH build per primitive: read the soup (36), write code + id (8),
# four 8-bit radix passes over (code, id) pairs (16 each), write the node and leaf records (60).
LBVH_BYTES_PER_PRIM = 36 + 8 + 16 * 4 + 60


def run_lbvh(args) -> dict:
    import ctypes
    from lighttransportsimulation_amd import Context, scenegen
    w = WORKLOADS[args.workload]
    n = w["n"]
    tri = scenegen.triangle_soup(n, seed=3)
    with Context(0) as ctx:
        ctx.bvh_build(tri, mode=0, export=False)  # uploads the soup once; it stays resident
        if args.warmup:
            ctx.bvh_build_device(None, n, 0, repeat=args.warmup)
        t0 = time.perf_counter()
        t = ctx.bvh_build_device(None, n, 0, repeat=args.steps)
        wall = time.perf_counter() - t0
        copy_bytes = 1 << 30
        copy_ms = ctx.copy_benchmark(copy_bytes, 10)
    value = n / (t["total_ms"] * 1e3)
    achieved = LBVH_BYTES_PER_PRIM * n / (t["total_ms"] * 1e-3) / 1e9
    out = {"metric": "LBVH build Mprims/s, 1 MI355X", "value": round(value, 2), "unit": "Mprims/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(t["total_ms"], 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": w["desc"], "primitives": n, "wall_ms_per_step": round(wall * 1e3 / args.steps, 3),
                      "stage_ms": {k: round(v, 4) for k, v in t.items()}},
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": "whole build",
                        "algorithmic_bytes_per_prim": LBVH_BYTES_PER_PRIM,
                        "copy_kernel_GBps": round(2 * copy_bytes / (copy_ms * 1e-3) / 1e9, 1)}}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib
        m = min(n, 1 << 20)
        sub = tri[:m]
        t1 = time.perf_counter()
        oracle_lib.lbvh_build(sub, 0)
        sec = time.perf_counter() - t1
        out["cpu_baseline"] = {"value": round(m / sec / 1e6, 3), "unit": "Mprims/s", "cores": 1, "kind": "port",
                               "sample": f"first {m} triangles of the same soup ({sec:.2f} s, std::stable_sort)"}
    return out


# Radiance-cache network work per sample (64 -> 64 x5 -> 3): 2 (5 * 64 * 64 + 64 * 3) flops forward;
# a training sample costs the forward, the activation gradients and the weight gradients (3x).
NRC_FLOPS_FORWARD = 2 * (5 * 64 * 64 + 64 * 3)
NRC_FLOPS_TRAIN = 3 * NRC_FLOPS_FORWARD
MFMA_BF16_PEAK_TFLOPS = 2500.0
# cached-integrator settings of the quality comparison (queries at bounce 2 as specified, and at 1)
CACHE_SETTINGS = dict(warmup_passes=2, train_steps=32, lr=3e-3, seed=1)


def relmse(x, r):
    import numpy as np
    return float(np.mean((x - r) ** 2 / (r * r + 0.01)))


def time_to_reach(curve, target):
    """Wall ms at which a (spp, ms, relmse) curve reaches `target`, interpolated in log-log between the
    bracketing points; None when the curve never gets there."""
    import math
    for (s0, t0, e0), (s1, t1, e1) in zip(curve, curve[1:]):
        if e0 >= target >= e1 and e0 > e1:
            f = (math.log(e0) - math.log(target)) / (math.log(e0) - math.log(e1))
            return math.exp(math.log(t0) + f * (math.log(t1) - math.log(t0)))
    if curve and curve[0][2] <= target:
        return curve[0][1] * curve[0][2] / target  # already below at the first point: scale back as 1/time
    return None


def relmse_at(curve, t):
    """relMSE of a curve at wall time t (log-log interpolation, 1/time extrapolation beyond the ends)."""
    import math
    pts = sorted(curve, key=lambda c: c[1])
    if t <= pts[0][1]:
        return pts[0][2] * pts[0][1] / t
    for (s0, t0, e0), (s1, t1, e1) in zip(pts, pts[1:]):
        if t0 <= t <= t1:
            f = (math.log(t) - math.log(t0)) / (math.log(t1) - math.log(t0))
            return math.exp(math.log(e0) + f * (math.log(e1) - math.log(e0)))
    return pts[-1][2] * pts[-1][1] / t


TRAFFIC_PROFILE = os.path.join("profiles", "traffic_wavefront_displaced_1080p.md")


def measured_traffic():
    """HBM bytes per frame of the traversal kernels, from the committed rocprofv3 counter summary
    (FETCH_SIZE and WRITE_SIZE, calibrated with the copy kernel); None when it is absent."""
    import re
    try:
        text = open(os.path.join(ROOT, TRAFFIC_PROFILE)).read()
    except OSError:
        return None
    m = re.search(r"^traversal_hbm_bytes_per_frame: ([0-9.eE+]+)$", text, re.M)
    return float(m.group(1)) if m else None


def progress(msg: str):
    """A step of the run finished (stderr; stdout carries only the JSON line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def quality(ctx, cfg, tuning, spps=(16, 32, 64, 128), ref_spp=4096) -> dict:
    """Equal-time relMSE of the cached and the plain wavefront tracer against a plain reference."""
    from lighttransportsimulation_amd import INTEGRATOR_CACHED, INTEGRATOR_WAVEFRONT
    base = dict(max_bounce=cfg["max_bounce"], rr_start=cfg["rr_start"], stratified=cfg["stratified"])
    tune = {k: tuning[k] for k in ("trace_block", "persistent_blocks", "sort_materials")} if tuning else {}
    t0 = time.perf_counter()
    ref, _, _ = ctx.render(ref_spp, seed=cfg["seed"] + 1000, integrator=INTEGRATOR_WAVEFRONT, **base, **tune)
    ref_s = time.perf_counter() - t0
    progress(f"quality: {ref_spp}-spp reference rendered in {ref_s:.1f} s")

    def timed(spp, integrator, **kw):
        t = time.perf_counter()
        img, _, st = ctx.render(spp, seed=cfg["seed"], integrator=integrator, **base, **tune, **kw)
        return img, (time.perf_counter() - t) * 1e3, st

    plain = []
    for spp in spps:
        img, ms, _ = timed(spp, INTEGRATOR_WAVEFRONT)
        plain.append((spp, round(ms, 2), relmse(img, ref)))
    progress("quality: plain curve done")
    out = {"relmse": "mean over pixels and channels of (x - r)^2 / (r^2 + 0.01)",
           "reference": f"plain wavefront, {ref_spp} spp, independent seed ({ref_s:.1f} s)",
           "time": "wall ms of the render call, fresh cache per point (training and warm-up included)",
           "plain": [list(p) for p in plain], "cached": {}}
    nominal = cfg["spp"]
    p_nom = next(p for p in plain if p[0] == nominal)
    nrc_perf = None
    for qb in (2, 1):
        curve = []
        for spp in spps:
            ctx.nrc_configure(query_bounce=qb, **CACHE_SETTINGS)
            img, ms, st = timed(spp, INTEGRATOR_CACHED, profile_phases=1)
            curve.append((spp, round(ms, 2), relmse(img, ref)))
            if qb == 2 and spp == nominal:
                nrc_perf = st
        progress(f"quality: cached curve, query at bounce {qb}, done")
        t_reach = time_to_reach(curve, p_nom[2])
        out["cached"][f"query_bounce_{qb}"] = {
            "curve": [list(c) for c in curve],
            "ms_to_plain_relmse_at_nominal_spp": None if t_reach is None else round(t_reach, 1),
            "equal_time_speedup": None if t_reach is None else round(p_nom[1] / t_reach, 3),
            "relmse_ratio_plain_over_cached_at_plain_nominal_time": round(relmse_at(plain, p_nom[1]) /
                                                                         relmse_at(curve, p_nom[1]), 3)}
    out["plain_relmse_at_nominal_spp"] = p_nom[2]
    out["cache_settings"] = dict(CACHE_SETTINGS, batch=65536, train_paths=262144, precision="bf16")
    if nrc_perf:
        # MFMA work of the cache in the 64-spp cached render: queries and training samples of the
        # last pass times the passes, over the phase times summed by HIP events
        passes = nrc_perf["launches"]
        q = nrc_perf["nrc_queries"] * passes
        n_total, batch, steps = nrc_perf["nrc_train_samples"], 65536, CACHE_SETTINGS["train_steps"]
        chunks = max(1, -(-n_total // batch))
        per_pass = sum(max(0, min(batch, n_total - (k % chunks) * batch)) for k in range(steps))
        inf_ms, tr_ms = nrc_perf["phase_ms"]["cache_infer"], nrc_perf["phase_ms"]["cache_train"]
        out["cache_mfma"] = {
            "infer_tflops": round(NRC_FLOPS_FORWARD * q / (inf_ms * 1e-3) / 1e12, 2) if inf_ms > 0 else None,
            "train_tflops": round(NRC_FLOPS_TRAIN * per_pass * passes / (tr_ms * 1e-3) / 1e12, 2) if tr_ms > 0 else None,
            "peak_tflops": MFMA_BF16_PEAK_TFLOPS, "infer_ms": round(inf_ms, 3), "train_ms": round(tr_ms, 3),
            "queries": q, "train_samples": per_pass * passes,
            "flops_per_query": NRC_FLOPS_FORWARD, "flops_per_train_sample": NRC_FLOPS_TRAIN}
    return out


def run_batch(args) -> dict:
    """Aggregate throughput of many scenes rendered together (lts_render_batch), against the same scenes
    rendered one after another."""
    from lighttransportsimulation_amd import INTEGRATOR_WAVEFRONT, Context, render_batch, scenegen
    w = WORKLOADS[args.workload]
    n = args.batch_scenes or w["n_scenes"]
    with tempfile.TemporaryDirectory() as d:
        t0 = time.perf_counter()
        cfgs = scenegen.gen_batch(d, n_scenes=n, spp=w["spp"], bounces=w["bounces"])
        gen_s = time.perf_counter() - t0
        progress(f"batch: {n} scenes generated in {gen_s:.1f} s")
        ctxs = [Context(0) for _ in cfgs]
        try:
            t0 = time.perf_counter()
            for c, cfg in zip(ctxs, cfgs):
                c.load_scene(d, cfg["scene"], cfg["subject"], cfg["width"], cfg["height"])
            load_s = time.perf_counter() - t0
            progress(f"batch: scenes loaded and built in {load_s:.1f} s")
            tris = [c.scene_info()["tris"] for c in ctxs]
            kw = dict(spp=w["spp"], max_bounce=w["bounces"], rr_start=w["rr"], seed=0x5EED)
            for _ in range(args.warmup):
                render_batch(ctxs, want_images=False, **kw)
            t0 = time.perf_counter()
            dev = 0.0
            for _ in range(args.steps):
                _, st = render_batch(ctxs, want_images=False, **kw)
                dev += st["render_ms"] + st["resolve_ms"]
            wall = time.perf_counter() - t0
            groups = st["batch_groups"]
            progress(f"batch: {args.steps} timed batches, {wall / args.steps * 1e3:.1f} ms each")
            # every scene in one group: one pass sequence for all 64
            render_batch(ctxs, want_images=False, group_pixels=-1, **kw)
            t1 = time.perf_counter()
            render_batch(ctxs, want_images=False, group_pixels=-1, **kw)
            one_group_ms = (time.perf_counter() - t1) * 1e3
            progress("batch: one-group render done")
            # the same scenes one at a time (each in its own context, warmed up once)
            single_ms = []
            for c in ctxs:
                c.render(integrator=INTEGRATOR_WAVEFRONT, want_image=False, **kw)
                t1 = time.perf_counter()
                c.render(integrator=INTEGRATOR_WAVEFRONT, want_image=False, **kw)
                single_ms.append((time.perf_counter() - t1) * 1e3)
        finally:
            for c in ctxs:
                c.close()
    samples = sum(cfg["width"] * cfg["height"] for cfg in cfgs) * w["spp"]
    ms = wall * 1e3 / args.steps
    return {"metric": "aggregate Msamples/s of a batched multi-scene render, 1 MI355X",
            "value": round(samples / (ms * 1e-3) / 1e6, 3), "unit": "Msamples/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": w["desc"], "scenes": n, "samples_per_step": samples,
                       "triangles_total": int(sum(tris)), "device_ms_per_step": round(dev / args.steps, 3),
                       "groups": groups, "one_group_ms": round(one_group_ms, 3),
                       "one_at_a_time_ms": round(sum(single_ms), 3),
                       "batch_speedup_over_one_at_a_time": round(sum(single_ms) / ms, 3),
                       "per_scene": [{"triangles": t, "res": cfg["width"], "ms_alone": round(m, 3)}
                                     for t, cfg, m in zip(tris, cfgs, single_ms)],
                       "generate_s": round(gen_s, 1), "load_and_build_s": round(load_s, 1)}}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="displaced_1080p", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of oracle work for cpu_baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--integrator", default="wavefront", choices=("megakernel", "wavefront"))
    ap.add_argument("--autotune", type=int, default=1, help="wavefront: tune block/grid/sort during warm-up")
    ap.add_argument("--spp-per-pass", type=int, default=0, help="samples per pixel per pass (0: the library's choice)")
    ap.add_argument("--no-quality", action="store_true", help="skip the cached-vs-plain equal-time comparison")
    ap.add_argument("--batch-scenes", type=int, default=0, help="batch64: number of scenes (default 64)")
    ap.add_argument("--cached-frame", action="store_true",
                    help="render one frame of the workload with the cached integrator (quality-block settings) "
                         "and print its stats; for kernel profiles of the cache")
    ap.add_argument("--copy-only", action="store_true",
                    help="run only the plain 1 GiB device copy (traffic-counter calibration)")
    args = ap.parse_args()
    if args.copy_only:
        from lighttransportsimulation_amd import Context
        with Context(0) as ctx:
            ms = ctx.copy_benchmark(1 << 30, 3)
        print(json.dumps({"copy_bytes": 1 << 30, "launches": 4, "ms_per_copy": round(ms, 4),
                          "GBps_read_plus_write": round(2 * (1 << 30) / (ms * 1e-3) / 1e9, 1)}), flush=True)
        return 0
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        return 2
    if WORKLOADS[args.workload]["gen"] == "soup":
        print(json.dumps(run_lbvh(args)), flush=True)
        return 0
    if WORKLOADS[args.workload]["gen"] == "batch":
        print(json.dumps(run_batch(args)), flush=True)
        return 0
    from lighttransportsimulation_amd import INTEGRATOR_CACHED, INTEGRATOR_MEGAKERNEL, INTEGRATOR_WAVEFRONT, Context
    if args.cached_frame:
        with tempfile.TemporaryDirectory() as d:
            cfg = make_workload(args.workload, d)
            with Context(0) as ctx:
                ctx.load_scene(d, cfg["scene"], cfg["subject"], cfg["width"], cfg["height"])
                ctx.nrc_configure(query_bounce=2, **CACHE_SETTINGS)
                t0 = time.perf_counter()
                _, _, st = ctx.render(cfg["spp"], cfg["max_bounce"], cfg["rr_start"], cfg["seed"], cfg["stratified"],
                                      integrator=INTEGRATOR_CACHED, want_image=False)
                ms = (time.perf_counter() - t0) * 1e3
        print(json.dumps({"cached_frame_ms": round(ms, 2), "render_ms": round(st["render_ms"], 2),
                          "passes": st["launches"], "queries_last_pass": st["nrc_queries"],
                          "train_samples_last_pass": st["nrc_train_samples"], "adam_steps": st["nrc_steps"],
                          "settings": dict(CACHE_SETTINGS, query_bounce=2)}), flush=True)
        return 0
    wavefront = args.integrator == "wavefront"
    with tempfile.TemporaryDirectory() as d:
        cfg = make_workload(args.workload, d)
        W, H, spp = cfg["width"], cfg["height"], cfg["spp"]
        with Context(0) as ctx:
            ctx.load_scene(d, cfg["scene"], cfg["subject"], W, H)
            info = ctx.scene_info()
            kw = dict(spp=spp, max_bounce=cfg["max_bounce"], rr_start=cfg["rr_start"], seed=cfg["seed"],
                      stratified=cfg["stratified"], want_image=False,
                      integrator=INTEGRATOR_WAVEFRONT if wavefront else INTEGRATOR_MEGAKERNEL,
                      autotune=int(wavefront and args.autotune), profile_phases=int(wavefront),
                      spp_per_pass=args.spp_per_pass)
            tuning = None
            progress(f"scene loaded: {info['tris']} triangles")
            for _ in range(max(args.warmup, 1 if wavefront and args.autotune else 0)):
                _, _, st = ctx.render(**kw)
                if wavefront:
                    tuning = {k: st[k] for k in ("trace_block", "persistent_blocks", "sort_materials", "use_graph")}
                    tuning["autotune_ms"] = round(st["autotune_ms"], 1)
            t0 = time.perf_counter()
            dev_ms, kern_ms, launches = 0.0, 0.0, 0
            phase = {}
            for _ in range(args.steps):
                _, _, st = ctx.render(**kw)
                dev_ms += st["render_ms"] + st["resolve_ms"]
                kern_ms += st["render_ms"]
                launches += st["launches"]
                for k, v in st["phase_ms"].items():
                    phase[k] = phase.get(k, 0.0) + v
            wall = time.perf_counter() - t0
            progress(f"{args.steps} timed frames: {wall / args.steps * 1e3:.1f} ms per frame")
            if wavefront and tuning is None:
                tuning = {k: st[k] for k in ("trace_block", "persistent_blocks", "sort_materials", "use_graph")}
            graph_rate = None
            if wavefront:
                # the same frames replayed as captured graphs, without the phase events
                gk = dict(kw, autotune=0, profile_phases=0, trace_block=tuning["trace_block"],
                          persistent_blocks=tuning["persistent_blocks"], sort_materials=tuning["sort_materials"])
                ctx.render(**gk)
                t1 = time.perf_counter()
                for _ in range(args.steps):
                    ctx.render(**gk)
                graph_rate = W * H * spp * args.steps / (time.perf_counter() - t1) / 1e6
            mega = None
            if wavefront:
                # the megakernel on the same frame, for comparison (not the headline value)
                mk = dict(kw, integrator=INTEGRATOR_MEGAKERNEL, autotune=0, profile_phases=0)
                ctx.render(**mk)
                t1 = time.perf_counter()
                for _ in range(2):
                    ctx.render(**mk)
                mega = W * H * spp * 2 / (time.perf_counter() - t1) / 1e6
            tree = None
            if wavefront:
                # render tree and its traversal statistics: node fetches and triangle tests of the
                # path rays of one sample per pixel, counted by the megakernel
                tree = ctx.render_tree_info()
                _, _, tst = ctx.render(**dict(kw, spp=1, stratified=0, integrator=INTEGRATOR_MEGAKERNEL, autotune=0,
                                              profile_phases=0, count_traversal=1))
                tree["path_ray_node_fetches_per_sample"] = round(tst["node_visits"] / tst["samples"], 3)
                tree["path_ray_tri_tests_per_sample"] = round(tst["tri_tests"] / tst["samples"], 3)
            qual = quality(ctx, cfg, tuning) if (wavefront and not args.no_quality
                                                  and args.workload == "displaced_1080p") else None
        samples = W * H * spp
        ms_per_step = wall * 1e3 / args.steps
        value = samples * args.steps / wall / 1e6
        out = {
            "metric": METRIC, "value": round(value, 3), "unit": "Msamples/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": WORKLOADS[args.workload]["desc"], "width": W, "height": H, "spp": spp,
                       "max_bounce": cfg["max_bounce"], "rr_start": cfg["rr_start"], "triangles": info["tris"],
                       "integrator": args.integrator, "device_ms_per_step": round(dev_ms / args.steps, 3)},
        }
        if tuning:
            out["config"]["tuning"] = tuning
        if mega:
            out["config"]["megakernel_msamples_per_s"] = round(mega, 3)
        if tree:
            out["config"]["render_tree"] = tree
        if graph_rate:
            out["config"]["graph_replay_msamples_per_s"] = round(graph_rate, 3)
            out["config"]["timed_region"] = "passes launched directly, HIP events between the phases"
        if phase:
            out["config"]["phase_ms_per_step"] = {k: round(v / args.steps, 3) for k, v in phase.items()}
        if not args.no_cpu:
            progress("cpu baseline starts")
            cb = cpu_baseline(d, cfg, args.cpu_budget)
            progress("cpu baseline done")
            bytes_per_sample = 32.0 * cb["node_visits_per_sample"] + 36.0 * cb["tri_tests_per_sample"]
            if wavefront:
                # the traversal kernel (closest hit of path rays, any hit of shadow rays, closest hit
                # of MIS rays, traced together) against all node and triangle tests of the frame
                trav_ms = (phase["extend"] + phase["shadow"] + phase["mis"]) / args.steps
                achieved = bytes_per_sample * samples / (trav_ms * 1e-3) / 1e9
                kernel = "k_wf_trace (per frame)"
                kernel_ms = trav_ms
            else:
                launch_s = kern_ms / launches / 1e3
                achieved = bytes_per_sample * (samples * args.steps / launches) / launch_s / 1e9
                kernel = "k_megakernel"
                kernel_ms = kern_ms / launches
            traffic = measured_traffic() if (wavefront and args.workload == "displaced_1080p") else None
            out["roofline"] = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                               "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic, "kernel": kernel,
                               "kernel_ms": round(kernel_ms, 3),
                               "algorithmic_bytes_per_sample": round(bytes_per_sample, 1)}
            if traffic is not None:
                out["roofline"]["traffic_source"] = TRAFFIC_PROFILE + " (HBM bytes per frame of the same kernels)"
            out["cpu_baseline"] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in cb.items()
                                   if k in ("value", "unit", "cores", "kind", "sample")}
        if qual:
            out["quality"] = qual
        print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
