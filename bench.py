"""Throughput of the path tracer on one MI355X, as one JSON line.

A step is one full frame of the workload (all samples of every pixel) through the wavefront
integrator (passes launched directly with HIP events between the phases, autotuned in the warm-up)
and the film gather, with the scene and BVH already resident in HBM. Workloads are generated from
fixed seeds in the reference's own file formats (lighttransportsimulation_amd/scenegen.py).

Extra fields of the default run (the 1080p workload):
  roofline        the traversal kernel (k_wf_trace: path, shadow and MIS rays, one launch per bounce)
                  against HBM bandwidth. Algorithmic bytes are the kernel's own work on its own
                  tree: 64 B per quantized 4-wide node fetch and 48 B per triangle test, counted on
                  the GPU for every ray kind of the same frame by the counting instantiation of the
                  kernel (an untimed run), over the kernel's device time per frame from HIP events.
                  `traffic` is the HBM bytes per frame that rocprofv3's FETCH_SIZE / WRITE_SIZE
                  counters read for the same kernel (profiles/), and `issue` its VALU issue rate.
  cpu_baseline    the oracle (the reference integrator restated in C++) on all host cores over a
                  bounded band of rows of the same frame.
  quality         equal-time relMSE of the cached integrator (learned radiance cache) against the
                  plain wavefront integrator, both against a 4096-spp plain reference: with the
                  cache carried from frame to frame (steady state) and from a fresh cache.
  lbvh            the LBVH build alone on the 1M- and 8M-triangle soups (Mprims/s, HBM fraction).
  batch64         64 scenes of uneven size rendered together (aggregate Msamples/s).
  reference_scene the reference's own bigBox scene with its diamond subject at 1920x1080 with the
                  settings of its app.params, beside the render time the reference quotes.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Msamples/s & equal-time relMSE: cached vs plain path tracer, 1 MI355X"
HBM_PEAK_GBS = 8000.0
# bytes the traversal kernel reads per render-node fetch (a quantized 4-wide node, four 16-byte
# loads) and per triangle test (three 16-byte loads)
NODE_BYTES, TRI_BYTES = 64, 48
# radix passes of the LBVH sort (10-bit digits over the 30-bit Morton codes)
LBVH_RADIX_PASSES = 3

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


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


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
                       f"({st['samples']} samples, {st['seconds']:.1f} s, {st['threads']} threads of {cpu_model()})",
                cpu_model=cpu_model())


# Algorithmic HBM bytes of one LBVH build per primitive: read the soup (36), write code + id (8),
# the radix passes over (code, id) pairs (16 each: read and write 8 B), write the node and leaf
# records (60).
def lbvh_bytes_per_prim(passes: int) -> int:
    return 36 + 8 + 16 * passes + 60


def lbvh_measure(ctx, n: int, warmup: int, steps: int) -> dict:
    from lighttransportsimulation_amd import scenegen
    tri = scenegen.triangle_soup(n, seed=3)
    ctx.bvh_build(tri, mode=0, export=False)  # uploads the soup once; it stays resident
    if warmup:
        ctx.bvh_build_device(None, n, 0, repeat=warmup)
    t0 = time.perf_counter()
    t = ctx.bvh_build_device(None, n, 0, repeat=steps)
    wall = time.perf_counter() - t0
    bpp = lbvh_bytes_per_prim(LBVH_RADIX_PASSES)
    achieved = bpp * n / (t["total_ms"] * 1e-3) / 1e9
    return dict(tri=tri, value=n / (t["total_ms"] * 1e3), ms=t["total_ms"], wall_ms=wall * 1e3 / steps, stages=t,
                achieved=achieved, bytes_per_prim=bpp)


def run_lbvh(args) -> dict:
    from lighttransportsimulation_amd import Context
    w = WORKLOADS[args.workload]
    n = w["n"]
    with Context(0) as ctx:
        m = lbvh_measure(ctx, n, args.warmup, args.steps)
        copy_bytes = 1 << 30
        copy_ms = ctx.copy_benchmark(copy_bytes, 10)
    out = {"metric": "LBVH build Mprims/s, 1 MI355X", "value": round(m["value"], 2), "unit": "Mprims/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(m["ms"], 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": w["desc"], "primitives": n, "wall_ms_per_step": round(m["wall_ms"], 3),
                      "stage_ms": {k: round(v, 4) for k, v in m["stages"].items()}},
           "roofline": {"bound": "hbm", "achieved": round(m["achieved"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(m["achieved"] / HBM_PEAK_GBS, 4), "traffic": None, "kernel": "whole build",
                        "algorithmic_bytes_per_prim": m["bytes_per_prim"],
                        "copy_kernel_GBps": round(2 * copy_bytes / (copy_ms * 1e-3) / 1e9, 1)}}
    if not args.no_cpu:
        out["cpu_baseline"] = lbvh_cpu_baseline(m["tri"])
    return out


def lbvh_cpu_baseline(tri) -> dict:
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    m = min(tri.shape[0], 1 << 20)
    t1 = time.perf_counter()
    oracle_lib.lbvh_build(tri[:m], 0)
    sec = time.perf_counter() - t1
    return {"value": round(m / sec / 1e6, 3), "unit": "Mprims/s", "cores": 1, "kind": "port",
            "sample": f"first {m} triangles of the same soup ({sec:.2f} s, std::stable_sort, {cpu_model()})"}


def lbvh_block(no_cpu: bool) -> dict:
    """Both LBVH workloads, the build alone repeated on a resident soup, each measured by a child
    process of its own (this script with --workload lbvh_1m / lbvh_8m, 3 warm-up and 20 timed
    builds), i.e. in a fresh HIP context as the standalone LBVH lines are."""
    out = {"definition": "n / device time of one build (scene bounds, Morton codes, radix sort, hierarchy and node "
                         "bounds) on a resident soup, replayed as a captured HIP graph; achieved = algorithmic "
                         "bytes per primitive x n / build time",
           "radix_passes": LBVH_RADIX_PASSES,
           "process": "one child process per workload: bench.py --workload lbvh_1m / lbvh_8m --steps 20 --warmup 3"}
    for name, workload in (("1M", "lbvh_1m"), ("8M", "lbvh_8m")):
        cmd = [sys.executable, os.path.abspath(__file__), "--workload", workload, "--steps", "20", "--warmup", "3"]
        if no_cpu or name == "8M":
            cmd.append("--no-cpu")
        p = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
        if p.returncode != 0:
            raise RuntimeError(f"{' '.join(cmd)} failed ({p.returncode}): {p.stderr.strip()[-2000:]}")
        line = json.loads(p.stdout.strip().splitlines()[-1])
        rf = line["roofline"]
        out[name] = {"primitives": line["config"]["primitives"], "mprims_per_s": round(line["value"], 1),
                     "ms_per_build": line["ms_per_step"], "stage_ms": line["config"]["stage_ms"],
                     "achieved_GBps": rf["achieved"], "hbm_frac": rf["frac"],
                     "algorithmic_bytes_per_prim": rf["algorithmic_bytes_per_prim"]}
        if "cpu_baseline" in line:
            out[name]["cpu_baseline"] = line["cpu_baseline"]
        progress(f"lbvh {name}: {line['value']:.0f} Mprims/s")
    return out


# Radiance-cache network work per sample (64 -> 64 x5 -> 3): 2 (5 * 64 * 64 + 64 * 3) flops forward;
# a training sample costs the forward, the activation gradients and the weight gradients (3x).
NRC_FLOPS_FORWARD = 2 * (5 * 64 * 64 + 64 * 3)
NRC_FLOPS_TRAIN = 3 * NRC_FLOPS_FORWARD
MFMA_BF16_PEAK_TFLOPS = 2500.0
# the learned cache's settings on the 1080p workload: learning rate 1e-2, Adam steps on batches of
# 2^16 path-suffix samples (16 per frame, from up to 2^18 training paths), queries after bounce 2;
# the cache is carried from frame to frame, the first 4 frames after configuration warm it up
CACHE_SETTINGS = dict(lr=1e-2, batch=65536, train_steps=16, train_paths=262144, seed=1)
CACHE_WARM_FRAMES = 12


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
    if curve and curve[-1][2] > target:
        return curve[-1][1] * curve[-1][2] / target  # beyond the last point: error falls as 1/time
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
ISSUE_PROFILE = os.path.join("profiles", "stalls_traversal_displaced_1080p.md")


def profile_figure(path: str, key: str):
    """A machine-readable figure `key: value` of a committed rocprofv3 summary; None when absent."""
    import re
    try:
        text = open(os.path.join(ROOT, path)).read()
    except OSError:
        return None
    m = re.search(r"^" + re.escape(key) + r": ([0-9.eE+-]+)$", text, re.M)
    return float(m.group(1)) if m else None


def progress(msg: str):
    """A step of the run finished (stderr; stdout carries only the JSON line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def equal_time(plain, curve, nominal):
    """Equal-time figures of a cached curve against the plain curve at the plain tracer's nominal spp."""
    p_nom = next(p for p in plain if p[0] == nominal)
    t_reach = time_to_reach(curve, p_nom[2])
    return {"curve": [list(c) for c in curve],
            "ms_to_plain_relmse_at_nominal_spp": None if t_reach is None else round(t_reach, 2),
            "equal_time_speedup": None if t_reach is None else round(p_nom[1] / t_reach, 3),
            "relmse_ratio_plain_over_cached_at_plain_nominal_time": round(relmse_at(plain, p_nom[1]) /
                                                                         relmse_at(curve, p_nom[1]), 3)}


def quality(ctx, cfg, tuning, spps=(16, 32, 64, 128), ref_spp=4096) -> dict:
    """Equal-time relMSE of the cached and the plain wavefront tracer against a plain reference."""
    from lighttransportsimulation_amd import INTEGRATOR_CACHED, INTEGRATOR_WAVEFRONT
    base = dict(max_bounce=cfg["max_bounce"], rr_start=cfg["rr_start"], stratified=cfg["stratified"])
    tune = {k: tuning[k] for k in ("trace_block", "persistent_blocks", "sort_materials", "tile_rows")} if tuning else {}
    t0 = time.perf_counter()
    ref, _, _ = ctx.render(ref_spp, seed=cfg["seed"] + 1000, integrator=INTEGRATOR_WAVEFRONT, **base, **tune)
    ref_s = time.perf_counter() - t0
    progress(f"quality: {ref_spp}-spp reference rendered in {ref_s:.1f} s")

    def timed(spp, integrator, warm=True, **kw):
        if warm:  # the same call once untimed (graph capture), with another seed
            ctx.render(spp, seed=cfg["seed"] + 1, integrator=integrator, want_image=False, **base, **tune, **kw)
        t = time.perf_counter()
        img, _, st = ctx.render(spp, seed=cfg["seed"], integrator=integrator, **base, **tune, **kw)
        return img, (time.perf_counter() - t) * 1e3, st

    plain = []
    for spp in spps:
        img, ms, _ = timed(spp, INTEGRATOR_WAVEFRONT)
        plain.append((spp, round(ms, 2), relmse(img, ref)))
    progress("quality: plain curve done")
    nominal = cfg["spp"]
    out = {"relmse": "mean over pixels and channels of (x - r)^2 / (r^2 + 0.01)",
           "reference": f"plain wavefront, {ref_spp} spp, independent seed ({ref_s:.1f} s)",
           "time": "wall ms of the render call (one pass per frame), training included",
           "plain": [list(p) for p in plain], "steady_state": {}, "fresh_cache": {}}
    nrc_perf = None
    for qb in (2, 1):
        # steady state: the cache carried from frame to frame, as a renderer runs it
        ctx.nrc_configure(query_bounce=qb, **CACHE_SETTINGS)
        for f in range(CACHE_WARM_FRAMES):
            ctx.render(nominal, seed=100 + f, integrator=INTEGRATOR_CACHED, want_image=False, **base, **tune)
        curve = []
        for spp in spps:
            img, ms, st = timed(spp, INTEGRATOR_CACHED)
            curve.append((spp, round(ms, 2), relmse(img, ref)))
        out["steady_state"][f"query_bounce_{qb}"] = equal_time(plain, curve, nominal)
        if qb == 2:
            _, _, nrc_perf = ctx.render(nominal, seed=cfg["seed"] + 2, integrator=INTEGRATOR_CACHED, want_image=False,
                                        profile_phases=1, **base, **tune)
        progress(f"quality: steady-state cached curve, query at bounce {qb}, done")
    # fresh cache: each frame starts from a newly initialised cache, in passes of 16 spp whose
    # first two warm the cache up
    curve = []
    for spp in spps:
        ctx.nrc_configure(query_bounce=2, warmup_passes=2, **CACHE_SETTINGS)
        ctx.render(16, seed=cfg["seed"] + 3, integrator=INTEGRATOR_CACHED, spp_per_pass=16, want_image=False,
                   **base, **tune)  # graph capture of the 16-spp pass
        ctx.nrc_configure(query_bounce=2, warmup_passes=2, **CACHE_SETTINGS)
        img, ms, st = timed(spp, INTEGRATOR_CACHED, warm=False, spp_per_pass=16)
        curve.append((spp, round(ms, 2), relmse(img, ref)))
    out["fresh_cache"]["query_bounce_2"] = equal_time(plain, curve, nominal)
    progress("quality: fresh-cache curve done")
    out["headline"] = {"equal_time_speedup_steady_state_query_bounce_2":
                       out["steady_state"]["query_bounce_2"]["equal_time_speedup"],
                       "equal_time_speedup_steady_state_query_bounce_1":
                       out["steady_state"]["query_bounce_1"]["equal_time_speedup"],
                       "equal_time_speedup_fresh_cache_query_bounce_2":
                       out["fresh_cache"]["query_bounce_2"]["equal_time_speedup"]}
    out["plain_relmse_at_nominal_spp"] = next(p for p in plain if p[0] == nominal)[2]
    out["cache_settings"] = dict(CACHE_SETTINGS, warm_frames=CACHE_WARM_FRAMES, warmup_passes=4, precision="bf16",
                                 beta1=0.9, beta2=0.99, eps=1e-15)
    if nrc_perf:
        # MFMA work of the cache in one steady-state frame (one pass): its queries and the samples of
        # its Adam steps, over the inference and training phase times from HIP events
        q = nrc_perf["nrc_queries"]
        n_total, batch, steps = nrc_perf["nrc_train_samples"], CACHE_SETTINGS["batch"], CACHE_SETTINGS["train_steps"]
        chunks = max(1, -(-n_total // batch))
        trained = sum(max(0, min(batch, n_total - (k % chunks) * batch)) for k in range(steps))
        inf_ms, tr_ms = nrc_perf["phase_ms"]["cache_infer"], nrc_perf["phase_ms"]["cache_train"]
        out["cache_mfma"] = {
            "infer_tflops": round(NRC_FLOPS_FORWARD * q / (inf_ms * 1e-3) / 1e12, 2) if inf_ms > 0 else None,
            "train_tflops": round(NRC_FLOPS_TRAIN * trained / (tr_ms * 1e-3) / 1e12, 2) if tr_ms > 0 else None,
            "peak_tflops": MFMA_BF16_PEAK_TFLOPS, "infer_ms": round(inf_ms, 3), "train_ms": round(tr_ms, 3),
            "queries": q, "train_samples": trained, "frame_phase_ms": {k: round(v, 3) for k, v in
                                                                        nrc_perf["phase_ms"].items()},
            "flops_per_query": NRC_FLOPS_FORWARD, "flops_per_train_sample": NRC_FLOPS_TRAIN}
    return out


def batch_measure(n_scenes: int, steps: int, warmup: int, compare: bool = True) -> dict:
    """Aggregate throughput of many scenes rendered together (lts_render_batch), against the same scenes
    rendered one after another."""
    from lighttransportsimulation_amd import INTEGRATOR_WAVEFRONT, Context, render_batch, scenegen
    w = WORKLOADS["batch64"]
    with tempfile.TemporaryDirectory() as d:
        t0 = time.perf_counter()
        cfgs = scenegen.gen_batch(d, n_scenes=n_scenes, spp=w["spp"], bounces=w["bounces"])
        gen_s = time.perf_counter() - t0
        progress(f"batch: {n_scenes} scenes generated in {gen_s:.1f} s")
        ctxs = [Context(0) for _ in cfgs]
        try:
            t0 = time.perf_counter()
            for c, cfg in zip(ctxs, cfgs):
                c.load_scene(d, cfg["scene"], cfg["subject"], cfg["width"], cfg["height"])
            load_s = time.perf_counter() - t0
            progress(f"batch: scenes loaded and built in {load_s:.1f} s")
            tris = [c.scene_info()["tris"] for c in ctxs]
            kw = dict(spp=w["spp"], max_bounce=w["bounces"], rr_start=w["rr"], seed=0x5EED)
            for _ in range(max(1, warmup)):
                render_batch(ctxs, want_images=False, **kw)
            t0 = time.perf_counter()
            dev = 0.0
            for _ in range(steps):
                _, st = render_batch(ctxs, want_images=False, **kw)
                dev += st["render_ms"] + st["resolve_ms"]
            wall = time.perf_counter() - t0
            groups = st["batch_groups"]
            progress(f"batch: {steps} timed batches, {wall / steps * 1e3:.1f} ms each")
            one_group_ms, single_ms = None, []
            if compare:
                # every scene in one group: one pass sequence for all of them
                render_batch(ctxs, want_images=False, group_pixels=-1, **kw)
                t1 = time.perf_counter()
                render_batch(ctxs, want_images=False, group_pixels=-1, **kw)
                one_group_ms = (time.perf_counter() - t1) * 1e3
                # the same scenes one at a time (each in its own context, warmed up once)
                for c in ctxs:
                    c.render(integrator=INTEGRATOR_WAVEFRONT, want_image=False, **kw)
                    t1 = time.perf_counter()
                    c.render(integrator=INTEGRATOR_WAVEFRONT, want_image=False, **kw)
                    single_ms.append((time.perf_counter() - t1) * 1e3)
                progress("batch: one-group and one-at-a-time renders done")
        finally:
            for c in ctxs:
                c.close()
    samples = sum(cfg["width"] * cfg["height"] for cfg in cfgs) * w["spp"]
    ms = wall * 1e3 / steps
    out = {"workload": w["desc"], "scenes": n_scenes, "samples_per_batch": samples, "steps": steps,
           "msamples_per_s": round(samples / (ms * 1e-3) / 1e6, 3), "ms_per_batch": round(ms, 3),
           "triangles_total": int(sum(tris)), "device_ms_per_batch": round(dev / steps, 3), "groups": groups,
           "generate_s": round(gen_s, 1), "load_and_build_s": round(load_s, 1)}
    if compare:
        out.update(one_group_ms=round(one_group_ms, 3), one_at_a_time_ms=round(sum(single_ms), 3),
                   batch_speedup_over_one_at_a_time=round(sum(single_ms) / ms, 3),
                   per_scene=[{"triangles": t, "res": cfg["width"], "ms_alone": round(m, 3)}
                              for t, cfg, m in zip(tris, cfgs, single_ms)])
    return out


def run_batch(args) -> dict:
    b = batch_measure(args.batch_scenes or WORKLOADS["batch64"]["n_scenes"], args.steps, args.warmup)
    return {"metric": "aggregate Msamples/s of a batched multi-scene render, 1 MI355X",
            "value": b["msamples_per_s"], "unit": "Msamples/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": b["ms_per_batch"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic", "config": b}


def reference_scene_block() -> dict:
    """The reference's bigBox scene (res/platforms) with its diamond subject at 1920x1080 and the
    sample, bounce and roulette settings of its app.params (10, 5, 2). The reference quotes 14 s for
    the diamond at 1080p on a GTX 1650 Ti (app.params:20-22), in a closeupPlane scene that its tree
    does not contain, with the spp and bounces of that run unrecorded: context, not a comparison."""
    from lighttransportsimulation_amd import INTEGRATOR_MEGAKERNEL, INTEGRATOR_WAVEFRONT, Context
    proj = os.path.join(ROOT, "tests", "golden") + "/"
    W, H, spp, bounces, rr = 1920, 1080, 10, 5, 2
    out = {"scene": "res/platforms/bigBox + res/diamond (the reference's own files)", "width": W, "height": H,
           "spp": spp, "max_bounce": bounces, "rr_start": rr,
           "reference_quote": "diamond, 1080p: 14 s on a GTX 1650 Ti (app.params:20-22; closeupPlane scene, spp and "
                              "bounces not recorded)"}
    with Context(0) as ctx:
        ctx.load_scene(proj, "res/platforms/bigBox", "res/diamond", W, H)
        out["triangles"] = ctx.scene_info()["tris"]
        for integ, name in ((INTEGRATOR_WAVEFRONT, "wavefront"), (INTEGRATOR_MEGAKERNEL, "megakernel")):
            ctx.render(spp, bounces, rr, 0x5EED, integrator=integ, want_image=False)
            t0 = time.perf_counter()
            ctx.render(spp, bounces, rr, 0x5EED, integrator=integ, want_image=False)
            ms = (time.perf_counter() - t0) * 1e3
            out[name] = {"ms": round(ms, 3), "msamples_per_s": round(W * H * spp / (ms * 1e-3) / 1e6, 1)}
    return out


TUNING_KEYS = ("trace_block", "persistent_blocks", "sort_materials", "use_graph", "tile_rows", "spp_per_pass")


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
    ap.add_argument("--no-extras", action="store_true", help="skip the lbvh, batch64 and reference_scene blocks")
    ap.add_argument("--bdpt", action="store_true",
                    help="also time the bidirectional integrator on the workload's frame (4 spp)")
    ap.add_argument("--batch-scenes", type=int, default=0, help="batch64: number of scenes (default 64)")
    ap.add_argument("--cached-frame", action="store_true",
                    help="render steady-state frames of the workload with the cached integrator and print their "
                         "stats; for kernel profiles of the cache")
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
    from lighttransportsimulation_amd import (INTEGRATOR_BDPT, INTEGRATOR_CACHED, INTEGRATOR_MEGAKERNEL,
                                          INTEGRATOR_WAVEFRONT, Context)
    if args.cached_frame:
        with tempfile.TemporaryDirectory() as d:
            cfg = make_workload(args.workload, d)
            with Context(0) as ctx:
                ctx.load_scene(d, cfg["scene"], cfg["subject"], cfg["width"], cfg["height"])
                ctx.nrc_configure(query_bounce=2, warmup_passes=-1, **CACHE_SETTINGS)
                base = dict(max_bounce=cfg["max_bounce"], rr_start=cfg["rr_start"], stratified=cfg["stratified"],
                            integrator=INTEGRATOR_CACHED, want_image=False)
                for f in range(4):
                    ctx.render(cfg["spp"], seed=100 + f, **base)
                t0 = time.perf_counter()
                _, _, st = ctx.render(cfg["spp"], seed=cfg["seed"], **base)
                ms = (time.perf_counter() - t0) * 1e3
        print(json.dumps({"cached_frame_ms": round(ms, 2), "render_ms": round(st["render_ms"], 2),
                          "passes": st["launches"], "queries": st["nrc_queries"],
                          "train_samples": st["nrc_train_samples"], "adam_steps": st["nrc_steps"],
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
                    tuning = {k: st[k] for k in TUNING_KEYS}
                    tuning["autotune_ms"] = round(st["autotune_ms"], 1)
            t0 = time.perf_counter()
            dev_ms, kern_ms, launches = 0.0, 0.0, 0
            phase, phase_launches = {}, {}
            for _ in range(args.steps):
                _, _, st = ctx.render(**kw)
                dev_ms += st["render_ms"] + st["resolve_ms"]
                kern_ms += st["render_ms"]
                launches += st["launches"]
                for k, v in st["phase_ms"].items():
                    phase[k] = phase.get(k, 0.0) + v
                for k, v in st["phase_launches"].items():
                    phase_launches[k] = phase_launches.get(k, 0) + v
            wall = time.perf_counter() - t0
            progress(f"{args.steps} timed frames: {wall / args.steps * 1e3:.1f} ms per frame")
            if wavefront and tuning is None:
                tuning = {k: st[k] for k in TUNING_KEYS}
            graph_rate = mega = tree = counts = bdpt = None
            if wavefront:
                # the same frames replayed as captured graphs, without the phase events
                gk = dict(kw, autotune=0, profile_phases=0, trace_block=tuning["trace_block"],
                          persistent_blocks=tuning["persistent_blocks"], sort_materials=tuning["sort_materials"],
                          tile_rows=tuning["tile_rows"], spp_per_pass=tuning["spp_per_pass"])
                ctx.render(**gk)
                t1 = time.perf_counter()
                for _ in range(args.steps):
                    ctx.render(**gk)
                graph_rate = W * H * spp * args.steps / (time.perf_counter() - t1) / 1e6
                # the megakernel on the same frame, for comparison (not the headline value)
                mk = dict(kw, integrator=INTEGRATOR_MEGAKERNEL, autotune=0, profile_phases=0)
                ctx.render(**mk)
                t1 = time.perf_counter()
                for _ in range(2):
                    ctx.render(**mk)
                mega = W * H * spp * 2 / (time.perf_counter() - t1) / 1e6
                if args.bdpt:
                    # bidirectional path tracing on the same frame at 4 spp (a different estimator:
                    # every connection strategy per sample, so far more rays per sample; for scale)
                    bd_spp = 4
                    bk = dict(kw, integrator=INTEGRATOR_BDPT, autotune=0, profile_phases=0, spp=bd_spp,
                              spp_per_pass=0)
                    ctx.render(**bk)
                    t1 = time.perf_counter()
                    _, _, bst = ctx.render(**bk)
                    bdpt = {"spp": bd_spp, "max_bounce": cfg["max_bounce"],
                            "msamples_per_s": round(W * H * bd_spp / (time.perf_counter() - t1) / 1e6, 3),
                            "kernel_ms_per_frame": round(bst["render_ms"], 3)}
                # the traversal kernel's own work on the same frame: render-node fetches and triangle
                # tests of every ray kind, counted by its counting instantiation (an untimed run)
                tree = ctx.render_tree_info()
                _, _, cst = ctx.render(**dict(gk, count_traversal=1))
                counts = {k: (cst["ray_nodes"][k] / cst["samples"], cst["ray_tris"][k] / cst["samples"])
                          for k in cst["ray_nodes"]}
                tree["path_ray_node_fetches_per_sample"] = round(counts["path"][0], 3)
                tree["path_ray_tri_tests_per_sample"] = round(counts["path"][1], 3)
                progress("graph replay, megakernel and counting runs done")
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
        if bdpt:
            out["config"]["bdpt"] = bdpt
        if tree:
            out["config"]["render_tree"] = tree
        if graph_rate:
            out["config"]["graph_replay_msamples_per_s"] = round(graph_rate, 3)
            out["config"]["timed_region"] = "passes launched directly, HIP events between the phases"
        if phase:
            out["config"]["phase_ms_per_step"] = {k: round(v / args.steps, 3) for k, v in phase.items()}
        if counts:
            # one launch per bounce traces that bounce's path rays with the previous bounce's shadow
            # and MIS rays; achieved = all of the frame's node and triangle bytes over the launches'
            # summed device time (the phase events bracket exactly those launches)
            per_sample = {k: NODE_BYTES * n + TRI_BYTES * t for k, (n, t) in counts.items()}
            bytes_frame = sum(per_sample.values()) * samples
            trav_ms = phase["extend"] / args.steps
            n_launch = phase_launches.get("extend", 0) / args.steps
            achieved = bytes_frame / (trav_ms * 1e-3) / 1e9
            traffic = profile_figure(TRAFFIC_PROFILE, "traversal_hbm_bytes_per_frame")
            valu = profile_figure(ISSUE_PROFILE, "traversal_valu_issue_rate")
            out["roofline"] = {
                "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": "k_wf_trace",
                "kernel_ms_per_step": round(trav_ms, 3), "launches_per_step": n_launch,
                "kernel_avg_launch_ms": round(trav_ms / n_launch, 3) if n_launch else None,
                "algorithmic_bytes": "64 B per 4-wide node fetch + 48 B per triangle test, counted on the GPU for "
                                     "the same frame",
                "algorithmic_bytes_per_sample": {k: round(v, 1) for k, v in per_sample.items()},
                "node_fetches_per_sample": {k: round(n, 3) for k, (n, t) in counts.items()},
                "tri_tests_per_sample": {k: round(t, 3) for k, (n, t) in counts.items()},
                "algorithmic_bytes_per_step": round(bytes_frame),
                "hbm_traffic_frac": (round(traffic / (trav_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
                                     if traffic and args.workload == "displaced_1080p" else None),
                "valu_issue_rate": valu,
                "sources": f"traffic: {TRAFFIC_PROFILE} (FETCH_SIZE + WRITE_SIZE per frame, copy-kernel calibrated); "
                           f"valu_issue_rate: {ISSUE_PROFILE} (wave64 VALU instructions per SIMD per 4 cycles: SQ_INSTS_VALU over 1024 SIMDs x 2.4 GHz / 4)"}
        if not args.no_cpu:
            progress("cpu baseline starts")
            cb = cpu_baseline(d, cfg, args.cpu_budget)
            progress("cpu baseline done")
            out["cpu_baseline"] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in cb.items()}
        if qual:
            out["quality"] = qual
    if wavefront and args.workload == "displaced_1080p" and not args.no_extras:
        out["lbvh"] = lbvh_block(args.no_cpu)
        out["batch64"] = batch_measure(WORKLOADS["batch64"]["n_scenes"], 2, 1)
        out["reference_scene"] = reference_scene_block()
        progress("lbvh, batch64 and reference scene done")
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
