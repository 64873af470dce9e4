"""Bidirectional path tracing (LTS_INTEGRATOR_BDPT; the reference's BidirectionalPathIntegrator,
intergrators/BidirectionalPathIntegrator.cuh, is disabled upstream at
LightTransportSimulation.cu:116-121, so it has no output to compare with).

Pinned by convergence: on scenes where the plain path tracer (itself per-sample identical to the
oracle) is unbiased, both integrators estimate the same per-pixel expectation from the same film
and lens samples (the test pairs them sample by sample). The path tracer is unbiased here because each scene has one area light, a single
triangle (the reference's light search never picks the last of several lights, and its infinite
light samples a different density than it reports, Light.cu:68-90), and no infinite light. Block
means of the per-sample differences must be within 5 standard errors of zero, and the image means
within 1 % and 5 standard errors of each other. Scenes: a matte box with a matte sphere, and the same box with a mirror sphere at 8
bounces. No glass: the path tracer counts paths with two consecutive specular vertices twice (its
light sample takes the previous bounce's specular flag, PathTracingIntegrator.cuh:77), and a path
that enters and leaves a sphere has two. The path tracer also stops one vertex short of the
bidirectional integrator on paths whose last scattering vertex is specular (it adds emission after a
specular bounce only on the next loop iteration); at 8 bounces that term is far below the noise.
"""
import os

import numpy as np
import pytest

from lighttransportsimulation_amd import INTEGRATOR_BDPT, INTEGRATOR_MEGAKERNEL, LtsError
from lighttransportsimulation_amd.scenegen import SceneText, icosphere, mat_matte, quad, write_bin

W, H = 32, 24
SEED = 77


def _box_scene(dir_: str, sphere: str) -> tuple[str, str]:
    name = sphere
    d = os.path.join(dir_, name)
    os.makedirs(d, exist_ok=True)
    st = SceneText(camera=(0, 0, -5.4, 0, 0, 0, 0, 1, 0, 60, 5.4))
    walls = {"floor": ((0, -2, 0), (2, 0, 0), (0, 0, -2)), "ceiling": ((0, 2, 0), (2, 0, 0), (0, 0, 2)),
             "back": ((0, 0, 2), (2, 0, 0), (0, 2, 0)), "left": ((-2, 0, 0), (0, 0, 2), (0, 2, 0)),
             "right": ((2, 0, 0), (0, 2, 0), (0, 0, 2))}
    albedo = {"floor": (0.7, 0.7, 0.7), "ceiling": (0.7, 0.7, 0.7), "back": (0.6, 0.6, 0.5),
              "left": (0.7, 0.2, 0.2), "right": (0.2, 0.7, 0.2)}
    for wall, (c, u, v) in walls.items():
        P, F, N = quad(c, u, v)
        write_bin(os.path.join(d, wall), P.astype(np.float32), F, N=N.astype(np.float32), idx_n=F)
        st.meshes.append(f"{name}/{wall}")
        st.mats.append(mat_matte(albedo[wall]))
    # one emissive triangle under the ceiling, facing down
    P, F, N = quad((0.3, 1.98, 0.2), (0.6, 0, 0), (0, 0, 0.6))
    write_bin(os.path.join(d, "light"), P[:3].astype(np.float32), F[:1], N=N[:3].astype(np.float32), idx_n=F[:1])
    st.meshes.append(f"{name}/light")
    st.mats.append(mat_matte((0.5, 0.5, 0.5)))
    st.area_les = [(20.0, 18.0, 15.0)]
    V, Fs = icosphere(2)
    Ps = V * 0.8 + np.array([-0.4, -1.2, 0.3])
    write_bin(os.path.join(d, "sphere"), Ps.astype(np.float32), Fs, N=V.astype(np.float32), idx_n=Fs)
    st.subj_meshes = [f"{name}/sphere"]
    st.subj_mats = [{"matte": mat_matte((0.3, 0.5, 0.8)), "mirror": "R c 0.9 0.85 0.8 1",
                     "glass": "G c 1 1 1 c 1 1 1 1.5"}[sphere]]
    st.write(dir_, f"{name}/scene")
    return f"{name}/scene", f"{name}/scene"


@pytest.fixture(scope="module")
def scenes(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("bdpt")) + "/"
    return d, {k: _box_scene(d, k) for k in ("matte", "mirror", "glass")}


def _block_mean_se(x: np.ndarray, b: int = 8):
    """Per-block mean and standard error over [sample][H][W][3] values."""
    n, h, w, c = x.shape
    y = x.reshape(n, h // b, b, w // b, b, c).transpose(1, 3, 5, 0, 2, 4).reshape(h // b, w // b, c, -1)
    return y.mean(axis=-1), y.std(axis=-1, ddof=1) / np.sqrt(y.shape[-1])


@pytest.mark.gpu
@pytest.mark.parametrize("which,bounces", [("matte", 4), ("mirror", 8), ("matte", 1)])
def test_bdpt_converges_to_the_path_tracer(gpu_ctx, scenes, which, bounces):
    """Paired comparison: both integrators take the same seed, so each (pixel, sample) starts from
    the same film position and camera ray; the per-sample differences (BDPT - PT) have mean zero
    when both are unbiased, and their spread gives the standard error."""
    d, sc = scenes
    scene, subject = sc[which]
    print(f"[bdpt {which} bounces {bounces}]", flush=True)
    gpu_ctx.load_scene(d, scene, subject, W, H)
    spp = 4096
    _, pt_smp, _ = gpu_ctx.render(spp, bounces, 1000, SEED, integrator=INTEGRATOR_MEGAKERNEL, want_samples=True)
    bd_rgb, bd_smp, _ = gpu_ctx.render(spp, bounces, 0, SEED, integrator=INTEGRATOR_BDPT, want_samples=True)
    assert np.isfinite(bd_smp).all() and (bd_smp >= 0).all()
    diff = bd_smp.astype(np.float64) - pt_smp.astype(np.float64)
    m, se = _block_mean_se(diff)
    z = np.abs(m) / np.maximum(se, 1e-12)
    flat = diff.ravel()
    zmean = abs(flat.mean()) / (flat.std() / np.sqrt(flat.size))
    ratio = bd_smp.mean(dtype=np.float64) / pt_smp.mean(dtype=np.float64)
    print(f"  blocks: max z {z.max():.2f}; image mean ratio {ratio:.5f}, z {zmean:.2f}; "
          f"per-sample std pt {pt_smp.std():.3f} bdpt {bd_smp.std():.3f} diff {flat.std():.3f}")
    assert z.max() < 5.0, (which, bounces, float(z.max()))
    assert zmean < 5.0 and abs(ratio - 1) < 0.01, (ratio, zmean)
    assert bd_rgb.mean() > 0


@pytest.mark.gpu
def test_bdpt_is_deterministic_and_progressive(gpu_ctx, scenes):
    d, sc = scenes
    gpu_ctx.load_scene(d, *sc["glass"], W, H)
    a, sa, _ = gpu_ctx.render(16, 5, 0, SEED, integrator=INTEGRATOR_BDPT, want_samples=True)
    b, sb, _ = gpu_ctx.render(16, 5, 0, SEED, integrator=INTEGRATOR_BDPT, want_samples=True)
    assert np.array_equal(sa, sb) and np.array_equal(a, b)
    gpu_ctx.render(16, 5, 0, SEED, integrator=INTEGRATOR_BDPT, sample_end=6, want_image=False)
    c, _, _ = gpu_ctx.render(16, 5, 0, SEED, integrator=INTEGRATOR_BDPT, sample_begin=6)
    assert np.array_equal(a, c)
    # per-pass sizes do not change the samples
    e, _, _ = gpu_ctx.render(16, 5, 0, SEED, integrator=INTEGRATOR_BDPT, spp_per_pass=3)
    assert np.array_equal(a, e)


@pytest.mark.gpu
def test_bdpt_rejects_unsupported_parameters(gpu_ctx, scenes):
    d, sc = scenes
    gpu_ctx.load_scene(d, *sc["matte"], W, H)
    with pytest.raises(LtsError, match="max_bounce"):
        gpu_ctx.render(4, 9, 0, SEED, integrator=INTEGRATOR_BDPT)
    with pytest.raises(LtsError, match="count_traversal"):
        gpu_ctx.render(4, 3, 0, SEED, integrator=INTEGRATOR_BDPT, count_traversal=1)
