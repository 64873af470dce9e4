"""The reference's own scene files (tests/golden/res, copied from its res/ directory) rendered on
the GPU against the oracle, sample by sample, with the app.params settings the reference ships
(10 samples, 5 bounces, roulette from 2; app.params:1-5).

Scenes: platforms/bigBox (area light box + infinite light), platforms/plane (the dome mesh under
the infinite light only) and kernel_box_scene/kernel (a 64^2 PNG texture, Black wrap), with the
diamond (glass, eta 2.419), holder and empty subjects; the empty
scene (sky only) with the diamond; the folding box with its donut (an Oren-Nayar subject and 64
emissive triangles). The reference's JPEG texture (res/blackDragon.jpg, decoded by the oracle with
the reference's own stb_image) appears twice: on the kernel box in place of its PNG, and through
blackDragon.subject's image and material lines on the diamond mesh. The behelit and chicken subjects use the mirror material R
(Material.cuh:78-95) but their meshes are not in the reference tree, so a mirror subject on the
diamond mesh (behelit's material line) stands in for them.
"""
import os
import shutil

import numpy as np
import pytest

import oracle_lib
from lighttransportsimulation_amd import INTEGRATOR_MEGAKERNEL, INTEGRATOR_WAVEFRONT
from test_gpu_render import SEED, _compare

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

CASES = [
    ("res/platforms/bigBox", "res/diamond"),
    ("res/platforms/bigBox", "res/holder"),
    ("res/kernel_box_scene/kernel", "res/diamond"),
    ("res/kernel_box_scene/kernel", "res/holder"),
    ("res/empty", "res/diamond"),
    ("res/foldingbox/folding", "res/foldingbox/donut"),
    ("res/platforms/bigBox", "res/mirror_diamond"),
    ("res/platforms/plane", "res/diamond"),
    ("res/kernel_box_scene/kernel_jpeg", "res/diamond"),
    ("res/platforms/bigBox", "res/blackDragon_diamond"),
]


@pytest.fixture(scope="module")
def project(tmp_path_factory):
    d = tmp_path_factory.mktemp("refres")
    shutil.copytree(os.path.join(GOLDEN, "res"), os.path.join(d, "res"))
    # behelit.subject's material (R c 0.9 0.02 0.02 1) on the diamond mesh
    with open(os.path.join(d, "res", "mirror_diamond.subject"), "w") as f:
        f.write("0\n1\tres/diamond\n1\nR\tc\t0.9 0.02 0.02\t1\n")
    # the kernel box with the reference's JPEG texture (res/blackDragon.jpg) in place of its PNG
    with open(os.path.join(GOLDEN, "res", "kernel_box_scene", "kernel.scene")) as f:
        kernel = f.read()
    assert "res/kernel_box_scene/kernel.png" in kernel
    with open(os.path.join(d, "res", "kernel_box_scene", "kernel_jpeg.scene"), "w") as f:
        f.write(kernel.replace("res/kernel_box_scene/kernel.png", "res/blackDragon.jpg"))
    # blackDragon.subject (its JPEG texture and 'L i 0 0.5' material) on the diamond mesh: the
    # dragon mesh (2 MB) is not among the fixtures
    with open(os.path.join(GOLDEN, "res", "blackDragon.subject")) as f:
        dragon = f.read()
    assert "res/blackDragon\n" in dragon
    with open(os.path.join(d, "res", "blackDragon_diamond.subject"), "w") as f:
        f.write(dragon.replace("res/blackDragon\n", "res/diamond\n"))
    return str(d) + "/"


@pytest.mark.gpu
@pytest.mark.parametrize("scene,subject", CASES)
def test_reference_scene_matches_oracle(gpu_ctx, project, scene, subject):
    W, H, spp, bounces, rr = 120, 90, 10, 5, 2
    print(f"[reference scene {scene} + {subject}]", flush=True)
    gpu_ctx.load_scene(project, scene, subject, W, H)
    o = oracle_lib.OracleScene(project, scene, subject, W, H)
    assert gpu_ctx.scene_info()["tris"] == o.info()["tris"]
    o_rgb, o_smp, _ = o.render(spp, bounces, rr, SEED, want_samples=True)
    for integ, name in ((INTEGRATOR_MEGAKERNEL, "megakernel"), (INTEGRATOR_WAVEFRONT, "wavefront")):
        g_rgb, g_smp, _ = gpu_ctx.render(spp, bounces, rr, SEED, integrator=integ, want_samples=True)
        assert np.isfinite(g_rgb).all()
        assert g_rgb.mean() > 0
        _compare(g_smp, o_smp, g_rgb, o_rgb, f"{scene}+{subject} {name}")
