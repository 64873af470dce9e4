"""OBJ mesh import (parseMeshFile + binMeshWrite, core/FileIO.cuh:20-158) through the C ABI.

The expected .bin bytes are built here from the format itself (int32 nTri nP nUV nN, float32
points, uvs, normals, int32 p/uv/n indices per corner), so the converter is pinned independently
of its own code. The GPU test renders a scene whose meshes exist only as .obj and requires the
image of the same scene loaded from .bin.
"""
import os
import struct

import numpy as np
import pytest

from lighttransportsimulation_amd import LtsError, obj_to_bin


def _bin_bytes(P, UV, N, idx):
    P, UV, N = (np.asarray(a, np.float32).reshape(-1) for a in (P, UV, N))
    idx = np.asarray(idx, np.int32).reshape(-1)
    head = struct.pack("<4i", len(idx) // 9, len(P) // 3, len(UV) // 2, len(N) // 3)
    return head + P.tobytes() + UV.tobytes() + N.tobytes() + idx.tobytes()


def _convert(tmp_path, name, text):
    stem = str(tmp_path / name)
    with open(stem + ".obj", "w") as f:
        f.write(text)
    n = obj_to_bin(stem)
    with open(stem + ".bin", "rb") as f:
        return n, f.read()


def test_obj_with_uvs_normals_and_ignored_statements(tmp_path):
    text = """# exported quad
o quad
v 0 0 0
v 1.5 0 0
v 1.5 2 -0.25
v 0 2 -0.25
vt 0 0
vt 1 0
vt 1 1
vt 0 1
vn 0 0.124035 0.992278
s off
usemtl white
f 1/1/1 2/2/1 3/3/1
f 1/1/1 3/3/1 4/4/1
"""
    n, got = _convert(tmp_path, "quad", text)
    P = [0, 0, 0, 1.5, 0, 0, 1.5, 2, -0.25, 0, 2, -0.25]
    UV = [0, 0, 1, 0, 1, 1, 0, 1]
    N = [0, 0.124035, 0.992278]
    idx = [0, 0, 0, 1, 1, 0, 2, 2, 0, 0, 0, 0, 2, 2, 0, 3, 3, 0]
    assert n == 2 and got == _bin_bytes(P, UV, N, idx)


def test_obj_without_uvs_and_the_stale_field_quirk(tmp_path):
    # no vt: the uv index is 0 whatever the field holds
    n, got = _convert(tmp_path, "tri", "v 0 0 0\nv 1 0 0\nv 0 1 0\nvn 0 0 1\nf 1//1 2//1 3//1\n")
    assert n == 1 and got == _bin_bytes([0, 0, 0, 1, 0, 0, 0, 1, 0], [], [0, 0, 1], [0, 0, 0, 1, 0, 0, 2, 0, 0])
    # with a vt, an empty uv field keeps the integer parsed before it (the point index)
    n, got = _convert(tmp_path, "stale", "v 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0.5 0.5\nvt 1 1\nvt 0 1\nvn 0 0 1\n"
                                         "f 1//1 2/1/1 3//1\n")
    assert got == _bin_bytes([0, 0, 0, 1, 0, 0, 0, 1, 0], [0.5, 0.5, 1, 1, 0, 1], [0, 0, 1],
                             [0, 0, 0, 1, 0, 0, 2, 2, 0])


def _bin_to_obj_text(path_bin):
    orig = open(path_bin, "rb").read()
    nt, npnt, nuv, nn = struct.unpack_from("<4i", orig)
    off = 16
    P = np.frombuffer(orig, np.float32, 3 * npnt, off); off += 12 * npnt
    UV = np.frombuffer(orig, np.float32, 2 * nuv, off); off += 8 * nuv
    N = np.frombuffer(orig, np.float32, 3 * nn, off); off += 12 * nn
    idx = np.frombuffer(orig, np.int32, 9 * nt, off).reshape(nt, 3, 3)
    lines = [f"v {x:.9g} {y:.9g} {z:.9g}" for x, y, z in P.reshape(-1, 3)]
    lines += [f"vt {u:.9g} {v:.9g}" for u, v in UV.reshape(-1, 2)]
    lines += [f"vn {x:.9g} {y:.9g} {z:.9g}" for x, y, z in N.reshape(-1, 3)]
    lines += ["f " + " ".join(f"{p + 1}/{(t + 1) if nuv else ''}/{q + 1}" for p, t, q in tri) for tri in idx]
    return orig, nt, "\n".join(lines) + "\n"


def test_generated_mesh_round_trips_through_obj(tmp_path):
    """A generator mesh written out as OBJ text (9 significant digits, exact for float32) converts
    back to the generator's .bin bytes."""
    from lighttransportsimulation_amd import scenegen
    cfg = scenegen.gen_cornell_sphere(str(tmp_path), width=32, height=32, spp=1)
    stems = sorted(p[:-4] for p in os.listdir(tmp_path / os.path.dirname(cfg["scene"])) if p.endswith(".bin"))
    assert stems
    for s in stems:
        src = str(tmp_path / os.path.dirname(cfg["scene"]) / s)
        orig, nt, text = _bin_to_obj_text(src + ".bin")
        n, got = _convert(tmp_path, "rt_" + s, text)
        assert n == nt and got == orig, s


@pytest.mark.parametrize("text", [
    "v 0 0 0\nv 1 0 0\nv 0 1 0\nvn 0 0 1\nf 1/1 2/1 3/1\n",        # vertex without the normal field
    "v 0 0 0\nv 1 0 0\nv 0 1 0\nvn 0 0 1\nf 1//1 2//1 4//1\n",     # point index out of range
    "v 0 0 x\n",                                                   # float that does not parse
    "v 0 0 0\nv 1 0 0\nv 0 1 0\nvn 0 0 1\nf 1//1 2//1\n",          # truncated face
])
def test_malformed_obj_is_an_error(tmp_path, text):
    with pytest.raises(LtsError):
        _convert(tmp_path, "bad", text)


def test_missing_obj_is_an_error(tmp_path):
    with pytest.raises(LtsError):
        obj_to_bin(str(tmp_path / "nothing"))
    assert not os.path.exists(tmp_path / "nothing.bin")


@pytest.mark.gpu
def test_scene_with_obj_meshes_renders_like_bin(gpu_ctx, tmp_path):
    """lts_load_scene on a scene whose meshes exist only as .obj converts them (writing the .bin
    files, as parseMesh does) and renders the .bin scene's image."""
    import shutil
    from lighttransportsimulation_amd import scenegen
    a, b = tmp_path / "bin", tmp_path / "obj"
    cfg = scenegen.gen_cornell_sphere(str(a), width=48, height=32, spp=4)
    shutil.copytree(a, b)
    mesh_dir = b / os.path.dirname(cfg["scene"])
    originals = {}
    for p in sorted(mesh_dir.glob("*.bin")):
        orig, _, text = _bin_to_obj_text(str(p))
        originals[p] = orig
        p.with_suffix(".obj").write_text(text)
        p.unlink()
    print(f"[obj scene: {len(originals)} meshes as .obj]", flush=True)
    gpu_ctx.load_scene(str(a), cfg["scene"], cfg["subject"], 48, 32)
    ref, _, _ = gpu_ctx.render(4, 5, 3, 0x5EED, stratified=1)
    gpu_ctx.load_scene(str(b), cfg["scene"], cfg["subject"], 48, 32)
    img, _, _ = gpu_ctx.render(4, 5, 3, 0x5EED, stratified=1)
    assert ref.mean() > 0 and np.array_equal(img.view(np.uint32), ref.view(np.uint32))
    for p, orig in originals.items():
        assert p.read_bytes() == orig
