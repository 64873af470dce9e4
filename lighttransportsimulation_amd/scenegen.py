"""Seeded synthetic scenes written in the reference's own input formats.

Every workload is produced as files the reference's loader reads:
  * ``app.params``  (core/FileIO.cuh:234-257): W H / spp / bounces / roulette / format / out / scene / subject
  * ``.scene`` / ``.subject`` (core/SceneLoader.cuh:33-193)
  * ``.bin`` meshes (core/FileIO.cuh:20-45): int32 nTri nP nUV nN, float3 P, float2 UV, float3 N,
    int32 idx[9 nTri] ordered (p, uv, n) x 3
  * images as binary PPM (P6), which stb_image (the reference's decoder) also reads.

Generators use numpy's PCG64 streams seeded per scene part, so a workload is a pure
function of its name and seed.
"""
from __future__ import annotations

import os
import struct
from dataclasses import dataclass, field

import numpy as np


# ----------------------------------------------------------------------------- writers
def write_bin(path_noext: str, P, idx_p, N=None, idx_n=None, UV=None, idx_uv=None) -> None:
    """Writes a reference .bin mesh; ``idx_*`` are (nTri, 3) int arrays."""
    P = np.ascontiguousarray(P, dtype=np.float32).reshape(-1, 3)
    idx_p = np.asarray(idx_p, dtype=np.int32).reshape(-1, 3)
    n_tri = idx_p.shape[0]
    N = np.zeros((0, 3), np.float32) if N is None else np.ascontiguousarray(N, dtype=np.float32).reshape(-1, 3)
    UV = np.zeros((0, 2), np.float32) if UV is None else np.ascontiguousarray(UV, dtype=np.float32).reshape(-1, 2)
    idx_n = np.zeros_like(idx_p) if idx_n is None else np.asarray(idx_n, np.int32).reshape(-1, 3)
    idx_uv = np.zeros_like(idx_p) if idx_uv is None else np.asarray(idx_uv, np.int32).reshape(-1, 3)
    idx = np.stack([idx_p, idx_uv, idx_n], axis=2).reshape(n_tri, 9).astype(np.int32)
    with open(path_noext + ".bin", "wb") as f:
        f.write(struct.pack("<4i", n_tri, P.shape[0], UV.shape[0], N.shape[0]))
        f.write(P.astype("<f4").tobytes())
        f.write(UV.astype("<f4").tobytes())
        f.write(N.astype("<f4").tobytes())
        f.write(idx.astype("<i4").tobytes())


def write_ppm(path: str, rgb8: np.ndarray) -> None:
    h, w, _ = rgb8.shape
    with open(path, "wb") as f:
        f.write(b"P6\n%d %d\n255\n" % (w, h))
        f.write(np.ascontiguousarray(rgb8, dtype=np.uint8).tobytes())


def write_app_params(dir_: str, w, h, spp, bounces, rr, fmt, out, scene, subject) -> None:
    with open(os.path.join(dir_, "app.params"), "w") as f:
        f.write(f"{w} {h}\n{spp}\n{bounces}\n{rr}\n{fmt}\n{out}\n{scene}\n{subject}\n")


def _f(x) -> str:
    return repr(float(np.float32(x)))


@dataclass
class SceneText:
    """Builds the .scene / .subject token streams."""
    camera: tuple
    transform: tuple = (0, 0, 0, 1, 0, 0, 0, 1)
    subj_images: list = field(default_factory=list)
    subj_meshes: list = field(default_factory=list)
    subj_mats: list = field(default_factory=list)
    images: list = field(default_factory=list)
    meshes: list = field(default_factory=list)
    mats: list = field(default_factory=list)
    area_les: list = field(default_factory=list)
    infinite: tuple | None = None

    def write(self, dir_: str, name: str) -> None:
        sc = [" ".join(_f(v) for v in self.camera), " ".join(_f(v) for v in self.transform)]
        sc.append(" ".join([str(len(self.images))] + [f"{w} {p}" for w, p in self.images]))
        sc.append(" ".join([str(len(self.meshes))] + self.meshes))
        sc.append(str(len(self.mats)))
        sc += self.mats
        sc.append(" ".join([str(len(self.area_les))] + [" ".join(_f(v) for v in le) for le in self.area_les]))
        sc.append("1 " + " ".join(_f(v) for v in self.infinite) if self.infinite is not None else "0")
        with open(os.path.join(dir_, name + ".scene"), "w") as f:
            f.write("\n".join(sc) + "\n")
        sb = [" ".join([str(len(self.subj_images))] + [f"{w} {p}" for w, p in self.subj_images]),
              " ".join([str(len(self.subj_meshes))] + self.subj_meshes), str(len(self.subj_mats))]
        sb += self.subj_mats
        with open(os.path.join(dir_, name + ".subject"), "w") as f:
            f.write("\n".join(sb) + "\n")


def mat_matte(rgb, roughness=0.0) -> str:
    return f"L c {_f(rgb[0])} {_f(rgb[1])} {_f(rgb[2])} {_f(roughness)}"


# ----------------------------------------------------------------------------- geometry
def icosphere(subdiv: int):
    t = (1.0 + 5 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    verts = [tuple(x) for x in v]
    for _ in range(subdiv):
        cache = {}
        vl = verts

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = (np.array(vl[a]) + np.array(vl[b])) / 2.0
                m /= np.linalg.norm(m)
                cache[key] = len(vl)
                vl.append(tuple(m))
            return cache[key]

        nf = []
        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        f = np.array(nf, dtype=np.int64)
    return np.array(verts, dtype=np.float64), f


def quad(c, u, v):
    """Two triangles spanning c +- u +- v, wound so that the normal is cross(u, v)."""
    c, u, v = (np.asarray(a, np.float64) for a in (c, u, v))
    P = np.array([c - u - v, c + u - v, c + u + v, c - u + v])
    n = np.cross(u, v)
    n /= np.linalg.norm(n)
    return P, np.array([[0, 1, 2], [0, 2, 3]]), np.tile(n, (4, 1))


def _nonneg_zero(a):
    a = np.asarray(a, np.float32)
    return np.where(a == 0, np.float32(0), a)  # no -0.0 anywhere in generated data


# ----------------------------------------------------------------------------- C1 scene
def gen_cornell_sphere(dir_: str, width=256, height=256, spp=16, bounces=5, rr=3, subdiv=5) -> dict:
    """Seeded icosphere (20,480 tris at subdiv 5, vertex radius 1 + N(0, 0.01)) in a 5-quad
    Cornell-style box with one ceiling quad light of radiance (15, 15, 15); Lambertian
    albedos U[0.2, 0.8]^3; pinhole camera, 60 degree vertical field of view."""
    os.makedirs(os.path.join(dir_, "c1"), exist_ok=True)
    rng1 = np.random.Generator(np.random.PCG64(1))
    rng2 = np.random.Generator(np.random.PCG64(2))
    V, F = icosphere(subdiv)
    r = 1.0 + rng1.normal(0.0, 0.01, size=V.shape[0])
    P = V * r[:, None]
    P[:, 1] -= 0.9
    N = V.copy()
    write_bin(os.path.join(dir_, "c1", "sphere"), _nonneg_zero(P), F, N=_nonneg_zero(N), idx_n=F)
    walls = {
        "floor": ((0, -2, 0), (2, 0, 0), (0, 0, -2)),
        "ceiling": ((0, 2, 0), (2, 0, 0), (0, 0, 2)),
        "back": ((0, 0, 2), (2, 0, 0), (0, 2, 0)),
        "left": ((-2, 0, 0), (0, 0, 2), (0, 2, 0)),
        "right": ((2, 0, 0), (0, 2, 0), (0, 0, 2)),
    }
    st = SceneText(camera=(0, 0, -5.4, 0, 0, 0, 0, 1, 0, 60, 5.4))
    st.subj_meshes = ["c1/sphere"]
    st.subj_mats = [mat_matte(rng2.uniform(0.2, 0.8, 3))]
    for name, (c, u, v) in walls.items():
        Pq, Fq, Nq = quad(c, u, v)
        write_bin(os.path.join(dir_, "c1", name), _nonneg_zero(Pq), Fq, N=_nonneg_zero(Nq), idx_n=Fq)
        st.meshes.append(f"c1/{name}")
        st.mats.append(mat_matte(rng2.uniform(0.2, 0.8, 3)))
    Pl, Fl, Nl = quad((0, 1.98, 0), (0.5, 0, 0), (0, 0, 0.5))
    write_bin(os.path.join(dir_, "c1", "light"), _nonneg_zero(Pl), Fl, N=_nonneg_zero(Nl), idx_n=Fl)
    st.meshes.append("c1/light")
    st.mats.append(mat_matte((0.5, 0.5, 0.5)))
    st.area_les = [(15.0, 15.0, 15.0)]
    st.write(dir_, "c1/cornell")
    write_app_params(dir_, width, height, spp, bounces, rr, 1, "c1_out", "c1/cornell", "c1/cornell")
    return dict(scene="c1/cornell", subject="c1/cornell", width=width, height=height, spp=spp,
                max_bounce=bounces, rr_start=rr, stratified=1, seed=0x5EED)


# ----------------------------------------------------------------------------- soups (LBVH)
def triangle_soup(n: int, seed: int = 3) -> np.ndarray:
    """n triangles: centroids U[0,1)^3, edge scale LogNormal(-5, 0.7), uniformly random
    orientation (Shoemake quaternions). Returns float32 (n, 9) with no -0.0 entries."""
    rng = np.random.Generator(np.random.PCG64(seed))
    c = rng.random((n, 3))
    s = rng.lognormal(-5.0, 0.7, n)
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    q = np.stack([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                  np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3)], axis=1)
    x, y, z, w = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    R = np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], 1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], 1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], 1)], 1)
    base = np.array([[0.0, 0.57735027, 0.0], [-0.5, -0.28867513, 0.0], [0.5, -0.28867513, 0.0]])
    verts = c[:, None, :] + s[:, None, None] * np.einsum("nij,vj->nvi", R, base)
    return _nonneg_zero(verts.reshape(n, 9))
