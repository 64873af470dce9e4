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
    base = np.array([[0.0, 0.57735027, 0.0], [-0.5, -0.28867513, 0.0], [0.5, -0.28867513, 0.0]])
    out = np.empty((n, 9), np.float32)
    step = 1 << 20
    for a in range(0, n, step):
        b = min(n, a + step)
        t1, t2, t3 = u1[a:b], 2 * np.pi * u2[a:b], 2 * np.pi * u3[a:b]
        x, y = np.sqrt(1 - t1) * np.sin(t2), np.sqrt(1 - t1) * np.cos(t2)
        z, w = np.sqrt(t1) * np.sin(t3), np.sqrt(t1) * np.cos(t3)
        R = np.empty((b - a, 3, 3))
        R[:, 0, 0], R[:, 0, 1], R[:, 0, 2] = 1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)
        R[:, 1, 0], R[:, 1, 1], R[:, 1, 2] = 2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)
        R[:, 2, 0], R[:, 2, 1], R[:, 2, 2] = 2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)
        v = c[a:b, None, :] + s[a:b, None, None] * np.matmul(R[:, None, :, :], base[None, :, :, None])[..., 0]
        out[a:b] = v.reshape(b - a, 9)
    out[out == 0] = 0.0  # no -0.0 anywhere in generated data
    return out


# ----------------------------------------------------------------------------- C3 scene
def _value_noise3(p: np.ndarray, seed: int) -> np.ndarray:
    """Seeded trilinear value noise in [-1, 1] on the integer lattice (hash table of 256^3 cells)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    table = rng.uniform(-1.0, 1.0, 4096)
    perm = rng.permutation(4096)
    i = np.floor(p).astype(np.int64)
    f = p - i
    f = f * f * (3 - 2 * f)

    def h(ix, iy, iz):
        return table[perm[(perm[(perm[ix & 4095] + iy) & 4095] + iz) & 4095]]

    out = 0.0
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                w = (f[:, 0] if dx else 1 - f[:, 0]) * (f[:, 1] if dy else 1 - f[:, 1]) * (f[:, 2] if dz else 1 - f[:, 2])
                out = out + w * h(i[:, 0] + dx, i[:, 1] + dy, i[:, 2] + dz)
    return out


def fbm_displacement(p: np.ndarray, octaves=6, omega=0.5, seed=7) -> np.ndarray:
    s, amp, freq = 0.0, 1.0, 2.0
    for o in range(octaves):
        s = s + amp * _value_noise3(p * freq, seed + o)
        amp *= omega
        freq *= 2.0
    return s


def value_noise_image(size: int, seed: int) -> np.ndarray:
    """size x size RGB8 image of seeded multi-octave value noise (smoothstep-interpolated lattices)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    img = np.zeros((size, size, 3))
    amp = 1.0
    for o in range(5):
        g = 4 * 2 ** o
        grid = rng.random((g + 1, g + 1, 3))
        x = np.linspace(0, g, size, endpoint=False)
        i = np.floor(x).astype(int)
        f = x - i
        f = f * f * (3 - 2 * f)
        M = np.zeros((size, g + 1))  # interpolation weights, one row per output coordinate
        M[np.arange(size), i] = 1 - f
        M[np.arange(size), i + 1] += f
        for c in range(3):
            img[:, :, c] += amp * (M @ grid[:, :, c] @ M.T)
        amp *= 0.5
    img /= img.max()
    return (img * 255 + 0.5).astype(np.uint8)


def uv_sphere(n_lat: int, n_lon: int):
    """Vertices on a (n_lat + 1) x (n_lon + 1) latitude/longitude grid and 2 n_lat n_lon triangles."""
    th = np.linspace(0.0, np.pi, n_lat + 1)
    ph = np.linspace(0.0, 2 * np.pi, n_lon + 1)
    T, Pg = np.meshgrid(th, ph, indexing="ij")
    dirs = np.stack([np.sin(T) * np.cos(Pg), np.cos(T), np.sin(T) * np.sin(Pg)], -1).reshape(-1, 3)
    uv = np.stack([Pg / (2 * np.pi), 1 - T / np.pi], -1).reshape(-1, 2)
    idx = np.arange((n_lat + 1) * (n_lon + 1)).reshape(n_lat + 1, n_lon + 1)
    a, b = idx[:-1, :-1], idx[:-1, 1:]
    c, d = idx[1:, :-1], idx[1:, 1:]
    tris = np.stack([np.stack([a, c, b], -1), np.stack([b, c, d], -1)], 2)  # (n_lat, n_lon, 2, 3)
    return dirs, uv, tris


def vertex_normals(P: np.ndarray, F: np.ndarray) -> np.ndarray:
    fn = np.cross(P[F[:, 1]] - P[F[:, 0]], P[F[:, 2]] - P[F[:, 0]])
    N = np.zeros_like(P)
    for k in range(3):
        np.add.at(N, F[:, k], fn)
    ln = np.linalg.norm(N, axis=1, keepdims=True)
    return N / np.where(ln > 0, ln, 1.0)


def gen_displaced_sphere(dir_: str, width=1920, height=1080, spp=64, bounces=8, rr=3, n_lat=500, n_lon=1000,
                         image_size=2048, submeshes=8, seed=11) -> dict:
    """Seeded displaced sphere (radius 1 + 0.1 fbm, 6 octaves, omega 0.5; 2 n_lat n_lon triangles,
    1,000,000 at the defaults) split into longitude sectors with materials drawn from matte /
    glass (IOR 1.5) / microfacet mirror (roughness U[0.05, 0.5]) and matte textures from FBM /
    wrinkled / marble (octaves 4-8, omega U[0.3, 0.7]) / a value-noise image; a ground quad with
    the image texture; infinite light of radiance 0.5 and two quad area lights."""
    sub = os.path.join(dir_, "c3")
    os.makedirs(sub, exist_ok=True)
    rng = np.random.Generator(np.random.PCG64(seed))
    dirs, uv, tris = uv_sphere(n_lat, n_lon)
    r = 1.0 + 0.1 * fbm_displacement(dirs * 1.5, octaves=6, omega=0.5, seed=seed)
    P = dirs * r[:, None]
    F_all = tris.reshape(-1, 3)
    N = vertex_normals(P, F_all)
    img = value_noise_image(image_size, seed + 1)
    write_ppm(os.path.join(sub, "noise.ppm"), img)
    st = SceneText(camera=(0.0, 1.1, -4.2, 0.0, 0.05, 0.0, 0, 1, 0, 40, 4.2))
    st.images = [(0, "c3/noise.ppm")]
    per = n_lon // submeshes
    for k in range(submeshes):
        cols = slice(k * per, n_lon if k == submeshes - 1 else (k + 1) * per)
        F = tris[:, cols].reshape(-1, 3)
        used = np.unique(F)
        remap = np.full(P.shape[0], -1, np.int64)
        remap[used] = np.arange(used.size)
        Fi = remap[F]
        write_bin(os.path.join(sub, f"sphere{k}"), _nonneg_zero(P[used]), Fi, N=_nonneg_zero(N[used]), idx_n=Fi,
                  UV=_nonneg_zero(uv[used]), idx_uv=Fi)
        st.subj_meshes.append(f"c3/sphere{k}")
        kind = ["matte", "glass", "microfacet"][int(rng.integers(0, 3))] if k > 0 else "matte"
        if kind == "glass":
            st.subj_mats.append("G c 1 1 1 c 1 1 1 1.5")
        elif kind == "microfacet":
            col = rng.uniform(0.5, 0.95, 3)
            st.subj_mats.append(f"F c {_f(col[0])} {_f(col[1])} {_f(col[2])} {_f(rng.uniform(0.05, 0.5))}")
        else:
            tex = ["f", "w", "m", "i"][k % 4]
            octv, om = int(rng.integers(4, 9)), rng.uniform(0.3, 0.7)
            if tex == "i":
                spec_ = "i 0"
            elif tex == "m":
                spec_ = f"m {octv} {_f(om)} {_f(rng.uniform(1.5, 4.0))} {_f(rng.uniform(0.5, 2.0))}"
            else:
                spec_ = f"{tex} {octv} {_f(om)}"
            st.subj_mats.append(f"L {spec_} {_f(0.0 if k % 2 == 0 else rng.uniform(5.0, 30.0))}")
    # the subject file carries no images: the image index stays 0 in the scene stream
    Pq, Fq, Nq = quad((0, -1.15, 0), (6, 0, 0), (0, 0, -6))
    write_bin(os.path.join(sub, "ground"), _nonneg_zero(Pq), Fq, N=_nonneg_zero(Nq), idx_n=Fq,
              UV=np.array([[0, 0], [4, 0], [4, 4], [0, 4]], np.float32), idx_uv=Fq)
    st.meshes.append("c3/ground")
    st.mats.append("L w 6 0.5 0")
    for j, (c, s) in enumerate((((-2.0, 3.0, -1.0), 1.0), ((2.5, 2.5, 1.5), 0.7))):
        Pl, Fl, Nl = quad(c, (s, 0, 0), (0, 0, s))  # normal cross(x, z) = -y: emits downwards
        write_bin(os.path.join(sub, f"light{j}"), _nonneg_zero(Pl), Fl, N=_nonneg_zero(Nl), idx_n=Fl)
        st.meshes.append(f"c3/light{j}")
        st.mats.append("L c 0 0 0 0")
    st.area_les = [(40.0, 36.0, 30.0), (12.0, 18.0, 28.0)]
    st.infinite = (0.5, 0.5, 0.5)
    st.write(dir_, "c3/displaced")
    write_app_params(dir_, width, height, spp, bounces, rr, 1, "c3_out", "c3/displaced", "c3/displaced")
    return dict(scene="c3/displaced", subject="c3/displaced", width=width, height=height, spp=spp,
                max_bounce=bounces, rr_start=rr, stratified=0, seed=0x5EED)


def uv_torus(n_u: int, n_v: int, R: float = 1.0, r: float = 0.38):
    """Torus of major radius R and minor radius r on an (n_u + 1) x (n_v + 1) grid: positions, outward
    unit normals of the undisplaced surface, uvs and 2 n_u n_v triangles shaped (n_u, n_v, 2, 3)."""
    u = np.linspace(0.0, 2 * np.pi, n_u + 1)
    v = np.linspace(0.0, 2 * np.pi, n_v + 1)
    U, V = np.meshgrid(u, v, indexing="ij")
    nrm = np.stack([np.cos(V) * np.cos(U), np.sin(V), np.cos(V) * np.sin(U)], -1).reshape(-1, 3)
    ctr = np.stack([R * np.cos(U), np.zeros_like(U), R * np.sin(U)], -1).reshape(-1, 3)
    uv = np.stack([U / (2 * np.pi), V / (2 * np.pi)], -1).reshape(-1, 2)
    idx = np.arange((n_u + 1) * (n_v + 1)).reshape(n_u + 1, n_v + 1)
    a, b = idx[:-1, :-1], idx[:-1, 1:]
    c, d = idx[1:, :-1], idx[1:, 1:]
    tris = np.stack([np.stack([a, b, c], -1), np.stack([b, d, c], -1)], 2)
    return ctr + r * nrm, nrm, uv, tris


def gen_batch_scene(dir_: str, name: str, seed: int, n_tris: int, width: int, height: int, n_lights: int,
                    shape: str = "sphere", spp: int = 32, bounces: int = 8, rr: int = 3, image_size: int = 128,
                    submeshes: int = 4) -> dict:
    """One scene of the batched workload: an FBM-displaced sphere or torus of about n_tris triangles
    split into sectors with materials drawn from matte (FBM / wrinkled / marble / image texture) /
    glass / microfacet mirror, a ground quad, n_lights quad area lights of random size, position and
    radiance, and (for half of the seeds) an infinite light. Files go to <dir_>/<name>/."""
    sub = os.path.join(dir_, name)
    os.makedirs(sub, exist_ok=True)
    rng = np.random.Generator(np.random.PCG64(seed))
    n = max(4, int(round(np.sqrt(n_tris / 4.0))))
    if shape == "torus":
        P0, nrm, uv, tris = uv_torus(2 * n, n)
        P = P0 + nrm * (0.06 * fbm_displacement(P0 * 2.0, octaves=6, omega=0.5, seed=seed))[:, None]
    else:
        nrm, uv, tris = uv_sphere(n, 2 * n)
        P = nrm * (1.0 + 0.1 * fbm_displacement(nrm * 1.5, octaves=6, omega=0.5, seed=seed))[:, None]
    F_all = tris.reshape(-1, 3)
    N = vertex_normals(P, F_all)
    write_ppm(os.path.join(sub, "noise.ppm"), value_noise_image(image_size, seed + 1))
    st = SceneText(camera=(0.0, 1.1, -4.2, 0.0, 0.05, 0.0, 0, 1, 0, 40, 4.2))
    st.images = [(0, f"{name}/noise.ppm")]
    cols_total = tris.shape[1]
    per = max(1, cols_total // submeshes)
    for k in range(submeshes):
        cols = slice(k * per, cols_total if k == submeshes - 1 else (k + 1) * per)
        F = tris[:, cols].reshape(-1, 3)
        used = np.unique(F)
        remap = np.full(P.shape[0], -1, np.int64)
        remap[used] = np.arange(used.size)
        Fi = remap[F]
        write_bin(os.path.join(sub, f"part{k}"), _nonneg_zero(P[used]), Fi, N=_nonneg_zero(N[used]), idx_n=Fi,
                  UV=_nonneg_zero(uv[used]), idx_uv=Fi)
        st.subj_meshes.append(f"{name}/part{k}")
        kind = ["matte", "glass", "microfacet"][int(rng.integers(0, 3))] if k > 0 else "matte"
        if kind == "glass":
            st.subj_mats.append("G c 1 1 1 c 1 1 1 1.5")
        elif kind == "microfacet":
            col = rng.uniform(0.5, 0.95, 3)
            st.subj_mats.append(f"F c {_f(col[0])} {_f(col[1])} {_f(col[2])} {_f(rng.uniform(0.05, 0.5))}")
        else:
            tex = ["f", "w", "m", "i"][int(rng.integers(0, 4))]
            octv, om = int(rng.integers(4, 9)), rng.uniform(0.3, 0.7)
            if tex == "i":
                spec_ = "i 0"
            elif tex == "m":
                spec_ = f"m {octv} {_f(om)} {_f(rng.uniform(1.5, 4.0))} {_f(rng.uniform(0.5, 2.0))}"
            else:
                spec_ = f"{tex} {octv} {_f(om)}"
            st.subj_mats.append(f"L {spec_} {_f(0.0 if rng.random() < 0.5 else rng.uniform(5.0, 30.0))}")
    Pq, Fq, Nq = quad((0, -1.15, 0), (6, 0, 0), (0, 0, -6))
    write_bin(os.path.join(sub, "ground"), _nonneg_zero(Pq), Fq, N=_nonneg_zero(Nq), idx_n=Fq,
              UV=np.array([[0, 0], [4, 0], [4, 4], [0, 4]], np.float32), idx_uv=Fq)
    st.meshes.append(f"{name}/ground")
    st.mats.append("L w 6 0.5 0")
    for j in range(n_lights):
        c = (rng.uniform(-2.5, 2.5), rng.uniform(2.0, 3.5), rng.uniform(-2.5, 2.5))
        s = rng.uniform(0.5, 1.2)
        Pl, Fl, Nl = quad(c, (s, 0, 0), (0, 0, s))  # emits downwards
        write_bin(os.path.join(sub, f"light{j}"), _nonneg_zero(Pl), Fl, N=_nonneg_zero(Nl), idx_n=Fl)
        st.meshes.append(f"{name}/light{j}")
        st.mats.append("L c 0 0 0 0")
        st.area_les.append(tuple(float(x) for x in rng.uniform(8.0, 40.0) * rng.uniform(0.6, 1.0, 3)))
    st.infinite = (0.3, 0.3, 0.3) if seed % 2 == 0 else None
    st.write(dir_, f"{name}/scene")
    return dict(scene=f"{name}/scene", subject=f"{name}/scene", width=width, height=height, spp=spp,
                max_bounce=bounces, rr_start=rr, stratified=0, seed=0x5EED, triangles=2 * n * 2 * n + 2 + 2 * n_lights)


def gen_batch(dir_: str, n_scenes: int = 64, seed: int = 2024, min_tris: float = 1e4, max_tris: float = 2e6,
              resolutions=(128, 256, 512, 1024), spp: int = 32, bounces: int = 8) -> list:
    """The batched workload: n_scenes seeded scenes, triangle counts log-uniform in [min_tris, max_tris],
    square resolutions drawn from `resolutions`, 1-4 area lights, spheres and tori alternating."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for i in range(n_scenes):
        n_tris = int(np.exp(rng.uniform(np.log(min_tris), np.log(max_tris))))
        res = int(resolutions[int(rng.integers(0, len(resolutions)))])
        n_lights = int(rng.integers(1, 5))
        out.append(gen_batch_scene(dir_, f"b{i:02d}", seed + 17 * i, n_tris, res, res, n_lights,
                                   "sphere" if i % 2 == 0 else "torus", spp=spp, bounces=bounces))
    return out
