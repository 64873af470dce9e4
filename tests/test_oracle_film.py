"""The oracle's deterministic film gather against a literal scatter of Camera::addSample (no GPU).

addSample (rendering/Camera.cu:192-218) scatters each sample into the pixels x in
[ceil(p.x - 1), floor(p.x + 1) + 1) (clipped to the film), with the Gaussian table entry at
|x - p.x - 0.5| * 16. The gather visits the owners of each pixel instead; this checks that it sees
exactly the contributions the scatter makes, including the film positions that round to the owner
pixel itself (owner + f == owner for f below half an ulp, which happens from 1024 pixels on and
makes the owner cover the pixel before it), and f = 1 exactly (the owner then covers 3 pixels).
"""
import numpy as np

import oracle_lib as O


def _filter_table():
    t = np.zeros(256, np.float32)
    O.lib().oracle_filter_table(O._p(t))
    return t


def _scatter(W, H, pf, L, table):
    """addSample restated literally, in float32, one sample after another; sums in float64."""
    film = np.zeros((H, W, 4), np.float64)
    spp = pf.shape[0]
    for s in range(spp):
        for oy in range(H):
            for ox in range(W):
                px, py = pf[s, oy, ox]
                p0x = max(int(np.ceil(np.float32(px - np.float32(1)))), 0)
                p0y = max(int(np.ceil(np.float32(py - np.float32(1)))), 0)
                p1x = min(int(np.floor(np.float32(px + np.float32(1)))) + 1, W)
                p1y = min(int(np.floor(np.float32(py + np.float32(1)))) + 1, H)
                for y in range(p0y, p1y):
                    for x in range(p0x, p1x):
                        dx = abs(np.float32((np.float32(x) - px - np.float32(0.5)) * np.float32(16)))
                        dy = abs(np.float32((np.float32(y) - py - np.float32(0.5)) * np.float32(16)))
                        ix, iy = min(int(np.floor(dx)), 15), min(int(np.floor(dy)), 15)
                        w = float(table[iy * 16 + ix])
                        film[y, x, :3] += L[s, oy, ox] * w
                        film[y, x, 3] += w
    return film


def _gather(W, H, pf, L):
    film = np.zeros((H, W, 4), np.float32)
    pfc = np.ascontiguousarray(pf, np.float32)
    Lc = np.ascontiguousarray(L, np.float32)
    O.lib().oracle_film_gather.argtypes = [O.C.c_int32] * 3 + [O.C.c_void_p] * 3
    O.lib().oracle_film_gather(W, H, pf.shape[0], O._p(pfc), O._p(Lc), O._p(film))
    return film


def _positions(W, H, X0, Y0, f):
    """Film positions of a band whose pixel coordinates start at (X0, Y0): pixel + f (float32)."""
    xs = (np.arange(W) + X0).astype(np.float32)
    ys = (np.arange(H) + Y0).astype(np.float32)
    pf = np.zeros((f.shape[0], H, W, 2), np.float32)
    pf[..., 0] = xs[None, None, :] + f[..., 0]
    pf[..., 1] = ys[None, :, None] + f[..., 1]
    return pf


def _check(W, H, pf, L, band_origin=(0, 0)):
    table = _filter_table()
    # the gather runs on a film whose coordinates start at 0; shift the positions of a band that
    # sits at (X0, Y0) of a larger film back after the float32 rounding has happened there
    X0, Y0 = band_origin
    local = pf.copy()
    local[..., 0] -= np.float32(X0)
    local[..., 1] -= np.float32(Y0)
    assert np.array_equal(local[..., 0] + np.float32(X0), pf[..., 0])
    g = _gather(W, H, local, L)
    s = _scatter(W, H, local, L, table)
    assert np.allclose(g[..., 3], s[..., 3], rtol=2e-6, atol=0), np.abs(g[..., 3] - s[..., 3]).max()
    assert np.allclose(g[..., :3], s[..., :3], rtol=2e-5, atol=1e-7)
    return g, s


def test_gather_equals_scatter_for_interior_samples():
    rng = np.random.Generator(np.random.PCG64(11))
    W, H, spp = 9, 7, 3
    f = ((rng.integers(0, 2 ** 24, (spp, H, W, 2)) + 1) / 2 ** 24).astype(np.float32)
    L = rng.uniform(0, 2, (spp, H, W, 3)).astype(np.float32)
    _check(W, H, _positions(W, H, 0, 0, f), L)


def test_gather_sees_film_positions_that_round_onto_the_owner():
    """Owners at x >= 1024 with f = 2^-24 have p == owner exactly: their footprint starts one
    pixel to the left. The band keeps the rounding of the full-width film (coordinates 1030..)."""
    rng = np.random.Generator(np.random.PCG64(12))
    W, H, spp = 10, 8, 2
    X0, Y0 = 1030, 1040
    f = ((rng.integers(0, 2 ** 24, (spp, H, W, 2)) + 1) / 2 ** 24).astype(np.float32)
    f[0, :, ::2, 0] = 2.0 ** -24   # half of the owners round onto themselves in x
    f[1, ::3, :, 1] = 2.0 ** -24   # a third in y
    f[1, 2, 3] = (1.0, 1.0)        # and one film sample at exactly 1 in both axes
    pf = _positions(W, H, X0, Y0, f)
    assert (pf[0, :, ::2, 0] == np.arange(W)[::2] + X0).all()
    L = rng.uniform(0, 2, (spp, H, W, 3)).astype(np.float32)
    g, s = _check(W, H, pf, L, (X0, Y0))
    # those owners cover the pixel before them, which a window of owners x-2..x never visits
    local_x = pf[..., 0] - np.float32(X0)
    owner_x = np.arange(W)[None, None, :]
    starts_left = np.ceil(local_x - np.float32(1)).astype(np.int64) == owner_x - 1
    assert starts_left[:, :, 1:].sum() >= (W // 2 - 1) * H
