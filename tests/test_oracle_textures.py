"""The filtered texture pipeline of the oracle against closed forms (no GPU).

A box pyramid of a constant image is that constant at every level; at texel centres a lookup of a
tiny width returns the texel itself; a width of 2 / 8 on an 8x8 image selects level 1 exactly,
whose texels are the means of 2x2 blocks; any width of at least 1 returns the 1x1 top level, the
image mean; repeat wrapping makes the lookup periodic in s and t.
"""
import numpy as np

import oracle_lib


def _img(h=8, w=8, seed=3):
    return np.random.default_rng(seed).uniform(0, 1, (h, w, 3)).astype(np.float32)


def test_constant_image_is_constant_at_every_width():
    img = np.full((7, 13, 3), (0.25, 0.5, 0.75), np.float32)
    rng = np.random.default_rng(1)
    for _ in range(50):
        s, t = rng.uniform(0, 1, 2)
        w = 10.0 ** rng.uniform(-6, 1)
        v = oracle_lib.mip_lookup(img, s, t, w, wrap=2)
        assert np.allclose(v, (0.25, 0.5, 0.75), rtol=1e-6, atol=0)


def test_level_zero_texel_centres_are_exact():
    img = _img()
    for v in range(8):
        for u in range(8):
            got = oracle_lib.mip_lookup(img, (u + 0.5) / 8, (v + 0.5) / 8, 1e-9)
            assert np.array_equal(got, img[v, u])


def test_level_one_is_the_mean_of_2x2_blocks():
    img = _img()
    lvl1 = img.reshape(4, 2, 4, 2, 3).astype(np.float64).mean(axis=(1, 3))
    for t in range(4):
        for s in range(4):
            got = oracle_lib.mip_lookup(img, (s + 0.5) / 4, (t + 0.5) / 4, 2.0 / 8)
            assert np.allclose(got, lvl1[t, s], rtol=1e-6, atol=1e-7)


def test_wide_lookups_return_the_image_mean():
    img = _img()
    for w in (1.0, 3.0, 100.0):
        got = oracle_lib.mip_lookup(img, 0.3, 0.7, w)
        assert np.allclose(got, img.astype(np.float64).mean(axis=(0, 1)), rtol=1e-6, atol=1e-7)


def test_repeat_wrap_is_periodic():
    img = _img(16, 8)
    for u, v in ((0.5, 0.5), (3.5, 11.5), (7.5, 15.5)):
        s, t = u / 8, v / 16
        a = oracle_lib.mip_lookup(img, s, t, 1e-9, wrap=0)
        b = oracle_lib.mip_lookup(img, s + 1, t - 2, 1e-9, wrap=0)
        assert np.array_equal(a, b) and np.array_equal(a, img[int(v), int(u)])

