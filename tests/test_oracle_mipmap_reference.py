"""Pins of the oracle's bug-compatible image texture pipeline on hand-checkable images (no GPU).

What the reference does, read off its source (cited) and restated here as closed forms:

* BlockedArray::operator() (core/MemoryHelper.cuh:88-94) takes the in-block row from block(v)
  instead of offset(v), so the four rows of a 4-row band share one slot per column. The base level
  is written row by row, v ascending (MemoryHelper.cuh:53-58), so level-0 texel (u, v) holds image
  row min(4 floor(v / 4) + 3, h - 1).
* pyramidInitKernel (textures/Mipmap_kernel.cu:23-37) fills level l >= 1 only for s, t < 16 >> l,
  each texel the mean of its 2 x 2 parent block read with Black wrap; its rows alias the same way
  (the oracle resolves the colliding writes as "last row wins"). Everything else stays 0.
* Mipmap::lookUp (Mipmap.cuh:55-72) picks level (L - 1) + log2(width) with L = (int)log2(max(w, h));
  at or above the top it calls texel(level - 1, 0, 0): the arguments are swapped, so it reads texel
  (L - 1, 0) of level 0.
* ImageTexture::evaluate uses the fixed width 0.01 (ImageTexture.cuh:8, 26): on a 2048^2 image that
  is level 3.36, between levels 3 and 4, which hold only their 2 x 2 and 1 x 1 corners.
"""
import math

import numpy as np

import oracle_lib


def _lookup(img, s, t, width, wrap=1):
    return oracle_lib.mip_lookup(img, s, t, width, wrap=wrap, filtered=False)


def _img(h, w, f):
    v, u = np.mgrid[0:h, 0:w]
    val = f(u, v).astype(np.float32)
    return np.stack([val, val + 1000, -val], axis=-1).astype(np.float32)


def _centre(i, res):
    return (i + 0.5) / res


def _width_for_level(l, levels):
    return 2.0 ** (l - (levels - 1))


def test_level0_rows_alias_within_4_row_bands():
    img = _img(16, 16, lambda u, v: 100 * u + v)
    for v in range(16):
        for u in range(16):
            got = _lookup(img, _centre(u, 16), _centre(v, 16), 1e-9)
            row = min(4 * (v // 4) + 3, 15)
            assert np.array_equal(got, img[row, u]), (u, v, got)


def test_level0_partial_last_band_takes_the_last_row():
    # 18 rows: the band of rows 16..19 holds only rows 16 and 17, so its texels read row 17
    img = _img(18, 16, lambda u, v: 100 * u + v)
    got = _lookup(img, _centre(5, 16), _centre(16, 18), 1e-9)
    assert np.array_equal(got, img[17, 5])


def test_levels_hold_block_means_of_the_aliased_rows():
    """Row-constant image (value = row index), 16 x 16, L = 4: level 1 is 7 for t < 4 and 15 for
    t >= 4 (the means of aliased rows 6, 7 and 14, 15); level 2 reads level-1 rows 6, 7 (15); level 3
    (2 x 2, the top, reachable only through the lerp from level 2) reads level-2 rows 2, 3 (15)."""
    img = _img(16, 16, lambda u, v: v)
    L = 4
    for l, res, expect in ((1, 8, lambda t: 7.0 if t < 4 else 15.0), (2, 4, lambda t: 15.0)):
        for t in range(res):
            for s in range(res):
                got = _lookup(img, _centre(s, res), _centre(t, res), _width_for_level(l, L))
                assert got[0] == np.float32(expect(t)), (l, s, t, got)
    # halfway between levels 2 and 3 at the centre of level-3 texel (0, 0): 15 from both
    got = _lookup(img, 0.25, 0.25, 2.0 ** -0.5)
    assert got[0] == np.float32(15.0), got


def test_levels_are_column_box_means():
    """Column-constant image (value = column index): no aliasing across columns, so level l texel s
    is the mean of columns [2^l s, 2^l (s + 1)), i.e. 2^l s + (2^l - 1) / 2."""
    img = _img(16, 16, lambda u, v: u)
    for l, res in ((1, 8), (2, 4)):
        for s in range(res):
            got = _lookup(img, _centre(s, res), _centre(0, res), _width_for_level(l, 4))
            assert got[0] == np.float32(2 ** l * s + (2 ** l - 1) / 2), (l, s, got)
    # level 3 texel 0 (3.5) lerped halfway with level 2 at s = 0.5, between its texels 0 (1.5)
    # and 1 (5.5): 0.5 * 3.5 + 0.5 * 3.5
    got = _lookup(img, 0.25, 0.25, 2.0 ** -0.5)
    assert got[0] == np.float32(3.5), got


def _filled(l, s, t):
    """Level-l texel (s, t), l >= 1, was written (directly or through its aliased row band)."""
    lim = 16 >> l
    return s < lim and 4 * (t // 4) < lim


def test_only_the_corner_of_upper_levels_is_filled():
    """64 x 64 constant image, L = 6: level 1 (32 x 32) holds texels only for s < 8 and row bands
    starting below 8, level 2 for s < 4 and rows 0..3, level 3 for s < 2 and rows 0..3 (rows 2, 3
    alias rows 0, 1), level 4 for s = 0 and rows 0..3; every written texel of a constant image is
    the constant. Level 5 is the top and is never read by a lookup."""
    img = np.ones((64, 64, 3), np.float32)
    L = 6
    for l in (1, 2, 3, 4):
        res = 64 >> l
        for t in range(res):
            for s in range(res):
                got = _lookup(img, _centre(s, res), _centre(t, res), _width_for_level(l, L))
                assert got[0] == (1.0 if _filled(l, s, t) else 0.0), (l, s, t, got)


def test_top_of_the_pyramid_reads_level0_texel_L_minus_1_row_0():
    for h, w in ((16, 16), (64, 64), (32, 64)):
        img = _img(h, w, lambda u, v: 100 * u + v)
        L = int(math.log2(max(w, h)))
        for width in (1.0, 4.0, 2.0 ** -(L - 1) * 2 ** (L - 1)):
            got = _lookup(img, 0.73, 0.21, width)
            assert np.array_equal(got, img[3, L - 1]), (h, w, width, got)


def _bilinear_filled(l, res, st):
    """Bilinear lookup of level l of a constant image (1 where written, 0 elsewhere), Black wrap."""
    s = st[0] * res - 0.5
    t = st[1] * res - 0.5
    s0, t0 = math.floor(s), math.floor(t)
    ds, dt = s - s0, t - t0
    tex = lambda a, b: 1.0 if (0 <= a < res and 0 <= b < res and _filled(l, a, b)) else 0.0
    return ((1 - ds) * (1 - dt) * tex(s0, t0) + (1 - ds) * dt * tex(s0, t0 + 1) + ds * (1 - dt) * tex(s0 + 1, t0) +
            ds * dt * tex(s0 + 1, t0 + 1))


def test_fixed_width_on_a_2048_image_sees_only_the_corner():
    """lookUp(st, 0.01) on a constant 2048^2 image: level 10 + log2(0.01) = 3.356, a lerp of levels 3
    (256^2, filled for s < 2, rows 0..3) and 4 (128^2, filled for s = 0, rows 0..3). Away from that
    corner the texture is black."""
    img = np.ones((2048, 2048, 3), np.float32)
    lv = 10 + math.log2(0.01)
    d = lv - 3
    for s, t in ((0.5, 0.5), (0.1, 0.9), (0.05, 0.05), (0.9, 0.0), (0.01, 0.02), (0.002, 0.004), (0.0, 0.015)):
        expect = (1 - d) * _bilinear_filled(3, 256, (s, t)) + d * _bilinear_filled(4, 128, (s, t))
        got = _lookup(img, s, t, 0.01)
        assert abs(got[0] - expect) < 1e-5, (s, t, got, expect)
        if s >= 0.05 or t >= 0.05:
            assert got[0] == 0.0


def test_repeat_wrap_of_negative_indices_reads_black():
    """Repeat wraps with fmodf on ints, which keeps negative indices negative (Mipmap.cuh:19-20);
    the reference then reads out of bounds, the oracle reads 0. A lookup between texel -1 and 0 at
    level 0 therefore weights only texel 0."""
    img = np.ones((16, 16, 3), np.float32)
    s = 0.25 / 16  # s * 16 - 0.5 = -0.25: texels -1 (weight 0.25) and 0 (weight 0.75)
    got = _lookup(img, s, _centre(5, 16), 1e-9, wrap=0)
    assert abs(got[0] - 0.75) < 1e-6
