"""Closed-form pins of the oracle's texture sampler.

The sampler follows the CUDA path's spec
(include/ifrit/softgraphics/engine/math/ShaderBuiltinCuda.cuh:96-295), which
cannot be compiled here, so its formulas are pinned by inputs whose answer
the mathematics fixes:

* a texture whose mip levels are distinct constant colours turns the level
  of detail into a colour: trilinear gives c[l0] + (c[l1] - c[l0]) * frac(lod)
  with lod = log2 of the longer footprint axis (SBC:244-250), and the
  anisotropic sampler uses n = ceil(pmax / pmin) taps (capped by maxAniso)
  at lod = log2(pmax / n) (SBC:252-272); a footprint measured with the axes
  mixed, or the lod taken before dividing by n, gives another colour;
* on a linear ramp the anisotropic taps, placed symmetrically at
  (i + 0.5) / n - 0.5 of the major axis, average to the ramp's value at the
  centre (the spec's i / (n + 1) - 0.5 placement shifts that mean);
* on a V-shaped ramp along x, taps along the major axis y leave the value at
  the centre unchanged, while taps along x would raise it;
* bilinear filtering returns a texel at its centre, the mean of two texels
  halfway between them, and wraps at the edge (repeat addressing).
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

BILINEAR, TRILINEAR, ANISO = 0, 1, 2


def level_sizes(w, h, levels):
    return [(max(1, w >> l), max(1, h >> l)) for l in range(levels)]


def constant_levels(w, h, levels):
    """Level l holds the constant colour COL[l]."""
    parts = []
    for l, (lw, lh) in enumerate(level_sizes(w, h, levels)):
        parts.append(np.tile(np.array(COL[l], np.uint8), lw * lh))
    return np.concatenate(parts)


COL = [(8 + 30 * l, 250 - 25 * l, 3 + 17 * l, 255 - 9 * l) for l in range(10)]


def col(l):
    return np.array(COL[l], np.float64) / 255.0


def lerp_levels(lod, levels):
    lod = min(max(lod, 0.0), levels - 1)
    l0 = int(math.floor(lod))
    l1 = min(l0 + 1, levels - 1)
    f = lod - l0
    return col(l0) + (col(l1) - col(l0)) * f


@pytest.mark.parametrize("s", [0.25, 1.0, 1.5, 2.0, 3.0, 4.7, 8.0, 33.0, 64.0, 300.0])
def test_trilinear_lod_is_log2_of_the_isotropic_footprint(s):
    W = H = 64
    levels = 7
    tex = constant_levels(W, H, levels)
    c, flag = O.sample_tex(tex, W, H, levels, TRILINEAR, 16, 0.3, 0.6, s / W, 0.0, 0.0, s / H)
    assert not flag
    np.testing.assert_allclose(c, lerp_levels(math.log2(s), levels), rtol=0, atol=1e-12)


def test_trilinear_footprint_takes_each_derivative_as_a_vector():
    # d(uv)/dx = (4, 4) texels, d(uv)/dy = (-1, 1): |d/dx| = 4 sqrt 2, lod 2.5.  Mixing the axes
    # (|(du/dx, du/dy)|, |(dv/dx, dv/dy)|) would give sqrt 17 on both and lod 2.04.
    W = H = 64
    levels = 7
    tex = constant_levels(W, H, levels)
    c, _ = O.sample_tex(tex, W, H, levels, TRILINEAR, 16, 0.4, 0.4, 4 / W, 4 / H, -1 / W, 1 / H)
    np.testing.assert_allclose(c, lerp_levels(2.5, levels), rtol=0, atol=1e-12)


@pytest.mark.parametrize("major,minor,max_aniso,n,close", [
    (8.0, 2.0, 16, 4, True),     # ratio exactly 4 -> 4 taps, lod log2(8 / 4) = 1 (flagged: FP32 may round up)
    (8.0, 2.0, 2, 2, False),     # capped at 2 taps, lod log2(8 / 2) = 2
    (7.0, 2.0, 16, 4, False),    # ratio 3.5 -> ceil 4 taps, lod log2(7 / 4)
    (40.0, 1.0, 16, 16, False),  # ratio 40 -> capped at 16, lod log2(40 / 16)
    (3.0, 2.5, 16, 2, False),    # ratio 1.2 -> 2 taps
])
def test_anisotropic_tap_count_and_lod(major, minor, max_aniso, n, close):
    W = H = 128
    levels = 8
    tex = constant_levels(W, H, levels)
    for swap in (False, True):
        # the major axis along x, then along y
        dudx, dvdy = (major / W, minor / H) if not swap else (minor / W, major / H)
        c, flag = O.sample_tex(tex, W, H, levels, ANISO, max_aniso, 0.37, 0.52, dudx, 0.0, 0.0, dvdy)
        assert flag == close
        np.testing.assert_allclose(c, lerp_levels(math.log2(major / n), levels), rtol=0, atol=1e-12)


def test_anisotropic_footprint_of_a_rotated_ellipse():
    # |d/dx| = 4 sqrt 2, |d/dy| = 1.1 sqrt 2: ratio 3.64 -> 4 taps, lod log2(sqrt 2) = 0.5
    W = H = 64
    levels = 7
    tex = constant_levels(W, H, levels)
    c, flag = O.sample_tex(tex, W, H, levels, ANISO, 16, 0.5, 0.5, 4 / W, 4 / H, -1.1 / W, 1.1 / H)
    assert not flag
    np.testing.assert_allclose(c, lerp_levels(0.5, levels), rtol=0, atol=1e-12)


def ramp_texture(W, H, fn):
    """One level; channel 0 = fn(x), the other channels constant."""
    x = np.arange(W)
    row = np.stack([fn(x), np.full(W, 7), np.full(W, 9), np.full(W, 255)], axis=1).astype(np.uint8)
    return np.tile(row[None], (H, 1, 1)).reshape(-1)


@pytest.mark.parametrize("n_taps", [2, 3, 4, 7, 16])
def test_anisotropic_taps_are_symmetric_about_the_centre(n_taps):
    # a footprint of n - 0.5 texels by 1 gives n taps (a ratio that is not a whole number, so
    # the tap count is not close to call); one level (lod below 0 clamps to 0); red = x:
    # bilinear reproduces the ramp, so the taps' mean is the ramp at the centre texel, x = 100
    W, H = 256, 4
    tex = ramp_texture(W, H, lambda x: x)
    u, v = 100.5 / W, 1.5 / H
    c, flag = O.sample_tex(tex, W, H, 1, ANISO, 16, u, v, (n_taps - 0.5) / W, 0.0, 0.0, 1.0 / H)
    assert not flag
    assert abs(c[0] * 255.0 - 100.0) < 1e-9


def test_anisotropic_taps_follow_the_major_axis():
    # red = 20 |x - 100|; a 3.5 x 1 footprint gives 4 taps; with the major axis y all taps stay
    # at x = 100 and red is 0
    W, H = 256, 64
    tex = ramp_texture(W, H, lambda x: np.minimum(255, 20 * np.abs(x - 100)))
    c, flag = O.sample_tex(tex, W, H, 1, ANISO, 16, 100.5 / W, 20.5 / H, 1.0 / W, 0.0, 0.0, 3.5 / H)
    assert not flag
    assert abs(c[0]) < 1e-12
    # the same footprint turned onto x puts the taps at x = 100 + 3.5 * (-3/8, -1/8, 1/8, 3/8)
    # = 100 -+ 1.3125, 100 -+ 0.4375, where the bilinear V is exact: mean red 20 * 0.875 = 17.5
    c, flag = O.sample_tex(tex, W, H, 1, ANISO, 16, 100.5 / W, 20.5 / H, 3.5 / W, 0.0, 0.0, 1.0 / H)
    assert not flag
    assert abs(c[0] * 255.0 - 17.5) < 1e-9


def test_bilinear_centres_midpoints_and_wrap():
    W, H = 8, 4
    rng = np.random.default_rng(5)
    tex = rng.integers(0, 256, size=(H, W, 4), dtype=np.uint8)
    flat = tex.reshape(-1)
    for y in range(H):
        for x in range(W):
            c, _ = O.sample_tex(flat, W, H, 1, BILINEAR, 1, (x + 0.5) / W, (y + 0.5) / H, 0, 0, 0, 0)
            np.testing.assert_allclose(c, tex[y, x] / 255.0, rtol=0, atol=1e-12)
    # halfway between columns 2 and 3 of row 1
    c, _ = O.sample_tex(flat, W, H, 1, BILINEAR, 1, 3.0 / W, 1.5 / H, 0, 0, 0, 0)
    np.testing.assert_allclose(c, (tex[1, 2].astype(float) + tex[1, 3]) / 2 / 255.0, rtol=0, atol=1e-12)
    # u = 0 lies halfway between the last and the first column (repeat addressing)
    c, _ = O.sample_tex(flat, W, H, 1, BILINEAR, 1, 0.0, 2.5 / H, 0, 0, 0, 0)
    np.testing.assert_allclose(c, (tex[2, W - 1].astype(float) + tex[2, 0]) / 2 / 255.0, rtol=0, atol=1e-12)
