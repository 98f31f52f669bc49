"""Pins of the oracle's texture pipeline beyond repeat + bilinear: sampler
addressing and border colours, the nearest filter, cube-map face selection
and blitImage.

The CUDA path (ShaderBuiltinCuda.cuh:96-231, TileRasterImageOpInvocationsCuda.cu:36-131)
is the behavioural spec; nothing of it can be compiled here, so the oracle is
held to what the definitions fix:
* addressing: the texel index sequence each mode produces for a level four
  texels wide, written out as a table (repeat 0123 0123, mirrored repeat
  0123 3210, clamp 0000 0123 3333, border, mirror-clamp 3210 0123 3333);
* bilinear filtering across the border blends half the border colour in at
  u = 0 (the texel-centre convention puts texel 0's centre at 1/8);
* the cube-map face table (Vulkan's, which the CUDA code follows): a
  direction's major axis picks the face, the other two components give
  (u, v) as written per face in the table below;
* blit: an identity blit copies, a 2:1 nearest blit takes every other texel,
  a linear blit of a ramp is the ramp, a constant stays constant.
"""
import numpy as np
import pytest

from ifrit_hip import scenes
from oracle import oracle as O

# a 4x1 level whose texels are distinguishable: red = 10 + 50 i
TEX4 = np.array([[[10 + 50 * i, 7, 7, 255] for i in range(4)]], np.uint8)
BLACK, WHITE = (0.0, 0.0, 0.0, 1.0), (1.0, 1.0, 1.0, 1.0)

# texel index for i = -6 .. 9 (each entry the texel covering (i + 0.5) / 4), or -1 for the border
ADDRESS_TABLE = {
    0: [2, 3, 0, 1, 2, 3, 0, 1, 2, 3, 0, 1, 2, 3, 0, 1],     # repeat
    1: [2, 3, 3, 2, 1, 0, 0, 1, 2, 3, 3, 2, 1, 0, 0, 1],     # mirrored repeat
    2: [0, 0, 0, 0, 0, 0, 0, 1, 2, 3, 3, 3, 3, 3, 3, 3],     # clamp to edge
    3: [-1, -1, -1, -1, -1, -1, 0, 1, 2, 3, -1, -1, -1, -1, -1, -1],  # clamp to border
    4: [3, 3, 3, 2, 1, 0, 0, 1, 2, 3, 3, 3, 3, 3, 3, 3],     # mirror clamp to edge
}


@pytest.mark.parametrize("mode", sorted(ADDRESS_TABLE))
@pytest.mark.parametrize("border", [BLACK, WHITE])
def test_address_modes_nearest(mode, border):
    for k, i in enumerate(range(-6, 10)):
        u = (i + 0.5) / 4
        c, _ = O.sample_tex_mode(TEX4, 4, 1, 1, 0, 1, mode, 2, 1, border, u, 0.5, 0, 0, 0, 0)
        want = ADDRESS_TABLE[mode][k]
        expect = np.array(border) if want < 0 else TEX4[0, want] / 255.0
        assert np.allclose(c, expect), (mode, i, c, expect)


def test_bilinear_across_the_border():
    # u = 0 sits half way between texel centres -1/8 (border) and 1/8 (texel 0)
    for border in (BLACK, WHITE):
        c, _ = O.sample_tex_mode(TEX4, 4, 1, 1, 0, 1, 3, 2, 0, border, 0.0, 0.5, 0, 0, 0, 0)
        want = 0.5 * np.array(border) + 0.5 * TEX4[0, 0] / 255.0
        assert np.allclose(c, want)
    # clamp to edge: the same point is texel 0 itself
    c, _ = O.sample_tex_mode(TEX4, 4, 1, 1, 0, 1, 2, 2, 0, BLACK, 0.0, 0.5, 0, 0, 0, 0)
    assert np.allclose(c, TEX4[0, 0] / 255.0)


def test_repeat_bilinear_matches_the_default_sampler():
    rng = np.random.default_rng(3)
    lvl0 = rng.integers(0, 256, (8, 8, 4), dtype=np.uint8)
    tex = scenes.build_mips(lvl0)
    for _ in range(50):
        u, v = rng.uniform(-2, 3, 2)
        d = rng.normal(0, 0.05, 4)
        a, _ = O.sample_tex(tex, 8, 8, 4, 2, 8, u, v, *d)
        b, _ = O.sample_tex_mode(tex, 8, 8, 4, 2, 8, 0, 0, 0, BLACK, u, v, *d)
        assert np.array_equal(a, b)


# (face, major axis index, sign, sc, tc) with sc, tc as the components of the direction that give
# (u, v) = (sc / |ma| + 1) / 2, (tc / |ma| + 1) / 2 (Vulkan's cube-map face table)
FACE_TABLE = [(0, 0, +1, lambda d: -d[2], lambda d: -d[1]), (1, 0, -1, lambda d: d[2], lambda d: -d[1]),
              (2, 1, +1, lambda d: d[0], lambda d: d[2]), (3, 1, -1, lambda d: d[0], lambda d: -d[2]),
              (4, 2, +1, lambda d: d[0], lambda d: -d[1]), (5, 2, -1, lambda d: -d[0], lambda d: -d[1])]


def _face_texture(size=8):
    """Cube map whose texel (face, x, y) is (40 face + 10, 30 x, 30 y, 255): the lookup result names
    the face and the texel."""
    faces = []
    for f in range(6):
        t = np.zeros((size, size, 4), np.uint8)
        t[..., 0] = 40 * f + 10
        t[..., 1] = (30 * np.arange(size))[None, :]
        t[..., 2] = (30 * np.arange(size))[:, None]
        t[..., 3] = 255
        faces.append(t)
    return scenes.layered_mips(faces)


def test_cube_face_table():
    size = 8
    tex = _face_texture(size)
    rng = np.random.default_rng(4)
    for face, axis, sign, sc, tc in FACE_TABLE:
        for _ in range(20):
            d = rng.uniform(-0.9, 0.9, 3)
            d[axis] = sign * rng.uniform(1.0, 2.0)
            u = (sc(d) / abs(d[axis]) + 1) / 2
            v = (tc(d) / abs(d[axis]) + 1) / 2
            c, flag = O.cube_lookup(tex, size, scenes.num_levels(size, size), d, 0.0, nearest=1)
            assert not flag
            got = np.round(c * 255).astype(int)
            assert got[0] == 40 * face + 10, (face, d)
            assert got[1] == 30 * int(np.floor(u * size)) and got[2] == 30 * int(np.floor(v * size)), (face, d)


def test_cube_lookup_levels():
    # a constant colour per level: the lod blends the two levels around it
    size = 4
    levels = scenes.num_levels(size, size)
    chain = []
    for l in range(levels):
        n = max(1, size >> l)
        chain.append(np.full((6, n, n, 4), [20 + 60 * l, 0, 0, 255], np.uint8).reshape(-1))
    tex = np.concatenate(chain)
    c, _ = O.cube_lookup(tex, size, levels, (0.3, 1.0, -0.2), 1.25)
    assert np.isclose(c[0] * 255, 20 + 60 * 1.25)


def test_blit_identity_and_halving():
    rng = np.random.default_rng(5)
    src = rng.integers(0, 256, (16, 12, 4), dtype=np.uint8)
    dst = np.zeros_like(src)
    assert np.array_equal(O.blit(src, dst, (0, 0, 12, 16), (0, 0, 12, 16), linear=False), src)
    assert np.array_equal(O.blit(src, dst, (0, 0, 12, 16), (0, 0, 12, 16), linear=True), src)
    half = O.blit(src, np.zeros((8, 6, 4), np.uint8), (0, 0, 12, 16), (0, 0, 6, 8), linear=False)
    assert np.array_equal(half, src[::2, ::2])


def test_blit_linear_ramp_and_constant():
    ramp = np.zeros((4, 8, 4), np.uint8)
    ramp[..., 0] = (32 * np.arange(8))[None, :]
    ramp[..., 3] = 200
    out = O.blit(ramp, np.zeros((4, 16, 4), np.uint8), (0, 0, 8, 4), (0, 0, 16, 4))
    x = np.arange(16)
    assert np.array_equal(out[0, :, 0], np.minimum(np.floor(16 * x + 0.5), 224).astype(np.uint8))
    assert np.all(out[..., 3] == 200)
    # into a sub-region: texels outside it are kept
    dst = np.full((10, 10, 4), 9, np.uint8)
    out = O.blit(ramp, dst, (2, 1, 6, 3), (3, 4, 7, 8))
    assert np.all(out[:4] == 9) and np.all(out[:, :3] == 9)
    assert np.all(out[4:8, 3:7, 3] == 200)
