"""Workload generators, texture formats and the FP64 shading oracle.

The configs are the build's own synthetic workloads (the reference ships no
scenes for them); their sizes are fixed by the workload definitions.  Mip
generation has two independent implementations (numpy in the package, C in
the oracle) that must agree bit for bit; the shading oracle is checked
against a closed form.
"""
import math

import numpy as np
import pytest

from ifrit_hip import scenes
from oracle import oracle as O


def test_config1_sizes():
    s = scenes.config1()
    assert (s.width, s.height, s.samples) == (640, 480, 1)
    assert s.vertices.shape == (65 * 33 + 4, 8)
    assert s.num_triangles == 64 * 32 * 2 + 2
    assert s.indices.min() >= 0 and s.indices.max() < s.vertices.shape[0]
    t = s.textures[0]
    assert (t.width, t.height, t.levels) == (512, 512, 10)
    assert t.texels.size == 4 * sum(max(1, 512 >> l) ** 2 for l in range(10))


def test_config2_sizes_and_determinism():
    a = scenes.config2()
    assert (a.width, a.height, a.samples) == (1920, 1080, 8)
    assert a.vertices.shape == (655_872, 8)
    assert a.num_triangles == 1_310_720
    assert a.tris_per_material == 5120 and len(a.textures) == 256
    assert all(t.levels == 12 and t.texels is None for t in a.textures)
    assert a.indices.max() < a.vertices.shape[0]
    b = scenes.config2()
    assert np.array_equal(a.vertices, b.vertices) and np.array_equal(a.indices, b.indices)
    assert [t.fbm_seed for t in a.textures] == [t.fbm_seed for t in b.textures]
    assert len(a.lights) == 4


def test_config3_sizes():
    s = scenes.config3()
    assert (s.width, s.height, s.samples) == (1920, 1080, 4)
    assert s.vertices.shape[0] == 2049 * 1025 + 48
    assert s.num_triangles == 4_194_320
    assert s.indices.max() < s.vertices.shape[0]


def test_icosphere_level4_counts():
    p, uv, f = scenes.icosphere(4)
    assert p.shape[0] == 2562 and f.shape[0] == 5120
    assert np.allclose(np.linalg.norm(p, axis=1), 1.0, atol=1e-6)


def test_perspective_maps_near_and_far_planes():
    near, far = 0.1, 100.0
    m = scenes.perspective(math.radians(60), 16 / 9, near, far)
    for z, ndc in ((near, 0.0), (far, 1.0), (1.0, None)):
        c = m @ np.array([0.0, 0.0, z, 1.0])
        assert c[3] == pytest.approx(z)  # w = view-space z (left-handed)
        if ndc is not None:
            assert c[2] / c[3] == pytest.approx(ndc, abs=1e-12)
    # look_at puts the target on +z in view space
    v = scenes.look_at((1, 2, 3), (1, 2, 10), (0, 1, 0)) @ np.array([1, 2, 10, 1.0])
    assert v[:3] == pytest.approx([0, 0, 7])


@pytest.mark.parametrize("w,h", [(64, 64), (37, 20), (1, 9), (256, 8)])
def test_mip_chains_agree_between_implementations(w, h):
    rng = np.random.default_rng(w * 1000 + h)
    lvl0 = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    a = scenes.build_mips(lvl0)
    levels = scenes.num_levels(w, h)
    total = sum(max(1, w >> l) * max(1, h >> l) for l in range(levels)) * 4
    assert a.size == total
    b = np.zeros(total, np.uint8)
    b[: lvl0.size] = lvl0.reshape(-1)
    O.lib().oracle_build_mips(b.ctypes.data, w, h, levels)
    assert np.array_equal(a, b)


def test_mip_box_filter_known_answer():
    lvl0 = np.array([[[0, 10, 255, 1], [1, 11, 255, 2]], [[2, 12, 254, 3], [4, 13, 0, 4]]], np.uint8)
    m = scenes.build_mips(lvl0)
    # (a + b + c + d + 2) >> 2 per channel
    assert list(m[16:20]) == [(0 + 1 + 2 + 4 + 2) >> 2, (10 + 11 + 12 + 13 + 2) >> 2, (255 + 255 + 254 + 0 + 2) >> 2,
                              (1 + 2 + 3 + 4 + 2) >> 2]


def test_fbm_texture_is_deterministic_and_seeded():
    a = scenes.fbm_texture(64, 5)
    assert a.shape == (64, 64, 4) and a.dtype == np.uint8
    assert np.array_equal(a, scenes.fbm_texture(64, 5))
    assert not np.array_equal(a, scenes.fbm_texture(64, 6))
    assert np.all(a[..., 3] == 255)
    assert a[..., :3].std() > 5


def _facing_quad_scene(albedo, ambient, kd, light_color, W=48, H=40):
    """A quad at z = 5 facing a camera at the origin that looks along +z,
    normal towards the camera, a constant texture and one directional light
    shining along the view direction (n . l = 1 everywhere)."""
    verts = np.array([[-1, -1, 5, 0, 0, -1, 0, 0], [1, -1, 5, 0, 0, -1, 1, 0], [1, 1, 5, 0, 0, -1, 1, 1],
                      [-1, 1, 5, 0, 0, -1, 0, 1]], np.float32)
    tex = np.tile(np.array(albedo, np.uint8), (8, 8, 1))
    mvp = (scenes.perspective(math.radians(60), W / H, 0.1, 100) @ scenes.look_at((0, 0, 0), (0, 0, 1), (0, 1, 0)))
    s = scenes.Scene("quad", W, H, 1, verts, np.array([0, 1, 2, 0, 2, 3], np.int32), mvp.astype(np.float32),
                     program=scenes.PROGRAM_BLINN_PHONG, sampler=scenes.SAMPLER_TRILINEAR, camera=(0, 0, 0),
                     ambient=ambient, kd=kd, ks=0.0, lights=[((0.0, 0.0, -1.0, 0.0), light_color)],
                     textures=[scenes.Texture(8, 8, 4, scenes.build_mips(tex))])
    return s


def test_shading_oracle_closed_form_lambert():
    albedo = (200, 120, 40, 255)
    s = _facing_quad_scene(albedo, 0.1, 0.7, (0.9, 0.8, 1.0, 0.0))
    f, _ = O.render_scene(s)
    if (f.ids == 0xFFFFFFFF).all():
        s.indices = s.indices.reshape(2, 3)[:, [0, 2, 1]].reshape(-1).copy()
        f, _ = O.render_scene(s)
    rgba, flags = f.shade(s)
    hit = (f.ids != 0xFFFFFFFF).reshape(s.height, s.width)
    assert hit.sum() > 100
    lc = (0.9, 0.8, 1.0)
    expect = [math.floor(min(1.0, albedo[c] / 255 * (0.1 + 0.7 * lc[c])) * 255 + 0.5) for c in range(3)]
    got = rgba[hit][:, :3]
    assert np.all(got == np.array(expect, np.uint8)), (np.unique(got, axis=0), expect)
    assert np.all(rgba[hit][:, 3] == 255)


def test_flat_program_colours_by_primitive_hash():
    s = _facing_quad_scene((1, 2, 3, 255), 0.1, 0.7, (1, 1, 1, 0))
    s.program = scenes.PROGRAM_FLAT
    s.textures = []
    f, _ = O.render_scene(s)
    if (f.ids == 0xFFFFFFFF).all():
        s.indices = s.indices.reshape(2, 3)[:, [0, 2, 1]].reshape(-1).copy()
        f, _ = O.render_scene(s)
    rgba, _ = f.shade(s)
    ids = f.ids.reshape(s.height, s.width)
    for p in (0, 1):
        h = O.flat_hash(p)
        px = rgba[ids == p]
        assert px.size and np.all(px[:, 0] == (h & 255)) and np.all(px[:, 1] == ((h >> 8) & 255))


def test_orthographic_maps_the_box_to_the_unit_cube():
    m = scenes.orthographic(6.0, 3.0, 1.0, 121.0)
    for p, want in (((6.0, -3.0, 1.0), (1.0, -1.0, 0.0)), ((-6.0, 3.0, 121.0), (-1.0, 1.0, 1.0)),
                    ((0.0, 0.0, 61.0), (0.0, 0.0, 0.5))):
        c = m @ np.array([*p, 1.0])
        assert c[3] == 1.0
        assert np.allclose(c[:3], want)


def test_config4_views_shapes():
    v = scenes.config4_views()
    assert len(v) == 64
    kinds = [x["kind"] for x in v]
    assert kinds.count("cascade") == 4 and kinds.count("cube") == 6 and kinds.count("camera") == 54
    assert all(x["width"] == x["height"] == 4096 and not x["shade"] for x in v[:4])
    assert all(x["width"] == x["height"] == 1024 and not x["shade"] for x in v[4:10])
    assert all(256 <= x["width"] <= 1920 and 256 <= x["height"] <= 1080 and x["shade"] for x in v[10:])
    pixels = sum(x["width"] * x["height"] for x in v)
    assert 100e6 < pixels < 130e6
    # cube faces: the view direction maps to the image centre at w = distance
    for x in v[4:10]:
        m = np.asarray(x["mvp"], np.float64)
        fwd = np.linalg.inv(m) @ np.array([0.0, 0.0, 0.5, 1.0])
        d = fwd[:3] / fwd[3]
        assert np.isclose(np.abs(d).max(), np.linalg.norm(d))  # along one axis
    # the widest cascade holds the whole scene: sphere centres in [-20, 20]^3 plus radius 3
    m = np.asarray(v[3]["mvp"], np.float64)
    corners = np.array([[sx, sy, sz, 1.0] for sx in (-23, 23) for sy in (-23, 23) for sz in (-23, 23)])
    c = corners @ m.T
    assert np.all(np.abs(c[:, :2]) <= 1.0) and np.all((c[:, 2] >= 0) & (c[:, 2] <= 1))
