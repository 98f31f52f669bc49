"""Closed-form pins of the shading oracle's Blinn-Phong lighting and MSAA resolve.

The FP64 shading oracle (oracle/oracle_shade.cpp) is the judge of the GPU's
colour.  These tests fix the branches the config-2 workload uses -- point
lights with 1 / (1 + k d^2) attenuation, the half-vector specular term and
the sample-weighted resolve -- against geometry whose answer follows from
first principles, not from the oracle's code:

* a quad at z = 5 facing a camera at the origin: the world point under a
  pixel is the intersection of the pixel's ray with the plane z = 5, so the
  angle to the normal and the distance to a light are known per pixel;
* a point light on the quad's axis at distance d: n.l = d / sqrt(r^2 + d^2)
  and d^2 = r^2 + d^2 at radius r;
* a directional light along the normal seen from the origin: the half
  vector's angle is half the view angle, n.h = cos(theta / 2);
* a point light at the camera: l = v = h, n.l = n.h = cos(theta);
* an axis-aligned screen rectangle at 4x / 8x MSAA: the covered sample count
  k of an edge pixel follows from the sample table, and the resolved flat
  colour must be round(k / S * hash byte).

Each closed form is also run against the oracle with a planted mistake
(oracle_set_plant: reflection vector instead of the half vector, 1/(1+kd)
attenuation, a flipped view vector, resolve over covered samples only,
unweighted resolve); the test shows that the mistake fails it.
"""
import math

import numpy as np
import pytest

from ifrit_hip import scenes
from oracle import oracle as O

W, H = 96, 80
Z_QUAD = 5.0


def _quad_scene(albedo, ambient, kd, ks, shininess, lights, W=W, H=H):
    """A 4 x 4 quad at z = 5 with normal (0, 0, -1), camera at the origin
    looking along +z, a constant texture of `albedo`."""
    verts = np.array([[-2, -2, Z_QUAD, 0, 0, -1, 0, 0], [2, -2, Z_QUAD, 0, 0, -1, 1, 0],
                      [2, 2, Z_QUAD, 0, 0, -1, 1, 1], [-2, 2, Z_QUAD, 0, 0, -1, 0, 1]], np.float32)
    tex = np.tile(np.array(albedo, np.uint8), (8, 8, 1))
    mvp = (scenes.perspective(math.radians(60), W / H, 0.1, 100) @ scenes.look_at((0, 0, 0), (0, 0, 1), (0, 1, 0)))
    return scenes.Scene("quad", W, H, 1, verts, np.array([0, 1, 2, 0, 2, 3], np.int32), mvp.astype(np.float32),
                        program=scenes.PROGRAM_BLINN_PHONG, sampler=scenes.SAMPLER_TRILINEAR, camera=(0, 0, 0),
                        ambient=ambient, kd=kd, ks=ks, shininess=shininess, lights=lights,
                        textures=[scenes.Texture(8, 8, 4, scenes.build_mips(tex))])


def _render(scene):
    f, _ = O.render_scene(scene)
    if (f.ids == 0xFFFFFFFF).all():  # culled: the other winding faces the camera
        scene.indices = scene.indices.reshape(-1, 3)[:, [0, 2, 1]].reshape(-1).copy()
        f, _ = O.render_scene(scene)
    return f


def _plane_points(scene, o=0.0):
    """World point on z = 5 under every pixel's shading position (pixel corner
    + o), from the inverse of the float32 view-projection in float64."""
    m = np.asarray(scene.mvp, np.float64)
    inv = np.linalg.inv(m)
    px, py = np.meshgrid(np.arange(scene.width) + o, np.arange(scene.height) + o)
    nx = 2.0 * px / scene.width - 1.0   # x = (ndc + 1) W / 2
    ny = 2.0 * py / scene.height - 1.0  # row 0 is ndc y = -1
    def unproject(nz):
        c = np.stack([nx, ny, np.full_like(nx, nz), np.ones_like(nx)], -1) @ inv.T
        return c[..., :3] / c[..., 3:4]
    a, b = unproject(0.0), unproject(1.0)
    t = (Z_QUAD - a[..., 2]) / (b[..., 2] - a[..., 2])
    return a + (b - a) * t[..., None]


def _to_u8(x):
    return np.floor(np.clip(x, 0.0, 1.0) * 255.0 + 0.5)


def _check(scene, expect_rgb, min_pixels=1000):
    """The oracle's RGB must equal the closed form on every covered pixel
    that is not within 0.02 of a rounding boundary.  Returns the number of
    mismatching pixels (0 on success)."""
    f = _render(scene)
    rgba, flags = f.shade(scene)
    hit = (f.ids != 0xFFFFFFFF).reshape(scene.height, scene.width) & (flags == 0)
    v = np.clip(expect_rgb, 0, 1) * 255.0
    frac = v - np.floor(v)
    clear = np.all(np.abs(frac - 0.5) > 0.02, -1)
    use = hit & clear
    assert use.sum() >= min_pixels, use.sum()
    got = rgba[..., :3][use].astype(np.float64)
    want = _to_u8(expect_rgb[use])
    return int(np.any(got != want, -1).sum())


@pytest.fixture
def plant():
    def set_(k):
        O.lib().oracle_set_plant(k)
    yield set_
    O.lib().oracle_set_plant(0)


def _point_light_case(d, k):
    albedo = np.array([210, 150, 90]) / 255.0
    lc = np.array([0.9, 1.0, 0.8])
    ambient, kd = 0.05, 0.9
    s = _quad_scene((210, 150, 90, 255), ambient, kd, 0.0, 16.0,
                    [((0.0, 0.0, Z_QUAD - d, 1.0), (lc[0], lc[1], lc[2], k))])
    P = _plane_points(s)
    r2 = P[..., 0] ** 2 + P[..., 1] ** 2
    ndl = d / np.sqrt(r2 + d * d)          # light on the axis at height d above the plane
    att = 1.0 / (1.0 + k * (r2 + d * d))  # squared distance to the light
    expect = albedo * (ambient + kd * (ndl * att)[..., None] * lc)
    return s, expect


@pytest.mark.parametrize("d,k", [(1.5, 0.2), (3.0, 0.05), (0.8, 1.0)])
def test_point_light_attenuation_closed_form(d, k, plant):
    s, expect = _point_light_case(d, k)
    assert _check(s, expect) == 0
    # at the centre pixel the light is straight above: n.l = 1, att = 1 / (1 + k d^2)
    c = _plane_points(s)[s.height // 2, s.width // 2]
    assert abs(c[0]) < 1e-9 and abs(c[1]) < 1e-9
    plant(2)  # 1 / (1 + k d)
    assert _check(s, expect) > 100


def _specular_case(shininess, ks=0.5):
    albedo = np.array([60, 90, 30]) / 255.0
    lc = np.array([1.0, 0.9, 0.7])
    ambient, kd = 0.05, 0.3
    # directional light along the normal, towards the camera side
    s = _quad_scene((60, 90, 30, 255), ambient, kd, ks, shininess, [((0.0, 0.0, -1.0, 0.0), (lc[0], lc[1], lc[2], 0.0))])
    P = _plane_points(s)
    cos_t = Z_QUAD / np.linalg.norm(P, axis=-1)   # angle between the normal and the view vector
    ndh = np.sqrt((1.0 + cos_t) / 2.0)            # the half vector bisects it: cos(theta / 2)
    expect = albedo * ambient + lc * (kd * albedo + ks * (ndh ** shininess)[..., None])
    return s, expect, cos_t


@pytest.mark.parametrize("shininess", [8.0, 64.0])
def test_specular_half_vector_closed_form(shininess, plant):
    s, expect, cos_t = _specular_case(shininess)
    assert cos_t.min() < 0.985  # the frame spans view angles up to ~15 degrees
    assert _check(s, expect) == 0
    # the centre pixel sees the specular peak: n.h = 1 adds exactly ks * light colour
    plant(1)  # reflection vector
    assert _check(s, expect) > 100
    plant(3)  # flipped view vector
    assert _check(s, expect, min_pixels=0) > 100


def test_specular_exponent_is_pinned():
    s, expect, _ = _specular_case(64.0)
    _, other, _ = _specular_case(32.0)
    assert _check(s, expect) == 0
    assert _check(s, other) > 100


def test_point_light_at_camera_closed_form(plant):
    albedo = np.array([120, 200, 160]) / 255.0
    lc = np.array([1.0, 0.8, 0.9])
    k, ambient, kd, ks, sh = 0.01, 0.08, 0.8, 0.4, 20.0
    s = _quad_scene((120, 200, 160, 255), ambient, kd, ks, sh, [((0.0, 0.0, 0.0, 1.0), (lc[0], lc[1], lc[2], k))])
    P = _plane_points(s)
    dist2 = np.sum(P * P, -1)
    cos_t = Z_QUAD / np.sqrt(dist2)  # l = v = h: n.l = n.h = cos(theta)
    att = 1.0 / (1.0 + k * dist2)
    expect = albedo * ambient + (att[..., None] * lc) * (kd * cos_t[..., None] * albedo + ks * (cos_t ** sh)[..., None])
    assert _check(s, expect) == 0
    plant(2)
    assert _check(s, expect) > 100


# --- MSAA resolve ---------------------------------------------------------------------------------------------

RW, RH = 64, 48
X0, X1, Y0, Y1 = 10.3, 41.7, 7.55, 35.2


def _rect_scene(samples):
    """An axis-aligned rectangle in pixel coordinates, two triangles sharing
    the diagonal (X0, Y0)-(X1, Y1), flat colour by primitive hash."""
    m = np.array([[2.0 / RW, 0, 0, -1.0], [0, 2.0 / RH, 0, -1.0], [0, 0, 0, 0.5], [0, 0, 0, 1.0]], np.float32)
    verts = np.array([[X0, Y0, 0], [X1, Y0, 0], [X1, Y1, 0], [X0, Y1, 0]], np.float32)
    return scenes.Scene("rect", RW, RH, samples, verts, np.array([0, 1, 2, 0, 2, 3], np.int32), m,
                        program=scenes.PROGRAM_FLAT)


def _expected_resolve(samples):
    """Per pixel: the covered-sample count of each primitive from the sample
    table, and the pixels whose samples all lie clear of every edge."""
    pos = O.sample_positions(samples).astype(np.float64)
    counts = np.zeros((RH, RW, 2), np.int64)
    clean = np.ones((RH, RW), bool)
    # the diagonal is the edge shared by the two triangles
    for py in range(RH):
        for px in range(RW):
            for sx, sy in pos:
                X, Y = px + sx, py + sy
                dist = min(abs(X - X0), abs(X - X1), abs(Y - Y0), abs(Y - Y1))
                side = (X1 - X0) * (Y - Y0) - (Y1 - Y0) * (X - X0)  # > 0: the (X0, Y1) side
                if dist < 1e-3 or abs(side) < 1e-2:
                    clean[py, px] = False
                if X0 < X < X1 and Y0 < Y < Y1:
                    # primitive 0 holds vertex 1 (X1, Y0), where side < 0; primitive 1 vertex 3 (X0, Y1)
                    counts[py, px, 1 if side > 0 else 0] += 1
    return counts, clean


@pytest.mark.parametrize("samples", [4, 8])
def test_msaa_resolve_weights_by_sample_count(samples, plant):
    s = _rect_scene(samples)
    f = _render(s)
    counts, clean = _expected_resolve(samples)
    # the oracle's coverage agrees with the sample table
    ids = f.ids.reshape(RH, RW, samples)
    got_counts = np.stack([(ids == p).sum(-1) for p in (0, 1)], -1)
    assert np.array_equal(got_counts[clean], counts[clean])
    partial = clean & (counts.sum(-1) > 0) & (counts.sum(-1) < samples)
    assert partial.sum() >= 60  # the rectangle's edge pixels
    mixed = clean & (counts[..., 0] > 0) & (counts[..., 1] > 0)
    assert mixed.sum() >= 10  # pixels on the diagonal hold both triangles

    hb = np.array([[(O.flat_hash(p) >> (8 * c)) & 255 for c in range(3)] + [255] for p in (0, 1)], np.int64)
    num = counts[..., 0:1] * hb[0] + counts[..., 1:2] * hb[1]  # S * expected byte value

    def mismatches():
        rgba, _ = f.shade(s)
        got = rgba.astype(np.int64)
        near = num // samples + ((2 * (num % samples)) > samples)  # nearest, halves down
        tie = (2 * (num % samples)) == samples                      # exactly .5: either neighbour
        ok = (got == near) | (tie & (got == near + 1))
        return int((~np.all(ok, -1))[clean].sum())

    assert mismatches() == 0
    plant(4)  # average over covered samples only
    assert mismatches() >= partial.sum() // 2
    plant(5)  # each distinct triangle counted once
    assert mismatches() >= partial.sum() // 2
