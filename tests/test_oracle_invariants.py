"""Properties of the oracle's raster stage that hold independently of how it
is organised.

The oracle rasterizes hierarchically (tiles, 4x4 blocks, 4x2 groups, quads,
pixels: TileRasterWorker.cpp:787-1223) and resolves visibility in primitive
order (TileRasterWorker.cpp:1517-1580).  These tests recompute the same
frames flat, from the assembled triangles alone:

* a sample is covered iff its tile passed the binner's trivial-accept test or
  the sample passes the pixel test fma(ey, y, ex * x) < -ez (monotone
  rounding makes every intermediate level agree with that);
* at one sample per pixel the depth expression is chosen by the pixel's
  class (4x2 group fully covered or tile accepted -> SIMD256 form; 2x2 quad
  fully covered -> SIMD128 form; otherwise the scalar DPPS form), with MSAA
  the SIMD256 form at the sample;
* the surviving fragment is the lexicographic minimum of (depth, primitive,
  fan) over the fragments covering the sample;
and check closed forms for screen-aligned rectangles and the culling rules.
FMA results are formed in float64 (the product of two float32 values is
exact there) and rounded once to float32.
"""
import numpy as np
import pytest

from oracle import oracle as O

F32 = np.float32
EMPTY = 0xFFFFFFFF
CLEAR = np.uint32(0x7F7F7F7F).view(np.float32)


def fma32(a, b, c):
    return (np.asarray(a, np.float64) * np.asarray(b, np.float64) + np.asarray(c, np.float64)).astype(F32)


def mul32(a, b):
    return (np.asarray(a, F32) * np.asarray(b, F32)).astype(F32)


def add32(a, b):
    return (np.asarray(a, F32) + np.asarray(b, F32)).astype(F32)


def inside(e, x, y):
    """Pixel / sample test with edge rows e[k] = (ex, ey, ez)."""
    ok = np.ones(np.broadcast(x, y).shape, bool)
    for k in range(3):
        ok &= fma32(F32(e[k][1]), y, mul32(F32(e[k][0]), x)) < -F32(e[k][2])
    return ok


def tile_accept(e, tx, ty):
    """Binner trivial accept of 16-pixel tiles (SIMD128 form)."""
    x0, y0 = (tx * 16).astype(F32), (ty * 16).astype(F32)
    x1, y1 = x0 + F32(16), y0 + F32(16)
    ok = np.ones(np.broadcast(tx, ty).shape, bool)
    for k in range(3):
        ex, ey, ez = (F32(v) for v in e[k])
        xa = x0 if ex < 0 else x1
        ya = y0 if ey < 0 else y1
        ok &= fma32(ex, xa, fma32(ey, ya, ez)) < 0
    return ok


def bary(F, x, y):
    return [fma32(F32(f[0]), x, fma32(F32(f[1]), y, F32(f[2]))) for f in F]


def depth256(F, x, y):
    b = bary(F, x, y)
    return fma32(b[0], F32(F[0][3]), fma32(b[1], F32(F[1][3]), mul32(b[2], F32(F[2][3]))))


def depth128(F, x, y):
    b = bary(F, x, y)
    return fma32(b[2], F32(F[2][3]), fma32(b[0], F32(F[0][3]), mul32(b[1], F32(F[1][3]))))


def depth_scalar(F, x, y):
    # DPPS sums ((p0 + p1) + (p2 + p3)) of the masked products
    b = [add32(add32(mul32(x, F32(f[0])), mul32(y, F32(f[1]))), add32(mul32(F32(1), F32(f[2])), mul32(F32(0), F32(f[3]))))
         for f in F]
    return add32(add32(mul32(F32(F[0][3]), b[0]), mul32(F32(F[1][3]), b[1])),
                 add32(mul32(F32(F[2][3]), b[2]), F32(0)))


def flat_frame(W, H, S, tris):
    """Visibility recomputed sample by sample from the assembled triangles."""
    pos = O.sample_positions(S)
    py, px = np.mgrid[0:H, 0:W]
    best_d = np.full((H, W, S), CLEAR, F32)
    best_k = np.full((H, W, S), -1, np.int64)
    frags = 0
    covered_samples = 0
    for t in tris:
        e = [t["e1"], t["e2"], t["e3"]]
        F = [t["f1"], t["f2"], t["f3"]]
        key = int(t["prim"]) * 4 + int(t["fan"])
        ta = tile_accept(e, px // 16, py // 16)
        fx, fy = px.astype(F32), py.astype(F32)
        if S == 1:
            cov = ta | inside(e, fx, fy)
            # 4x2 groups and 2x2 quads of the pixel
            gx0, gy0 = (px // 4) * 4, (py // 2) * 2
            g = np.ones_like(cov)
            for dy in range(2):
                for dx in range(4):
                    g &= inside(e, (gx0 + dx).astype(F32), (gy0 + dy).astype(F32))
            qx0 = (px // 2) * 2
            q = np.ones_like(cov)
            for dy in range(2):
                for dx in range(2):
                    q &= inside(e, (qx0 + dx).astype(F32), (gy0 + dy).astype(F32))
            c1 = ta | g
            d = np.where(c1, depth256(F, fx, fy), np.where(q, depth128(F, fx, fy), depth_scalar(F, fx, fy)))
            covs = cov[..., None]
            ds = d[..., None]
        else:
            sx = (fx[..., None] + pos[:, 0]).astype(F32)
            sy = (fy[..., None] + pos[:, 1]).astype(F32)
            covs = ta[..., None] | inside(e, sx, sy)
            ds = depth256(F, sx, sy)
        frags += int(covs.any(-1).sum())
        covered_samples += int(covs.sum())
        dz = np.where(ds == 0, F32(0), ds)  # -0 and +0 compare equal
        bz = np.where(best_d == 0, F32(0), best_d)
        better = covs & ~np.isnan(ds) & ((dz < bz) | ((dz == bz) & (best_k >= 0) & (key < best_k)))
        best_d = np.where(better, ds, best_d)
        best_k = np.where(better, key, best_k)
    ids = np.where(best_k >= 0, best_k // 4, EMPTY).astype(np.uint32)
    return ids.reshape(-1), best_d.reshape(-1), frags, covered_samples


def _scene(seed, kind, W, H, n):
    rng = np.random.default_rng(seed)
    if kind == 0:  # small triangles in front of the camera
        ctr = rng.uniform(-1, 1, (n, 1, 2))
        xy = ctr + rng.uniform(-0.25, 0.25, (n, 3, 2))
        w = rng.uniform(0.5, 2, (n, 3, 1))
        z = rng.uniform(0, 1, (n, 3, 1))
        c = np.concatenate([xy * w, z * w, w], -1).reshape(-1, 4)
    elif kind == 1:  # crossing w = 0
        c = np.concatenate([rng.uniform(-2, 2, (3 * n, 2)), rng.uniform(-1, 2, (3 * n, 1)),
                            rng.uniform(-1, 2, (3 * n, 1))], 1)
    else:  # large overlapping triangles (tile accepts)
        c = np.concatenate([rng.uniform(-3, 3, (3 * n, 2)), rng.uniform(0, 1, (3 * n, 1)), np.ones((3 * n, 1))], 1)
    return c.astype(F32), np.arange(3 * n, dtype=np.int32)


@pytest.mark.parametrize("S", [1, 4, 8])
@pytest.mark.parametrize("kind", [0, 1, 2])
def test_hierarchy_equals_flat_recomputation(S, kind):
    W, H = 53, 37
    clip, idx = _scene(100 * S + kind, kind, W, H, 24)
    f = O.Frame().render(W, H, S, clip, idx, threads=2)
    tris = f.tris()
    assert len(tris) > 0
    ids, depth, frags, covs = flat_frame(W, H, S, tris)
    assert np.array_equal(f.ids, ids), f"{(f.ids != ids).sum()} id mismatches"
    assert np.array_equal(f.depth.view(np.uint32), depth.view(np.uint32)), "depth mismatches"
    assert f.stats["fragments"] == frags
    assert f.stats["covered_samples"] == covs


def test_tile_accept_path_is_exercised():
    clip, idx = _scene(2, 2, 64, 64, 24)
    tris = O.Frame().render(64, 64, 1, clip, idx, threads=2).tris()
    ty, tx = np.mgrid[0:4, 0:4]
    assert any(tile_accept([t["e1"], t["e2"], t["e3"]], tx, ty).any() for t in tris)


@pytest.mark.parametrize("S", [1, 8])
def test_each_sample_keeps_lexicographic_minimum_of_single_triangle_renders(S):
    """Order independence: rendering each triangle alone and keeping the
    (depth, primitive) minimum per sample gives the full frame."""
    W, H = 40, 29
    clip, idx = _scene(7 + S, 0, W, H, 16)
    full = O.Frame().render(W, H, S, clip, idx, threads=2)
    best_d = np.full(W * H * S, CLEAR, F32)
    best_id = np.full(W * H * S, EMPTY, np.uint32)
    frags = 0
    for p in range(idx.size // 3):
        only = idx.copy()
        keep = np.zeros(idx.size // 3, bool)
        keep[p] = True
        only[~np.repeat(keep, 3)] = idx[3 * p]  # other triangles degenerate (culled, same primitive numbering)
        one = O.Frame().render(W, H, S, clip, only, threads=2)
        frags += one.stats["fragments"]
        hit = one.ids != EMPTY
        assert np.all(one.ids[hit] == p)
        better = hit & ((one.depth < best_d) | ((one.depth == best_d) & (p < best_id)))
        best_d = np.where(better, one.depth, best_d)
        best_id = np.where(better, np.uint32(p), best_id)
    assert np.array_equal(full.ids, best_id)
    assert np.array_equal(full.depth.view(np.uint32), best_d.view(np.uint32))
    assert full.stats["fragments"] == frags


def _rect_scene(x0, y0, x1, y1, z, w):
    """Two triangles covering an axis-aligned NDC rectangle at constant depth,
    counter-clockwise as seen by the camera."""
    ndc = np.array([[x0, y0], [x1, y0], [x1, y1], [x0, y1]], np.float64)
    c = np.concatenate([ndc * w, np.full((4, 1), z * w), np.full((4, 1), w)], 1).astype(F32)
    return c, np.array([0, 1, 2, 0, 2, 3], np.int32)


@pytest.mark.parametrize("S", [1, 4, 8])
def test_screen_rectangle_closed_form(S):
    W, H = 64, 48
    x0, y0, x1, y1, z = -0.61, -0.43, 0.37, 0.71, 0.3
    clip, idx = _rect_scene(x0, y0, x1, y1, z, 2.0)
    f = O.Frame().render(W, H, S, clip, idx, threads=2)
    ids = f.ids.reshape(H, W, S)
    depth = f.depth.reshape(H, W, S)
    if (ids == EMPTY).all():  # wrong winding for this convention: flip it
        idx = idx[[0, 2, 1, 3, 5, 4]]
        f = O.Frame().render(W, H, S, clip, idx, threads=2)
        ids = f.ids.reshape(H, W, S)
        depth = f.depth.reshape(H, W, S)
    pos = O.sample_positions(S)
    py, px = np.mgrid[0:H, 0:W]
    sx = px[..., None] + pos[:, 0].astype(np.float64)
    sy = py[..., None] + pos[:, 1].astype(np.float64)
    # NDC -> pixels: x = (ndc + 1) * W / 2, row 0 at NDC y = -1
    X0, X1 = (x0 + 1) * W / 2, (x1 + 1) * W / 2
    Y0, Y1 = (y0 + 1) * H / 2, (y1 + 1) * H / 2
    m = 1e-3
    strictly_in = (sx > X0 + m) & (sx < X1 - m) & (sy > Y0 + m) & (sy < Y1 - m)
    strictly_out = (sx < X0 - m) | (sx > X1 + m) | (sy < Y0 - m) | (sy > Y1 + m)
    # the shared diagonal: a sample within 1e-3 of it may go either way
    diag = np.abs((sx - X0) * (Y1 - Y0) - (sy - Y0) * (X1 - X0)) / np.hypot(X1 - X0, Y1 - Y0) < m
    assert np.all(ids[strictly_in & ~diag] != EMPTY)
    assert np.all(ids[strictly_out] == EMPTY)
    hit = ids != EMPTY
    assert set(np.unique(ids[hit])) <= {0, 1}
    assert np.allclose(depth[hit], z, rtol=2e-6, atol=0)


def test_back_facing_and_all_behind_triangles_produce_nothing():
    W, H = 32, 32
    clip, idx = _rect_scene(-0.5, -0.5, 0.5, 0.5, 0.5, 1.0)
    a = O.Frame().render(W, H, 1, clip, idx, threads=1)
    b = O.Frame().render(W, H, 1, clip, idx[[0, 2, 1, 3, 5, 4]], threads=1)
    assert ((a.ids != EMPTY).sum() == 0) != ((b.ids != EMPTY).sum() == 0), "exactly one winding is front-facing"
    behind = clip.copy()
    behind[:, 3] = -1.0
    for order in (idx, idx[[0, 2, 1, 3, 5, 4]]):
        g = O.Frame().render(W, H, 1, behind, order, threads=1)
        assert (g.ids == EMPTY).all() and len(g.tris()) == 0


def test_clipped_triangle_fans_share_the_primitive():
    W, H = 48, 48
    clip = np.array([[-0.5, -0.5, 0.2, 1.0], [0.6, -0.4, 0.5, 1.2], [0.1, 0.7, -0.8, -0.6]], F32)
    for idx in (np.array([0, 1, 2], np.int32), np.array([0, 2, 1], np.int32)):
        f = O.Frame().render(W, H, 1, clip, idx, threads=1)
        t = f.tris()
        if len(t):
            assert np.all(t["prim"] == 0)
            assert sorted(t["fan"]) == list(range(len(t)))
            assert f.stats["clipped_inputs"] == 1
            return
    pytest.fail("neither winding produced a clipped triangle")


@pytest.mark.parametrize("S", [1, 8])
def test_equal_depth_keeps_lowest_primitive(S):
    """Coincident triangles tie on depth everywhere; the reference shades in
    primitive order and keeps a fragment only when !(depth >= stored)
    (TileRasterWorker.cpp:2286-2290), so the first primitive survives."""
    W, H = 40, 30
    rng = np.random.default_rng(11)
    tri = np.concatenate([rng.uniform(-0.9, 0.9, (3, 2)), rng.uniform(0.1, 0.9, (3, 1)), np.ones((3, 1))], 1)
    clip = np.tile(tri, (4, 1)).astype(F32)
    idx = np.arange(12, dtype=np.int32)
    if (O.Frame().render(W, H, S, clip, idx, threads=1).ids == EMPTY).all():
        idx = idx.reshape(4, 3)[:, [0, 2, 1]].reshape(-1).copy()
    f = O.Frame().render(W, H, S, clip, idx, threads=1)
    hit = f.ids != EMPTY
    assert hit.any() and np.all(f.ids[hit] == 0)
    ids, depth, _, _ = flat_frame(W, H, S, f.tris())
    assert np.array_equal(f.ids, ids)
    # a later primitive cannot displace an equal-depth earlier one
    g = O.Frame().render(W, H, S, clip, np.concatenate([idx[3:], idx[:3]]).astype(np.int32), threads=1)
    assert np.all(g.ids[g.ids != EMPTY] == 0)


# --- the layered-GGX material's BRDF (oracle_shade.cpp ggx_light) -------------------------------

def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


@pytest.mark.parametrize("alpha", [0.05, 0.3, 1.0])
def test_ggx_normal_incidence_closed_form(alpha):
    """n = v = l: D = 1/(pi alpha^2), G = 1, F = F0 = 0.04, so the term is
    albedo * 0.96 kd / pi + ks * 0.04 / (4 pi alpha^2)."""
    z = np.array([0.0, 0.0, 1.0])
    alb = np.array([0.2, 0.5, 0.9])
    got = O.ggx_light(z, z, z, alb, alpha, 0.7, 0.4)
    want = alb * 0.96 * 0.7 / np.pi + 0.4 * 0.04 / (4 * np.pi * alpha * alpha)
    assert np.allclose(got, want, rtol=1e-12)


def test_ggx_is_zero_below_the_horizon():
    n = np.array([0.0, 0.0, 1.0])
    assert not O.ggx_light(n, _unit([0.3, 0, 1]), _unit([0.2, 0.1, -0.5]), [1, 1, 1], 0.4, 1, 1).any()
    assert not O.ggx_light(n, _unit([0.3, 0, -1]), _unit([0.2, 0.1, 0.5]), [1, 1, 1], 0.4, 1, 1).any()


def test_ggx_brdf_is_reciprocal():
    """f(v, l) = f(l, v): the term divided by n.l must not change when view and light swap (a
    transposed operand, e.g. n.v where n.l belongs, breaks this)."""
    rng = np.random.default_rng(3)
    n = np.array([0.0, 0.0, 1.0])
    for _ in range(200):
        v = _unit(rng.normal(size=3) + [0, 0, 1.5])
        l = _unit(rng.normal(size=3) + [0, 0, 1.5])
        if v[2] <= 0.05 or l[2] <= 0.05:
            continue
        a = rng.uniform(0.05, 1.0)
        f1 = O.ggx_light(n, v, l, [0.3, 0.6, 0.8], a, 0.8, 0.5) / l[2]
        f2 = O.ggx_light(n, l, v, [0.3, 0.6, 0.8], a, 0.8, 0.5) / v[2]
        assert np.allclose(f1, f2, rtol=1e-10)


@pytest.mark.parametrize("alpha", [0.2, 0.5, 1.0])
def test_ggx_distribution_is_normalised(alpha):
    """Retro-reflection (v = l, so h = v and F = F0) isolates D(n.h) * G1(n.v)^2; with the
    Smith-Schlick G1 taken out, the recovered D must satisfy the textbook normalisation
    integral over the hemisphere of D(h) (n.h) d(omega_h) = 1 (Walter et al. 2007)."""
    n = np.array([0.0, 0.0, 1.0])
    k = alpha / 2
    c = np.linspace(1e-4, 1.0, 20001)
    d = np.empty_like(c)
    for i, ci in enumerate(c):
        v = np.array([np.sqrt(1 - ci * ci), 0.0, ci])
        spec = O.ggx_light(n, v, v, [0, 0, 0], alpha, 0.0, 1.0)[0]  # = D G F / (4 c c) * c
        g1 = ci / (ci * (1 - k) + k)
        d[i] = spec * 4 * ci / (0.04 * g1 * g1)
    integral = 2 * np.pi * np.trapezoid(d * c, c)
    assert abs(integral - 1.0) < 2e-3
