"""Pins of the oracle's point and line polygon modes (the CUDA spec's
PointPipeline and LinePipeline, TileRasterCoreInvocationCuda.cu:2780-3335,
which cannot be compiled here) by what the definitions fix:

* points: pixel (int(x (W - 1)), int(y (H - 1))) of the [0, 1]-mapped NDC
  position, depth z / w, points with w < 1e-8 or outside [0, 1] dropped, the
  nearest point per pixel kept (equal depths: the lowest index), and kept only
  below the stored depth;
* lines: the spec's Bresenham walk written out as pixel tables for slopes 0,
  1/2, 1/3, 1 and a steep one (its strict d > 0.5 step rule); depth linear in
  the pixel coordinate along the major axis with the ends scaled by W and H
  (not W - 1); an end behind w = 0 cut at w = 1; the spec's asymmetric bounds
  test (a far end beyond x = 1 is kept, a near one dropped); edge ids 3 t + e;
* shading records: the planes of a line give one-hot weights at its ends and
  perspective-correct ones in between.
"""
import numpy as np
import pytest

from oracle import oracle as O

W, H = 101, 81  # W - 1 = 100, H - 1 = 80


def _ndc(px, py, size=(W, H)):
    """NDC (x, y) whose [0, 1] mapping lands a quarter pixel into pixel (px, py)."""
    return 2 * (px + 0.25) / (size[0] - 1) - 1, 2 * (py + 0.25) / (size[1] - 1) - 1


def _pts(points):
    clip = np.array([[*p, 1.0] for p in points], np.float32)
    return clip


def test_points_project_cull_and_resolve():
    x0, y0 = _ndc(10, 20)
    x1, y1 = _ndc(50, 40)
    clip = np.array([[x0, y0, 0.5, 1], [x1 * 2, y1 * 2, 0.8, 2],        # (10, 20) z 0.5; (50, 40) z 0.4
                     [x0, y0, 0.3, 1], [x0, y0, 0.3, 1],               # nearer points on (10, 20), a tie
                     [1.5, 0.0, 0.1, 1], [0.0, 0.0, 0.1, 1e-9],        # outside; behind w = 1e-8
                     [1.0, -1.0, 0.2, 1],                              # x = 1, y = 0: pixel (100, 0)
                     [2.0, 0.0, 0.1, 1]], np.float32)                  # outside
    idx = np.array([0, 1, 2, 3, 4, 5, 6, 7, 7, 0], np.int32)  # the trailing index forms no triangle: ignored
    f = O.Frame().render_polygons(W, H, clip, idx, mode=2)
    ids = f.ids.reshape(H, W)
    depth = f.depth.reshape(H, W)
    assert ids[20, 10] == 2 and depth[20, 10] == np.float32(0.3)  # nearest, then the lowest index
    assert ids[40, 50] == 1 and depth[40, 50] == np.float32(0.4)
    assert ids[0, 100] == 6
    assert (ids != 0xFFFFFFFF).sum() == 3
    assert f.stats["assembled_triangles"] == 5  # points 4 and 5 are dropped before the raster
    # kept only below the stored depth
    g = O.Frame().render_polygons(W, H, clip, idx, mode=2, clear=False, depth_in=np.full(W * H, 0.35, np.float32))
    g_ids = g.ids.reshape(H, W)
    assert g_ids[20, 10] == 2 and g_ids[40, 50] == 0xFFFFFFFF and g_ids[0, 100] == 6


# (end pixels, the spec's Bresenham pixels from the first end)
BRESENHAM = [
    (((10, 30), (16, 30)), [(10, 30), (11, 30), (12, 30), (13, 30), (14, 30), (15, 30), (16, 30)]),
    (((10, 30), (14, 32)), [(10, 30), (11, 30), (12, 31), (13, 31), (14, 32)]),
    (((10, 30), (16, 32)), [(10, 30), (11, 30), (12, 31), (13, 31), (14, 31), (15, 32), (16, 32)]),
    (((10, 30), (14, 34)), [(10, 30), (11, 31), (12, 32), (13, 33), (14, 34)]),
    (((40, 10), (41, 14)), [(40, 10), (40, 11), (40, 12), (41, 13), (41, 14)]),  # d: .25 .5 .75 > .5
    (((16, 32), (10, 30)), [(10, 30), (11, 30), (12, 31), (13, 31), (14, 31), (15, 32), (16, 32)]),
]


def _edge_frame(e0, e1, z=(0.1, 0.1)):
    """A triangle whose edge 0 runs e0 -> e1 and whose third vertex lies outside the frame at
    x, y < 0, so that the spec's bounds test drops the other two edges (edge 1 ends there, edge 2
    starts there) and edge 0 is drawn alone."""
    a, b = _ndc(*e0), _ndc(*e1)
    clip = np.array([[a[0], a[1], z[0], 1], [b[0], b[1], z[1], 1], [-1.5, -1.5, 0.5, 1]], np.float32)
    return O.Frame().render_polygons(W, H, clip, np.array([0, 1, 2], np.int32), mode=1), clip


@pytest.mark.parametrize("case", range(len(BRESENHAM)))
def test_line_pixels_follow_the_spec_walk(case):
    (e0, e1), want = BRESENHAM[case]
    f, _ = _edge_frame(e0, e1)
    ids = f.ids.reshape(H, W)
    got = sorted((int(x), int(y)) for y, x in zip(*np.nonzero(ids == 0)))
    assert got == sorted(want)


def test_line_depth_is_linear_along_the_major_axis():
    e0, e1 = (10, 30), (70, 45)
    f, clip = _edge_frame(e0, e1, z=(0.2, 0.6))
    ids, depth = f.ids.reshape(H, W), f.depth.reshape(H, W)
    xs = (clip[:2, 0] / clip[:2, 3] * 0.5 + 0.5) * W  # the ends scaled by W (getPercent)
    ys, xs_ = np.nonzero(ids == 0)
    assert xs_.size > 50
    p = (xs_ - xs[0]) / (xs[1] - xs[0])
    want = 0.2 * (1 - p) + 0.6 * p
    assert np.allclose(depth[ys, xs_], want, atol=2e-6)
    # with (W - 1) in place of W the depths would be off by far more than that
    p2 = (xs_ - xs[0] * (W - 1) / W) / ((xs[1] - xs[0]) * (W - 1) / W)
    assert np.abs(0.2 * (1 - p2) + 0.6 * p2 - want).max() > 1e-4


def test_line_ids_culling_and_clipping():
    a, b, c = _ndc(10, 10), _ndc(30, 10), _ndc(20, 30)
    clip = np.array([[a[0], a[1], 0.5, 1], [b[0], b[1], 0.5, 1], [c[0], c[1], 0.5, 1],
                     [0.0, 0.0, 0.5, 1], [1.8, 0.0, 0.5, 1], [0.0, 0.6, 0.5, 1]], np.float32)
    # triangle 1: edge 3 = (0,0)->(1.8,0) far end beyond x = 1: kept; edge 4 = (1.8,0)->(0,0.6): first end
    # beyond x = 1: dropped; edge 5 = (0,0.6)->(0,0): kept
    f = O.Frame().render_polygons(W, H, clip, np.array([0, 1, 2, 3, 4, 5], np.int32), mode=1)
    seen = set(int(i) for i in np.unique(f.ids) if i != 0xFFFFFFFF)
    assert seen == {0, 1, 2, 3, 5}
    assert f.stats["assembled_triangles"] == 5
    # an end behind w = 0 is cut at w = 1: ends (0.6, 0.6, 0.6, 3) and (-1, -1, -1, -1) meet w = 1 at
    # t = 1/2, the clip-space midpoint (-0.2, -0.2, -0.2, 1): NDC (-0.2, -0.2), pixel (40, 32)
    clip = np.array([[0.6, 0.6, 0.6, 3], [-1.0, -1.0, -1.0, -1.0], [0.6, 0.6, 0.6, 3]], np.float32)
    f = O.Frame().render_polygons(W, H, clip, np.array([0, 1, 2], np.int32), mode=1)
    ids, depth = f.ids.reshape(H, W), f.depth.reshape(H, W)
    # pixel 40 lies 0.4 before the cut end's scaled x (0.4 W = 40.4): getPercent extrapolates there
    p = (40 - 0.4 * W) / (0.6 * W - 0.4 * W)
    assert ids[32, 40] in (0, 1) and abs(depth[32, 40] - (-0.2 * (1 - p) + 0.2 * p)) < 1e-5
    # the rest of the line lies between the cut end and the visible end (0.2, 0.2): x in [40, 60]
    ys, xs = np.nonzero(ids != 0xFFFFFFFF)
    assert xs.min() == 40 and xs.max() <= 60 and np.all(depth[ys, xs] < 0.21)


def test_line_records_weight_their_ends():
    a, b = _ndc(10, 30), _ndc(70, 45)
    clip = np.array([[a[0] * 2, a[1] * 2, 0.2, 2], [b[0] * 4, b[1] * 4, 0.4, 4], [0.5, 0.9, 0.9, 1]], np.float32)
    f = O.Frame().render_polygons(W, H, clip, np.array([0, 1, 2], np.int32), mode=1)
    t = f.tris()
    rec = t[(t["prim"] == 0) & (t["fan"] == 0)][0]
    ax, bx = (a[0] * 0.5 + 0.5) * W, (b[0] * 0.5 + 0.5) * W

    def weights(x):
        b1 = rec["f1"][0] * x + rec["f1"][1] * 0 + rec["f1"][2]
        b2 = rec["f2"][0] * x + rec["f2"][1] * 0 + rec["f2"][2]
        B = np.array([b1, b2]) / (b1 + b2)
        return B[0] * rec["bx"][0] + B[1] * rec["bx"][1], B[0] * rec["by"][0] + B[1] * rec["by"][1]

    assert np.allclose(weights(ax), (1, 0), atol=1e-6) and np.allclose(weights(bx), (0, 1), atol=1e-6)
    # half way in screen space the perspective-correct weight of the far end (w 4) is (1/4) / (1/2 + 1/4)
    w0, w1 = weights((ax + bx) / 2)
    assert abs(w1 - (0.5 / 4) / (0.5 / 2 + 0.5 / 4)) < 1e-5 and abs(w0 + w1 - 1) < 1e-6
