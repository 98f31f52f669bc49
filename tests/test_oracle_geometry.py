"""Pins of the oracle's geometry stage (the demo's clip-offset geometry
shader): with a zero offset the emitted list draws the frame of the plain
draw (unsharing keeps triangles and primitive ids); an offset of 2k / W in
clip x on w = 1 vertices whose coordinates and sums are exact moves the
frame k pixels to the right, which fixes the offset's sign and component.
Pixels on a triangle's edge may resolve differently after the move (the edge
functions' rounding depends on absolute coordinates), so the shift is
checked on the pixels whose 3 x 3 neighbourhood lies in one triangle.
"""
import numpy as np

from ifrit_hip import scenes
from oracle import oracle as O


def test_zero_offset_is_the_plain_draw():
    s = scenes.config1(160, 120)
    f0, clip = O.render_scene(s)
    c, recs, idx = O.geometry_clip_offset(s.vertices, s.indices, s.mvp, (0, 0, 0, 0))
    assert np.array_equal(c, clip[s.indices])
    assert np.array_equal(recs, s.vertices[s.indices])
    f1 = O.Frame().render(s.width, s.height, 1, c, idx)
    assert np.array_equal(f1.ids, f0.ids)
    assert np.array_equal(f1.depth.view(np.uint32), f0.depth.view(np.uint32))


def test_offset_moves_the_frame():
    W = H = 128
    k = 16
    rng = np.random.default_rng(7)
    n = 12
    # w = 1 clip positions on a 1/64 grid inside [-0.75, 0.75]: every sum below is exact
    xy = rng.integers(-48, 49, size=(3 * n, 2)).astype(np.float32) / 64
    z = rng.integers(4, 28, size=(3 * n, 1)).astype(np.float32) / 32
    clip = np.concatenate([xy, z, np.ones((3 * n, 1), np.float32)], axis=1)
    idx = np.arange(3 * n, dtype=np.int32)
    plain = O.Frame().render(W, H, 1, clip, idx, cull_mode=0)
    c, _, eidx = O.geometry_clip_offset(clip, idx, np.eye(4, dtype=np.float32), (2 * k / W, 0, 0, 0),
                                        passthrough=True)
    moved = O.Frame().render(W, H, 1, c, eidx, cull_mode=0)
    a = plain.ids.reshape(H, W)
    b = moved.ids.reshape(H, W)
    inner = np.zeros((H, W), bool)
    inner[1:-1, 1:-1] = True
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            inner[1:-1, 1:-1] &= a[1 + dy:H - 1 + dy, 1 + dx:W - 1 + dx] == a[1:-1, 1:-1]
    inner &= a != 0xFFFFFFFF
    inner[:, W - k:] = False
    ys, xs = np.nonzero(inner)
    assert ys.size > W * H // 8, ys.size
    assert np.array_equal(b[ys, xs + k], a[ys, xs])
    # depth: the plain frame's values where the moved triangle's samples land (the planes move with it)
    da = plain.depth.reshape(H, W)
    db = moved.depth.reshape(H, W)
    assert np.abs(db[ys, xs + k] - da[ys, xs]).max() <= 1e-6
    # a wrong sign or component gives another frame
    assert (b[ys, xs] != a[ys, xs]).sum() > ys.size // 4
