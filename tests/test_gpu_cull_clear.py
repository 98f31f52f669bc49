"""Cull modes (and clear values) on the GPU against the oracle.

Every cull mode on random clip-space scenes (triangles crossing w = 0
included) at 1 and 4 samples: primitive IDs and depth bits per sample equal
the oracle's; with attributes (config 1, Blinn-Phong, cull mode front and
none) the colour is within one RGBA8 step of the FP64 oracle, which shades
exchanged triangles with their own vertices.
"""
import dataclasses

import numpy as np
import pytest

from ifrit_hip import scenes
from oracle import oracle as O
from test_gpu_parity import _random_scene

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("samples", [1, 4])
@pytest.mark.parametrize("cull", [0, 1, 2, 3])
def test_cull_modes_random_scenes(renderer_factory, samples, cull):
    r = renderer_factory()
    r.set_cull_mode(cull)
    for it in range(12):
        W, H, clip, idx = _random_scene(5000 + 100 * cull + 10 * samples + it, it % 4)
        f = O.Frame().render(W, H, samples, clip, idx, cull_mode=cull)
        r.set_target(W, H, samples)
        r.bind_geometry(clip, idx, np.eye(4, dtype=np.float32), clip_passthrough=True)
        s = scenes.Scene("rand", W, H, samples, clip, idx, np.eye(4, dtype=np.float32), program=scenes.PROGRAM_FLAT)
        r.bind_fragment_program(s)
        st = r.draw()
        ids, depth = r.visibility()
        bad_id = int((ids != f.ids).sum())
        bad_d = int((depth.view(np.uint32) != f.depth.view(np.uint32)).sum())
        print(f"[cull {cull} s={samples} it={it} kind={it % 4} {W}x{H}] covered {(f.ids != 0xffffffff).sum()} "
              f"id/depth mismatches {bad_id}/{bad_d}", flush=True)
        assert bad_id == 0 and bad_d == 0
        assert st["fragments"] == f.stats["fragments"]
        assert st["assembledTriangles"] == f.stats["assembled_triangles"]
        ref, _ = f.shade(s)
        assert np.abs(r.color().astype(int) - ref.astype(int)).max() <= 1


@pytest.mark.parametrize("cull", [0, 1])
def test_cull_modes_shade_exchanged_triangles(renderer_factory, cull):
    s = scenes.config1(320, 240)
    f, _ = O.render_scene(s, cull_mode=cull)
    r = renderer_factory()
    r.set_cull_mode(cull)
    r.load_scene(s)
    r.draw()
    ids, depth = r.visibility()
    assert np.array_equal(ids, f.ids) and np.array_equal(depth.view(np.uint32), f.depth.view(np.uint32))
    ref, flags = f.shade(s)
    diff = np.abs(r.color().astype(int) - ref.astype(int)).max(-1)
    print(f"[cull {cull} config1] covered {(f.ids != 0xffffffff).sum()} colour max diff {diff.max()}", flush=True)
    assert ((diff > 1) & (flags == 0)).sum() == 0


def _flat_draw(r, W, H, S, clip, idx):
    r.set_target(W, H, S)
    r.bind_geometry(clip, idx, np.eye(4, dtype=np.float32), clip_passthrough=True)
    s = scenes.Scene("rand", W, H, S, clip, idx, np.eye(4, dtype=np.float32), program=scenes.PROGRAM_FLAT)
    r.bind_fragment_program(s)
    r.draw()
    ids, depth = r.visibility()
    return s, ids, depth, r.color()


@pytest.mark.parametrize("samples", [1, 4])
def test_clear_values(renderer_factory, samples):
    """setClearValues: visibility bit-exact against the oracle with that clear depth (a fragment
    must be nearer than it), colour within one step of the FP64 oracle whose empty samples
    resolve to the clear colour."""
    r = renderer_factory()
    C = (0.2, 0.5, 0.9, 0.75)
    for it in range(8):
        W, H, clip, idx = _random_scene(7100 + 10 * samples + it, it % 4)
        D = (0.6, 0.3, 1.0, 3.3961514e38)[it % 4]
        r.set_clear_values(C, D)
        s, ids, depth, rgba = _flat_draw(r, W, H, samples, clip, idx)
        f = O.Frame().render(W, H, samples, clip, idx, clear_depth=D)
        bad_id = int((ids != f.ids).sum())
        bad_d = int((depth.view(np.uint32) != f.depth.view(np.uint32)).sum())
        ref, _ = f.shade(s, clear_color=C)
        diff = np.abs(rgba.astype(int) - ref.astype(int)).max()
        print(f"[clear values s={samples} it={it} D={D}] id/depth mismatches {bad_id}/{bad_d} colour max diff {diff}",
              flush=True)
        assert bad_id == 0 and bad_d == 0 and diff <= 1
    r.set_clear_values()


def _scissor_mask(W, H, rects):
    y, x = np.mgrid[0:H, 0:W].astype(np.float64)
    m = np.zeros((H, W), bool)
    for rx, ry, rw, rh in rects:
        m |= (x >= rx) & (x <= rx + rw) & (y >= ry) & (y <= ry + rh)
    return m


@pytest.mark.parametrize("samples", [1, 4])
def test_scissor_leaves_outside_pixels(renderer_factory, samples):
    """A draw with the scissor test on writes only pixels inside one of the rectangles
    (inclusive bounds); outside them the previous frame's IDs, depth and colour stay."""
    r = renderer_factory()
    for it in range(6):
        W, H, clip, idx = _random_scene(8200 + 10 * samples + it, (0, 2, 3)[it % 3])  # no clip fans
        rects = [(W * 0.1 + 0.5, H * 0.2 + 0.25, W * 0.4, H * 0.3), (W * 0.7, 1.0, 3.0, H * 0.8)]
        r.set_cull_mode(2)
        r.set_scissors([], enable=False)
        _, ids_a, depth_a, rgba_a = _flat_draw(r, W, H, samples, clip, idx)
        r.set_cull_mode(1)  # a different frame: the back faces
        r.set_scissors(rects, enable=True)
        s, ids_b, depth_b, rgba_b = _flat_draw(r, W, H, samples, clip, idx)
        f = O.Frame().render(W, H, samples, clip, idx, cull_mode=1)
        ref, _ = f.shade(s)
        m = _scissor_mask(W, H, rects)
        ms = np.repeat(m.reshape(-1), samples)
        want_ids = np.where(ms, f.ids, ids_a)
        want_d = np.where(ms, f.depth, depth_a)
        bad_id = int((ids_b != want_ids).sum())
        bad_d = int((depth_b.view(np.uint32) != want_d.view(np.uint32)).sum())
        dc = np.abs(rgba_b.astype(int) - np.where(m[..., None], ref, rgba_a).astype(int)).max()
        print(f"[scissor s={samples} it={it} {W}x{H}] inside {m.sum()} id/depth mismatches {bad_id}/{bad_d} "
              f"colour max diff {dc}", flush=True)
        assert bad_id == 0 and bad_d == 0 and dc <= 1
        assert np.array_equal(rgba_b[~m], rgba_a[~m])
    r.set_scissors([], enable=False)
    r.set_cull_mode(2)
