"""Point and line polygon modes on the GPU against the oracle (whose pixel
walk, depth and culling rules are pinned by closed forms in
test_oracle_polygons.py): primitive IDs and depth bits per pixel identical,
colour within one RGBA8 step of the FP64 shading of the same records.
"""
import dataclasses

import numpy as np
import pytest

from ifrit_hip import scenes
from oracle import oracle as O

from test_gpu_parity import _random_scene

pytestmark = pytest.mark.gpu


def _compare(r, f, s, label):
    ids, depth = r.visibility()
    bad_id = int((ids != f.ids).sum())
    bad_d = int((depth.view(np.uint32) != f.depth.view(np.uint32)).sum())
    print(f"[{label}] drawn {(f.ids != 0xffffffff).sum()} id/depth mismatches {bad_id}/{bad_d}", flush=True)
    assert bad_id == 0 and bad_d == 0
    ref, flags = f.shade(s)
    diff = np.abs(r.color().astype(int) - ref.astype(int)).max(-1)
    assert ((diff > 1) & (flags == 0)).sum() == 0, (label, diff.max())


@pytest.mark.parametrize("mode", [1, 2])
def test_polygon_modes_random_scenes(renderer_factory, mode):
    r = renderer_factory()
    r.set_polygon_mode(mode)
    for it in range(12):
        W, H, clip, idx = _random_scene(7300 + 100 * mode + it, it % 4)
        f = O.Frame().render_polygons(W, H, clip, idx, mode)
        r.set_target(W, H, 1)
        r.bind_geometry(clip, idx, np.eye(4, dtype=np.float32), clip_passthrough=True)
        s = scenes.Scene("rand", W, H, 1, clip, idx, np.eye(4, dtype=np.float32), program=scenes.PROGRAM_FLAT)
        r.bind_fragment_program(s)
        st = r.draw()
        assert st["fragments"] == f.stats["fragments"]
        assert st["assembledTriangles"] == f.stats["assembled_triangles"]
        _compare(r, f, s, f"mode {mode} it {it} kind {it % 4} {W}x{H}")


@pytest.mark.parametrize("mode", [1, 2])
def test_polygon_modes_config1_textured(renderer_factory, mode):
    s = scenes.config1()
    clip = O.vertex_transform(s.vertices, s.mvp)
    f = O.Frame().render_polygons(s.width, s.height, clip, s.indices, mode)
    r = renderer_factory()
    r.load_scene(s)
    r.set_polygon_mode(mode)
    r.draw()
    _compare(r, f, s, f"config1 mode {mode}")
    # back to filled triangles: the ordinary frame
    r.set_polygon_mode(0)
    r.draw()
    g, _ = O.render_scene(s)
    assert np.array_equal(r.visibility()[0], g.ids)


def test_polygon_modes_keep_a_nearer_stored_depth(renderer_factory):
    r = renderer_factory()
    W, H, clip, idx = _random_scene(7600, 0)
    r.set_target(W, H, 1)
    r.bind_geometry(clip, idx, np.eye(4, dtype=np.float32), clip_passthrough=True)
    s = scenes.Scene("rand", W, H, 1, clip, idx, np.eye(4, dtype=np.float32), program=scenes.PROGRAM_FLAT)
    r.bind_fragment_program(s)
    r.draw()  # filled triangles first
    _, d0 = r.visibility()
    r.set_polygon_mode(1)
    r.draw(clear=False)  # lines over them, kept only where nearer than the triangles
    f = O.Frame().render_polygons(W, H, clip, idx, 1, clear=False, depth_in=d0)
    ids, depth = r.visibility()
    won = f.ids != 0xFFFFFFFF
    assert won.any() and (~won).any()
    assert np.array_equal(ids[won], f.ids[won])
    assert np.array_equal(depth.view(np.uint32), f.depth.view(np.uint32))


def test_polygon_modes_refuse_msaa(renderer_factory):
    from ifrit_hip import IfritError

    r = renderer_factory()
    s = scenes.config1(160, 120)
    s = dataclasses.replace(s, samples=4)
    r.load_scene(s)
    r.set_polygon_mode(2)
    with pytest.raises(IfritError, match="1 sample per pixel"):
        r.draw()
