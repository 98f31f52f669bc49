"""The built-in geometry program (the reference demo's clip-offset geometry
shader) on the GPU: a draw with it bound must match the oracle's geometry
stage drawn by the oracle (primitive IDs and depth bits identical, colour
within one step of the FP64 shading oracle away from flagged texel
choices); with a zero offset it gives the ordinary draw's frame bit for bit;
graph replays, asynchronous draws and unbinding keep working; the checks
report errors.
"""
import dataclasses

import numpy as np
import pytest

from ifrit_hip import scenes
from oracle import oracle as O

pytestmark = pytest.mark.gpu

DEMO_OFFSET = (0.03, 0.0, 0.0, 0.0)  # demo/shader/DefaultDemoShaders.cu:92


def _oracle_frame(s, offset):
    c, recs, eidx = O.geometry_clip_offset(s.vertices, s.indices, s.mvp, offset)
    f = O.Frame().render(s.width, s.height, s.samples, c, eidx)
    ref, flags = f.shade(dataclasses.replace(s, vertices=recs, indices=eidx))
    return f, ref, flags


@pytest.mark.parametrize("samples", [1, 4])
def test_demo_geometry_program(renderer_factory, samples):
    s = dataclasses.replace(scenes.config1(320, 240), samples=samples)
    f, ref, flags = _oracle_frame(s, DEMO_OFFSET)
    r = renderer_factory()
    r.load_scene(s)
    r.bind_clip_offset_geometry(DEMO_OFFSET)
    st = r.draw()
    ids, depth = r.visibility()
    diff = np.abs(r.color().astype(int) - ref.astype(int)).max(-1)
    print(f"[geometry {samples} spp] covered {(f.ids != 0xffffffff).sum()} colour max diff {diff.max()}", flush=True)
    assert np.array_equal(ids, f.ids)
    assert np.array_equal(depth.view(np.uint32), f.depth.view(np.uint32))
    assert ((diff > 1) & (flags == 0)).sum() == 0
    assert st["inputTriangles"] == s.num_triangles
    assert st["fragments"] == f.stats["fragments"]
    # the frame moved: the plain draw differs
    plain, _ = O.render_scene(s)
    assert not np.array_equal(plain.ids, f.ids)


def test_zero_offset_is_the_ordinary_draw(renderer_factory):
    s = scenes.config1(320, 240)
    plain = renderer_factory()
    plain.load_scene(s)
    plain.draw()
    ids0, d0 = plain.visibility()
    c0 = plain.color()
    r = renderer_factory()
    r.load_scene(s)
    r.bind_clip_offset_geometry((0.0, 0.0, 0.0, 0.0))
    r.draw()
    ids1, d1 = r.visibility()
    assert np.array_equal(ids1, ids0)
    assert np.array_equal(d1.view(np.uint32), d0.view(np.uint32))
    assert np.array_equal(r.color(), c0)


def test_clip_passthrough_graph_async_and_unbind(renderer_factory):
    """A clip pass-through vertex program before the geometry program; graph replays and an
    asynchronous draw give the same frame; unbinding returns to the ordinary draw."""
    s = scenes.config1(320, 240)
    clip = O.vertex_transform(s.vertices, s.mvp)
    recs = np.concatenate([clip, s.vertices[:, 3:]], axis=1).astype(np.float32)
    off = (-0.05, 0.02, 0.0, 0.0)
    c, _, eidx = O.geometry_clip_offset(recs, s.indices, None, off, passthrough=True)
    f = O.Frame().render(s.width, s.height, 1, c, eidx)
    r = renderer_factory()
    r.set_target(s.width, s.height, 1)
    r.bind_geometry(recs, s.indices, np.eye(4, dtype=np.float32), clip_passthrough=True)
    r.bind_fragment_program(dataclasses.replace(s, program=scenes.PROGRAM_FLAT, vertices=recs))
    r.bind_clip_offset_geometry(off)
    r.set_graph_mode(True)
    for _ in range(3):
        r.draw()
        ids, depth = r.visibility()
        assert np.array_equal(ids, f.ids)
        assert np.array_equal(depth.view(np.uint32), f.depth.view(np.uint32))
    r.set_graph_mode(False)
    r.draw_async()
    r.synchronize()
    assert np.array_equal(r.visibility()[0], f.ids)
    r.bind_clip_offset_geometry(None)
    r.draw()
    g = O.Frame().render(s.width, s.height, 1, clip, s.indices)
    assert np.array_equal(r.visibility()[0], g.ids)


def test_geometry_errors(renderer_factory):
    from ifrit_hip import IfritError

    s = scenes.config1(160, 120)
    r = renderer_factory()
    r.load_scene(s)
    r.bind_clip_offset_geometry(DEMO_OFFSET)
    bad = s.indices.copy()
    bad[7] = s.vertices.shape[0] + 5
    r.bind_geometry(s.vertices, bad, s.mvp)
    with pytest.raises(IfritError, match="outside the bound vertex buffer"):
        r.draw()
    r.bind_geometry(s.vertices, s.indices, s.mvp)
    r.set_views([dict(width=64, height=64, mvp=s.mvp)])
    with pytest.raises(IfritError, match="single view"):
        r.draw()
    r.set_views([])
    r.draw()
