"""DrawMeshTasks with the built-in meshlet mesh and task programs (the
reference's meshlet demo) on the GPU: the meshlets of config 1, emitted by
the mesh work groups and drawn, must give the frame the ordinary draw gives
for the same triangle list (primitive IDs, depth bits and colour identical),
which the oracle also draws; a draw from a later first work group draws the
remaining meshlets' triangles, numbered from 0; the buffer and program checks
report errors.
"""
import dataclasses

import numpy as np
import pytest

from ifrit_hip import scenes
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _meshlet_scene():
    s = scenes.config1()
    recs, loc, vo, io, glob = scenes.build_meshlets(s.vertices, s.indices)
    flat = dataclasses.replace(s, vertices=recs, indices=glob)  # the same triangles as an ordinary draw
    return s, flat, recs, loc, vo, io


def test_meshlets_equal_the_ordinary_draw(renderer_factory):
    s, flat, recs, loc, vo, io = _meshlet_scene()
    plain = renderer_factory()
    plain.load_scene(flat)
    plain.draw()
    ids0, d0 = plain.visibility()
    c0 = plain.color()
    f, _ = O.render_scene(flat)
    assert np.array_equal(ids0, f.ids)

    r = renderer_factory()
    r.load_scene(flat)
    for slot, a in enumerate((recs, loc, vo, io)):
        r.create_buffer(slot, a)
    r.bind_meshlets(s.mvp, recs.shape[1])
    st = r.draw_mesh_tasks(len(vo) - 1)
    ids1, d1 = r.visibility()
    print(f"[meshlets] {len(vo) - 1} meshlets, {len(loc) // 3} triangles, covered {(ids1 != 0xffffffff).sum()}",
          flush=True)
    assert np.array_equal(ids1, ids0)
    assert np.array_equal(d1.view(np.uint32), d0.view(np.uint32))
    assert np.array_equal(r.color(), c0)
    assert st["inputTriangles"] == len(loc) // 3
    # the bound inputs come back after the mesh draw
    r.draw()
    assert np.array_equal(r.visibility()[0], ids0)


def test_meshlets_from_a_later_work_group(renderer_factory):
    s, flat, recs, loc, vo, io = _meshlet_scene()
    k = 20
    r = renderer_factory()
    r.load_scene(flat)
    for slot, a in enumerate((recs, loc, vo, io)):
        r.create_buffer(slot, a)
    r.bind_meshlets(s.mvp, recs.shape[1])
    r.draw_mesh_tasks(len(vo) - 1 - k, first=k)
    ids, depth = r.visibility()
    # the remaining triangles, rebased onto the remaining vertices, numbered from 0
    sub = flat.indices[io[k]:] - vo[k]
    g, _ = O.render_scene(dataclasses.replace(flat, vertices=recs[vo[k]:], indices=sub))
    assert np.array_equal(ids, g.ids)
    assert np.array_equal(depth.view(np.uint32), g.depth.view(np.uint32))


def test_mesh_draw_errors(renderer_factory):
    from ifrit_hip import IfritError

    s, flat, recs, loc, vo, io = _meshlet_scene()
    r = renderer_factory()
    r.load_scene(flat)
    with pytest.raises(IfritError, match="no mesh program"):
        r.draw_mesh_tasks(4)
    r.bind_meshlets(s.mvp, recs.shape[1])
    with pytest.raises(IfritError, match="buffer slot"):
        r.draw_mesh_tasks(4)
    for slot, a in enumerate((recs, loc, vo, io)):
        r.create_buffer(slot, a)
    with pytest.raises(IfritError, match="beyond the offset buffers"):
        r.draw_mesh_tasks(len(vo))
    r.set_views([dict(width=64, height=64, mvp=s.mvp)])
    with pytest.raises(IfritError, match="single view"):
        r.draw_mesh_tasks(len(vo) - 1)
    r.set_views([])
    bad = loc.copy()
    bad[5] = 200  # beyond its meshlet's 64 vertices
    r.create_buffer(1, bad)
    with pytest.raises(IfritError, match="outside its meshlet"):
        r.draw_mesh_tasks(len(vo) - 1)
