"""Multi-view batches on the GPU.

* Every view of a batch (shadow cascades and cube faces depth-only, seeded
  cameras shaded, sizes that are not multiples of the 16-pixel tile) equals
  the same view drawn alone: primitive IDs and depth bits per sample, RGBA8
  colour per pixel, all bit for bit, with and without hipGraph replay.
* Two views are checked against the oracle directly (a cube face whose
  triangles cross w = 0, and a shaded camera).
* A replayed graph picks up new cameras from the view table.
* Every tuning configuration gives the same outputs (digest), and the
  autotuner reports them identical.
"""
import dataclasses

import numpy as np
import pytest

from ifrit_hip import scenes
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _scene(samples):
    return scenes.config2(width=64, height=64, samples=samples, spheres=12, level=2, tex_size=64, host_textures=True)


def _views(seed=7):
    return scenes.config4_views(seed=seed, cascade_size=200, cube_size=72, cameras=6, max_size=(300, 180),
                                min_size=24)


def _batch(renderer_factory, s, views, graph):
    r = renderer_factory()
    r.load_scene(s)
    r.set_views(views)
    r.set_graph_mode(graph)
    st = r.draw()
    out = []
    for i, v in enumerate(views):
        ids, depth = r.view_visibility(i)
        col = r.view_color(i) if v["shade"] else None
        out.append((ids, depth, col))
    return r, st, out


def _alone(r, s, v):
    """The view drawn as the renderer's single target."""
    r.set_views([])
    r.set_target(v["width"], v["height"], s.samples)
    r.bind_geometry(s.vertices, s.indices, v["mvp"])
    if v["shade"]:
        r.bind_fragment_program(dataclasses.replace(s, camera=v["camera"]))
    else:
        r.bind_fragment_program(s, program=scenes.PROGRAM_DEPTH_ONLY)
    r.draw()
    ids, depth = r.visibility()
    return ids, depth, (r.color() if v["shade"] else None)


def _same(a, b, label):
    ids_a, d_a, c_a = a
    ids_b, d_b, c_b = b
    bad_id = int((ids_a != ids_b).sum())
    bad_d = int((d_a.view(np.uint32) != d_b.view(np.uint32)).sum())
    bad_c = 0 if c_a is None else int((c_a != c_b).any(-1).sum())
    print(f"[{label}] samples {ids_a.size} covered {(ids_a != 0xffffffff).sum()} id/depth/colour mismatches "
          f"{bad_id}/{bad_d}/{bad_c}", flush=True)
    assert bad_id == 0 and bad_d == 0 and bad_c == 0


@pytest.mark.parametrize("samples,graph", [(1, False), (1, True), (4, True)])
def test_batch_views_equal_single_view_draws(renderer_factory, samples, graph):
    s = _scene(samples)
    views = _views()
    _, st, got = _batch(renderer_factory, s, views, graph)
    print(f"[batch S={samples} graph={graph}] {len(views)} views, stats {st}", flush=True)
    assert st["clippedTriangles"] > 0, "the cube faces should clip triangles crossing w = 0"
    r1 = renderer_factory()
    r1.load_scene(s)
    for i, v in enumerate(views):
        _same(got[i], _alone(r1, s, v), f"view {i} {v['kind']} {v['width']}x{v['height']}")


def test_batch_views_match_the_oracle(renderer_factory):
    s = _scene(1)
    views = _views()
    _, _, got = _batch(renderer_factory, s, views, True)
    for i in (4, 10):  # the first cube face, the first camera
        v = views[i]
        sv = dataclasses.replace(s, width=v["width"], height=v["height"], mvp=v["mvp"], camera=v["camera"])
        f, _ = O.render_scene(sv)
        ids, depth, col = got[i]
        bad_id = int((ids != f.ids).sum())
        bad_d = int((depth.view(np.uint32) != f.depth.view(np.uint32)).sum())
        print(f"[oracle view {i} {v['kind']}] covered {(f.ids != 0xffffffff).sum()} id/depth mismatches "
              f"{bad_id}/{bad_d}", flush=True)
        assert bad_id == 0 and bad_d == 0
        if col is not None:
            ref, flags = f.shade(sv)
            diff = np.abs(col.astype(int) - ref.astype(int)).max(-1)
            assert ((diff > 1) & (flags == 0)).sum() == 0


def test_graph_replay_picks_up_new_cameras(renderer_factory):
    s = _scene(1)
    views = _views()
    r, _, first = _batch(renderer_factory, s, views, True)
    moved = [dict(v) for v in views]
    for v in moved[10:]:
        v["mvp"] = (v["mvp"] @ np.diag([1.0, 1.0, 1.0, 1.0]).astype(np.float32)).astype(np.float32)
        v["mvp"][0, 3] += 0.25  # shift the image sideways
    r.set_views(moved)
    r.draw()
    r1 = renderer_factory()
    r1.load_scene(s)
    for i in (0, 10, 15):
        ids, depth = r.view_visibility(i)
        col = r.view_color(i) if moved[i]["shade"] else None
        _same((ids, depth, col), _alone(r1, s, moved[i]), f"moved view {i}")
    assert not np.array_equal(r.view_visibility(10)[0], first[10][0])


def test_tuning_never_changes_results(renderer_factory):
    s = _scene(1)
    r = renderer_factory()
    r.load_scene(s)
    r.set_views(_views())
    r.set_graph_mode(True)
    r.draw()
    ref = r.digest()
    for kw in (dict(shadeTile=16), dict(shadeTile=32), dict(rasterPartEntries=256), dict(rasterPartEntries=1024),
               dict(rasterGroupsPerCU=2), dict(rasterGroupsPerCU=8), dict(bigTriangleTiles=16),
               dict(rasterTileThreads=64), dict(rasterTileThreads=64, rasterPartEntries=256)):
        r.set_tuning(**kw)
        r.draw()
        d = r.digest()
        print(f"[tuning {kw}] digest {d:#018x} (reference {ref:#018x})", flush=True)
        assert d == ref
    rep = r.autotune(frames=2)
    print(f"[autotune] best {rep['best']} {rep['best_ms']:.3f} ms (start {rep['default_ms']:.3f} ms), "
          f"{rep['candidates']} candidates, bin demand {rep['bin_demand']}", flush=True)
    assert rep["identical"]
    assert rep["candidates"] >= 9
    assert rep["best"]["binCapacity"] >= rep["bin_demand"]
    r.draw()
    assert r.digest() == ref
