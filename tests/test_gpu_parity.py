"""GPU parity: the HIP pipeline against the oracle on the same inputs.

Visibility (primitive id per sample), depth bits and the 1-spp coverage
class must be bit-identical; colour must be within one RGBA8 step of the
FP64 shading oracle outside pixels the oracle flags as a too-close call.
"""
import numpy as np
import pytest

from ifrit_hip import scenes
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _render_gpu(renderer_factory, scene):
    r = renderer_factory()
    r.load_scene(scene)
    st = r.draw()
    ids, depth = r.visibility()
    return r, st, ids, depth


def _compare_vis(scene, f, ids, depth, label):
    bad_id = np.flatnonzero(ids != f.ids)
    bad_d = np.flatnonzero(depth.view(np.uint32) != f.depth.view(np.uint32))
    print(f"[{label}] pixels {ids.size} covered {(f.ids != 0xffffffff).sum()} id mismatches {bad_id.size} "
          f"depth mismatches {bad_d.size}", flush=True)
    assert bad_id.size == 0, f"first id mismatches at {bad_id[:8]}: gpu {ids[bad_id[:8]]} oracle {f.ids[bad_id[:8]]}"
    assert bad_d.size == 0, f"first depth mismatches at {bad_d[:8]}"


def test_config1_visibility_and_color(renderer_factory):
    s = scenes.config1()
    f, _ = O.render_scene(s)
    r, st, ids, depth = _render_gpu(renderer_factory, s)
    _compare_vis(s, f, ids, depth, "config1")
    cls = r.coverage_class()
    assert np.array_equal(cls, f.cls), f"class mismatches {(cls != f.cls).sum()}"
    assert st["fragments"] == f.stats["fragments"]
    assert st["assembledTriangles"] == f.stats["assembled_triangles"]
    rgba = r.color()
    ref, flags = f.shade(s)
    diff = np.abs(rgba.astype(int) - ref.astype(int)).max(-1)
    bad = (diff > 1) & (flags == 0)
    print(f"[config1] colour max diff {diff.max()} >1LSB unflagged {bad.sum()} flagged {flags.sum()}", flush=True)
    assert bad.sum() == 0
