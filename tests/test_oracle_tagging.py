"""The tagging (visibility-buffer) mode's depth arithmetic.

With optsetForceDeterministic(false) and blending off, the reference draws
each tile in two passes (TileRasterWorker.cpp:1581-1675): a tagging pass that
keeps (primitive, barycentrics, depth) per pixel and a shading pass from the
tags.  Its depth expressions differ from the deterministic path's:
tile- and block-accepted fragments go through pixelShadingSIMD256Grouped,
whose barycentrics add the x term first (TileRasterWorker.cpp:2672-2684), and
single-pixel fragments through pixelShadingSingleQuad, which uses the
SIMD128 expression (TileRasterWorker.cpp:2123-2165) instead of the scalar DPPS
one.  Its traversal order depends on the worker threads, so only equal-depth
ties are left open; the oracle (and the GPU) resolve them to the lowest
primitive.

Pins: the reference build drawing in its tagging mode, on scenes without
equal-depth ties (random depths), with the oracle using this host's own RCPPS
so that the comparison holds on any x86 host.
"""
import numpy as np
import pytest

from oracle import oracle as O
from test_oracle_pins import _assert_same, _random_clip_scene

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="reference build (oracle/_ref) not present")


def _has_tagging():
    return O.ref_available() and hasattr(O.ref(), "ref_render_visibility_tagging")


@needs_ref
@pytest.mark.parametrize("seed", range(12))
def test_tagging_mode_matches_reference_build(seed):
    if not _has_tagging():
        pytest.skip("reference build predates the tagging entry point")
    W, H, clip, idx = _random_clip_scene(7100 + seed, seed % 3)  # kinds 0-2: continuous depths, no ties
    ids, depth, bad = O.ref_visibility_tagging(W, H, clip, idx)
    assert bad == 0
    f = O.Frame().render(W, H, 1, clip, idx, threads=4, live_rcp=True, tagging=True)
    _assert_same(f"tagging seed {seed}", f.ids, f.depth, ids, depth)


@needs_ref
def test_tagging_mode_matches_reference_build_config1():
    if not _has_tagging():
        pytest.skip("reference build predates the tagging entry point")
    from ifrit_hip import scenes

    s = scenes.config1()
    f, clip = O.render_scene(s, threads=8, live_rcp=True, tagging=True)
    ids, depth, bad = O.ref_visibility_tagging(s.width, s.height, clip, s.indices)
    assert bad == 0
    # the sphere's neighbouring triangles both cover some pixels on their shared edges at equal
    # depth; the reference's traversal order decides those, the oracle takes the lower primitive
    cov = ids != 0xFFFFFFFF
    assert np.array_equal(cov, f.ids != 0xFFFFFFFF)
    assert np.array_equal(depth.view(np.uint32)[cov], f.depth.view(np.uint32)[cov])
    ties = np.flatnonzero(ids != f.ids)
    print(f"config1 tagging: {ties.size} pixels where the reference took the other of two equal-depth fragments")
    assert ties.size < 200
    tri = s.indices.reshape(-1, 3)
    for p in ties:
        # the reference's primitive alone covers the pixel with the same depth bits, and its id is the larger
        alone = O.Frame().render(s.width, s.height, 1, clip, tri[ids[p]].copy(), threads=1, live_rcp=True,
                                 tagging=True)
        assert alone.ids[p] == 0 and alone.depth.view(np.uint32)[p] == depth.view(np.uint32)[p]
        assert ids[p] > f.ids[p]


def test_tagging_mode_changes_depth_bits_and_classes():
    """The two modes cover the same fragments; they differ in the depth bits
    of some tile/block-accepted and single-pixel fragments, and the tagging
    mode reports the grouped form (4) and never the scalar one (3).  Without
    these differences the comparison above would not tell the modes apart."""
    from ifrit_hip import scenes

    s = scenes.config1()
    det, _ = O.render_scene(s, threads=8)
    tag, _ = O.render_scene(s, threads=8, tagging=True)
    assert det.stats == tag.stats
    cov = det.ids != 0xFFFFFFFF
    assert np.array_equal(cov, tag.ids != 0xFFFFFFFF)
    assert set(np.unique(tag.cls[cov])) <= {1, 2, 4} and (tag.cls == 4).any()
    assert (det.cls == 3).any() and not (tag.cls == 3).any()
    # where the same fragment wins with the same class (SIMD256 or SIMD128) the expressions agree
    same = cov & (det.cls == tag.cls) & (det.ids == tag.ids)
    assert np.array_equal(det.depth[same].view(np.uint32), tag.depth[same].view(np.uint32))
    changed = cov & (det.depth.view(np.uint32) != tag.depth.view(np.uint32))
    print(f"tagging vs deterministic: {int(changed.sum())} of {int(cov.sum())} covered pixels' depth bits differ")
    assert changed.any()
    # the accepted fragments' grouped depth differs from SIMD256 only by rounding
    acc = cov & (tag.cls == 4) & (det.ids == tag.ids)
    assert np.allclose(det.depth[acc], tag.depth[acc], rtol=1e-5, atol=1e-6)


def test_tagging_mode_is_ignored_with_msaa_and_blending():
    W, H, clip, idx = _random_clip_scene(7200, 0)
    a = O.Frame().render(W, H, 4, clip, idx, threads=2)
    b = O.Frame().render(W, H, 4, clip, idx, threads=2, tagging=True)
    assert np.array_equal(a.ids, b.ids) and np.array_equal(a.depth.view(np.uint32), b.depth.view(np.uint32))
