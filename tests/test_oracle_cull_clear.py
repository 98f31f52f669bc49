"""Pins of the oracle's cull modes and clear values.

The reference CPU path has one cull (back faces) and one clear value
(depth 0x7f7f7f7f, colour 0); the CUDA path's SetCullMode and setClearValues
(TileRasterRendererCuda.h:109, :111; TileRasterCoreInvocationCuda.cu:1433-1462)
are the spec of the others.  What fixes them:

* cull mode FRONT on a scene equals BACK on the same scene with every
  triangle's first and third index exchanged, bit for bit (visibility,
  depth and shaded colour) -- and that reversed scene is drawn by the
  reference build itself;
* cull mode NONE keeps, per sample, the lexicographic minimum of
  (depth, primitive, fan) over the fragments of the BACK and FRONT frames
  (triangles in front of w = 0; one that crosses it decides per clip fan,
  and the GPU is held to the oracle there);
* FRONT_AND_BACK draws nothing;
* a clear depth D keeps exactly the samples whose default-clear winner is
  nearer than D (the depth test is strict), and leaves D elsewhere;
* a clear colour C weighs the samples no triangle covers: an MSAA edge pixel
  with k of S samples covered resolves to (k c + (S - k) C) / S.
"""
import dataclasses

import numpy as np
import pytest

from ifrit_hip import scenes
from oracle import oracle as O
from test_oracle_pins import _random_clip_scene

NONE, FRONT, BACK, FRONT_AND_BACK = 0, 1, 2, 3


def _reverse(idx):
    return idx.reshape(-1, 3)[:, ::-1].reshape(-1).copy()


def _render(W, H, S, clip, idx, cull):
    return O.Frame().render(W, H, S, clip, idx, threads=4, cull_mode=cull)


@pytest.mark.parametrize("samples", [1, 4, 8])
def test_front_equals_back_of_reversed_winding(samples):
    for seed in range(12):
        W, H, clip, idx = _random_clip_scene(700 + seed, seed % 4)
        a = _render(W, H, samples, clip, idx, FRONT)
        b = _render(W, H, samples, clip, _reverse(idx), BACK)
        assert np.array_equal(a.ids, b.ids), f"seed {seed}"
        assert np.array_equal(a.depth.view(np.uint32), b.depth.view(np.uint32)), f"seed {seed}"
        assert a.stats == b.stats


def test_front_shades_like_back_of_reversed_winding():
    """Attributes follow the exchanged vertex order: Blinn-Phong on config 1
    with FRONT equals BACK on the reversed index buffer, byte for byte."""
    s = scenes.config1(200, 150)
    clip = O.vertex_transform(s.vertices, s.mvp)
    a = O.Frame().render(s.width, s.height, 1, clip, s.indices, cull_mode=FRONT)
    rs = dataclasses.replace(s, indices=_reverse(s.indices))
    b = O.Frame().render(s.width, s.height, 1, clip, rs.indices, cull_mode=BACK)
    assert np.array_equal(a.ids, b.ids) and (a.ids != 0xFFFFFFFF).sum() > 1000
    ca, _ = a.shade(s)
    cb, _ = b.shade(rs)
    assert np.array_equal(ca, cb)


@pytest.mark.skipif(not O.ref_available(), reason="reference build (oracle/_ref) not present")
def test_front_matches_reference_build_on_reversed_winding():
    from test_oracle_pins import _assert_same

    for seed in range(16):
        W, H, clip, idx = _random_clip_scene(900 + seed, seed % 4)
        ids, depth, bad = O.ref_visibility(W, H, clip, _reverse(idx))
        assert bad == 0
        f = _render(W, H, 1, clip, idx, FRONT)
        _assert_same(f"front cull seed {seed}", f.ids, f.depth, ids, depth)


def _keys(f, S):
    tris = f.tris()
    fan = np.where(f.tri == 0xFFFFFFFF, 0, tris["fan"][np.minimum(f.tri, max(len(tris) - 1, 0))])
    key = np.where(f.ids == 0xFFFFFFFF, np.uint64(0xFFFFFFFFFFFFFFFF),
                   f.ids.astype(np.uint64) * np.uint64(4) + fan.astype(np.uint64))
    return key


@pytest.mark.parametrize("samples", [1, 4])
def test_none_is_the_nearer_of_back_and_front(samples):
    # triangles in front of w = 0 (scene kinds 0, 2, 3); a triangle crossing w = 0 decides per clip
    # fan, whose triangulation differs between the two windings
    for seed in range(12):
        W, H, clip, idx = _random_clip_scene(1300 + seed, (0, 2, 3)[seed % 3])
        n = _render(W, H, samples, clip, idx, NONE)
        b = _render(W, H, samples, clip, idx, BACK)
        fr = _render(W, H, samples, clip, idx, FRONT)
        kb, kf = _keys(b, samples), _keys(fr, samples)
        db, df = b.depth.astype(np.float64), fr.depth.astype(np.float64)
        take_f = (df < db) | ((df == db) & (kf < kb))
        want_ids = np.where(take_f, fr.ids, b.ids)
        want_d = np.where(take_f, fr.depth, b.depth)
        assert np.array_equal(n.ids, want_ids), f"seed {seed}"
        assert np.array_equal(n.depth.view(np.uint32), want_d.view(np.uint32)), f"seed {seed}"
        assert n.stats["assembled_triangles"] == b.stats["assembled_triangles"] + fr.stats["assembled_triangles"]


def test_front_and_back_draws_nothing():
    W, H, clip, idx = _random_clip_scene(77, 0)
    f = _render(W, H, 4, clip, idx, FRONT_AND_BACK)
    assert np.all(f.ids == 0xFFFFFFFF) and f.stats["assembled_triangles"] == 0
    assert np.all(f.depth.view(np.uint32) == 0x7F7F7F7F)


@pytest.mark.parametrize("samples", [1, 8])
def test_clear_depth_keeps_the_nearer_winners(samples):
    for seed in range(8):
        W, H, clip, idx = _random_clip_scene(2100 + seed, seed % 4)
        ref = _render(W, H, samples, clip, idx, BACK)
        for D in (0.5, 0.25, np.float32(ref.depth[ref.ids != 0xFFFFFFFF].min()) if (ref.ids != 0xFFFFFFFF).any()
                  else 0.5):
            f = O.Frame().render(W, H, samples, clip, idx, threads=4, clear_depth=float(D))
            keep = ref.depth < np.float32(D)
            assert np.array_equal(f.ids, np.where(keep, ref.ids, 0xFFFFFFFF)), (seed, D)
            want = np.where(keep, ref.depth, np.float32(D)).astype(np.float32)
            assert np.array_equal(f.depth.view(np.uint32), want.view(np.uint32)), (seed, D)


@pytest.mark.parametrize("samples", [4, 8])
def test_clear_colour_weighs_uncovered_samples(samples):
    from test_oracle_shading_pins import _expected_resolve, _rect_scene
    from test_oracle_shading_pins import _render as render_rect

    s = _rect_scene(samples)
    f = render_rect(s)
    counts, clean = _expected_resolve(samples)
    C = np.array([40, 200, 90, 255], np.int64)  # clear colour in RGBA8 steps
    hb = np.array([[(O.flat_hash(p) >> (8 * c)) & 255 for c in range(3)] + [255] for p in (0, 1)], np.int64)
    empty = samples - counts.sum(-1, keepdims=True)
    num = counts[..., 0:1] * hb[0] + counts[..., 1:2] * hb[1] + empty * C
    rgba, _ = f.shade(s, clear_color=tuple(C / 255.0))
    got = rgba.astype(np.int64)
    near = num // samples + ((2 * (num % samples)) > samples)
    tie = (2 * (num % samples)) == samples
    ok = (got == near) | (tie & (got == near + 1))
    assert np.all(ok[clean])
    assert (clean & (empty[..., 0] == samples)).sum() > 100  # background pixels take the clear colour
    assert np.all(got[clean & (empty[..., 0] == samples)] == C)
