"""Blending on the GPU (the ordered path, 1 sample per pixel).

* Translucent hash-coloured triangles blended in primitive order against the
  oracle's blend, which the CPU tests pin bit for bit to the reference build:
  the FP32 colour must be identical, as must the IDs and depth of the last
  writer, for several factor sets (with the reference's encoding of the alpha
  factors) and depth functions.
* Blending without clearing accumulates onto the previous colour.
* Opaque blending (ONE, ZERO) of the textured Blinn-Phong config 1 gives the
  same colour, FP32 and RGBA8, as the deferred shading path.
"""
import numpy as np
import pytest

from ifrit_hip import scenes
from oracle import oracle as O

from test_gpu_parity import _random_scene
from test_oracle_pins import BF_ONE, BF_ZERO, BLENDS

pytestmark = pytest.mark.gpu


def _setup(r, W, H, clip, idx, factors, func=1):
    r.set_target(W, H, 1)
    r.bind_geometry(clip, idx, np.eye(4, dtype=np.float32), clip_passthrough=True)
    s = scenes.Scene("rand", W, H, 1, clip, idx, np.eye(4, dtype=np.float32), program=scenes.PROGRAM_FLAT_ALPHA)
    r.bind_fragment_program(s)
    r.set_depth_func(func)
    r.set_blend(factors)


@pytest.mark.parametrize("func", [1, 7])
@pytest.mark.parametrize("blend", range(len(BLENDS)))
def test_translucent_blending_matches_oracle(renderer_factory, blend, func):
    r = renderer_factory()
    for it in range(6):
        W, H, clip, idx = _random_scene(5100 + 10 * blend + it, it % 4)
        _setup(r, W, H, clip, idx, BLENDS[blend], func)
        r.draw()
        got = r.color_f32()
        ids, depth = r.visibility()
        f = O.Frame().render_blend(W, H, clip, idx, BLENDS[blend], depth_func=func)
        bad = (got.view(np.uint32) != f.rgba.view(np.uint32)).any(-1)
        print(f"[blend {BLENDS[blend]} func {func} it {it}] {W}x{H} covered {(f.ids != 0xffffffff).sum()} "
              f"colour mismatches {bad.sum()}", flush=True)
        assert bad.sum() == 0
        assert np.array_equal(ids, f.ids)
        assert np.array_equal(depth.view(np.uint32), f.depth.view(np.uint32))
        q = np.floor(np.clip(got, 0, 1) * 255.0 + 0.5).astype(np.uint8)
        assert np.array_equal(r.color(), q)
    r.set_blend(None)
    r.set_depth_func(1)


def test_blending_without_clear_accumulates(renderer_factory):
    r = renderer_factory()
    W, H, clip, idx = _random_scene(5300, 3)
    factors = (BF_ONE, BF_ONE, BF_ONE, BF_ONE)
    _setup(r, W, H, clip, idx, factors, func=7)
    r.draw()
    first = r.color_f32()
    r.draw(clear=False)
    second = r.color_f32()
    f1 = O.Frame().render_blend(W, H, clip, idx, factors, depth_func=7)
    f2 = O.Frame().render_blend(W, H, clip, idx, factors, depth_func=7, clear=False, depth_in=f1.depth,
                                rgba_in=f1.rgba)
    assert np.array_equal(first.view(np.uint32), f1.rgba.view(np.uint32))
    assert np.array_equal(second.view(np.uint32), f2.rgba.view(np.uint32))
    assert (second > first).any()
    r.set_blend(None)
    r.set_depth_func(1)


def test_opaque_blending_equals_deferred_shading(renderer_factory):
    s = scenes.config1()
    r = renderer_factory()
    r.load_scene(s)
    r.set_float_output(True)
    r.draw()
    want8, want32 = r.color(), r.color_f32()
    ids0, _ = r.visibility()
    r.set_blend((BF_ONE, BF_ZERO, BF_ONE, BF_ZERO))
    r.draw()
    got8, got32 = r.color(), r.color_f32()
    ids1, _ = r.visibility()
    r.set_blend(None)
    cov = ids0 != 0xFFFFFFFF
    print(f"[opaque blend] covered {cov.sum()} colour mismatches {(got8 != want8).any(-1).sum()}", flush=True)
    assert np.array_equal(ids0, ids1)
    assert np.array_equal(got8, want8)
    assert np.array_equal(got32.reshape(-1, 4)[cov].view(np.uint32), want32.reshape(-1, 4)[cov].view(np.uint32))
