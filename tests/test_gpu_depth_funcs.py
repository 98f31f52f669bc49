"""Depth functions other than LESS on the GPU (the ordered raster path).

Each function against the oracle's primitive-order restatement, which the
CPU tests pin to the reference build for all eight functions: primitive IDs,
depth bits and 1-spp coverage classes per sample must be identical,
including the reference's scalar-path GREATER that tests nothing.  Also a
depth pre-pass followed by EQUAL without clearing, and a tile whose list is
longer than the sorting window.
"""
import numpy as np
import pytest

from ifrit_hip import scenes
from oracle import oracle as O

from test_gpu_parity import _compare_vis, _random_scene

pytestmark = pytest.mark.gpu

FUNCS = {0: "NEVER", 2: "EQUAL", 3: "LESS_OR_EQUAL", 4: "GREATER", 5: "NOT_EQUAL", 6: "GREATER_OR_EQUAL", 7: "ALWAYS"}


def _draw(r, W, H, S, clip, idx, func, clear=True):
    r.set_target(W, H, S)
    r.bind_geometry(clip, idx, np.eye(4, dtype=np.float32), clip_passthrough=True)
    s = scenes.Scene("rand", W, H, S, clip, idx, np.eye(4, dtype=np.float32), program=scenes.PROGRAM_FLAT)
    r.bind_fragment_program(s)
    r.set_depth_func(func)
    st = r.draw(clear=clear)
    return s, st


@pytest.mark.parametrize("samples", [1, 4])
@pytest.mark.parametrize("func", sorted(FUNCS))
def test_depth_function_matches_oracle(renderer_factory, func, samples):
    r = renderer_factory()
    for it in range(8):
        W, H, clip, idx = _random_scene(3000 + 100 * func + it, it % 4)
        f = O.Frame().render(W, H, samples, clip, idx, depth_func=func)
        s, st = _draw(r, W, H, samples, clip, idx, func)
        ids, depth = r.visibility()
        _compare_vis(s, f, ids, depth, f"{FUNCS[func]} s={samples} it={it}")
        assert st["fragments"] == f.stats["fragments"]
        assert st["coveredSamples"] == f.stats["covered_samples"]
        if samples == 1:
            assert np.array_equal(r.coverage_class(), f.cls)
    r.set_depth_func(1)


def test_depth_prepass_then_equal(renderer_factory):
    """LESS with clear, then EQUAL without clear over the same triangles: a pixel keeps the last
    fragment (in primitive order) whose depth equals the stored one."""
    r = renderer_factory()
    W, H, clip, idx = _random_scene(4242, 3)  # quantised depths: many equal-depth ties
    _draw(r, W, H, 1, clip, idx, 1)
    _, d0 = r.visibility()
    _draw(r, W, H, 1, clip, idx, 2, clear=False)
    ids, depth = r.visibility()
    f = O.Frame().render(W, H, 1, clip, idx, depth_func=2, clear=False, depth_in=d0)
    cov = f.ids != 0xFFFFFFFF
    print(f"[prepass + EQUAL] covered {cov.sum()} id mismatches {(ids[cov] != f.ids[cov]).sum()}", flush=True)
    assert cov.sum() > 100
    assert np.array_equal(ids[cov], f.ids[cov])
    assert np.array_equal(depth.view(np.uint32), d0.view(np.uint32))
    r.set_depth_func(1)


@pytest.mark.parametrize("func", [3, 7])
def test_long_tile_list_takes_sorted_windows(renderer_factory, func):
    """6000 small triangles inside one 16x16 tile: the list exceeds the 2048-entry sorting window."""
    rng = np.random.default_rng(17)
    n = 6000
    W = H = 32
    c = rng.uniform(-0.95, -0.05, (n, 1, 2))  # the lower-left quarter: tile (0, 0) of a 32x32 target
    d = rng.uniform(-0.08, 0.08, (n, 3, 2))
    xy = (c + d).reshape(-1, 2)
    z = np.repeat(rng.choice([0.2, 0.4, 0.6], (n, 1)), 3, 0)
    clip = np.concatenate([xy, z, np.ones((3 * n, 1))], 1).astype(np.float32)
    idx = np.arange(3 * n, dtype=np.int32).reshape(-1, 3)
    flip = rng.random(n) < 0.5
    idx[flip] = idx[flip][:, ::-1]
    idx = idx.reshape(-1)
    f = O.Frame().render(W, H, 1, clip, idx, depth_func=func)
    s, st = _draw(r := renderer_factory(), W, H, 1, clip, idx, func)
    ids, depth = r.visibility()
    print(f"[long list {FUNCS[func]}] bin entries {st['binEntries']}", flush=True)
    assert st["binEntries"] > 2048
    _compare_vis(s, f, ids, depth, f"long list {FUNCS[func]}")
    assert np.array_equal(r.coverage_class(), f.cls)
    r.set_depth_func(1)
