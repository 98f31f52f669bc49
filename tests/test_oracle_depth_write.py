"""Fragment programs that write depth (the FragmentShader depth
out-parameter, Shaders.h:64), pinned to the reference build.

In the CPU path a fragment is tested with its interpolated depth; when it
passes, the fragment shader runs and the depth it leaves in its out-parameter
is what the depth cache and attachment keep (TileRasterWorker.cpp:2272-2384,
2880-2916), so later fragments are tested against the written depth.  The
reference build draws with a fragment shader writing depth * scale + bias
(oracle/ref/ref_harness.cpp DepthAffineFS); the oracle must give the same
primitive and the same written depth bits at every pixel.
"""
import numpy as np
import pytest

from oracle import oracle as O
from test_oracle_pins import HOST_RCP, _assert_same, _random_clip_scene

AFFINE = [(1.0, 0.05), (0.5, 0.25), (1.0, -0.03)]


def _has_entry():
    return O.ref_available() and hasattr(O.ref(), "ref_render_visibility_depthwrite")


@pytest.mark.skipif(not O.ref_available(), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("func", [1, 3, 4, 6, 7])
@pytest.mark.parametrize("affine", range(len(AFFINE)))
def test_depth_writing_program_matches_reference_build(func, affine):
    if not _has_entry():
        pytest.skip("reference build predates the depth-writing entry point")
    scale, bias = AFFINE[affine]
    changed = 0
    for seed in range(6):
        W, H, clip, idx = _random_clip_scene(8100 + 10 * func + seed, seed % 4)
        ids, depth, bad = O.ref_visibility_depthwrite(W, H, clip, idx, scale, bias, depth_func=func)
        assert bad == 0
        f = O.Frame().render(W, H, 1, clip, idx, depth_func=func, threads=4, live_rcp=HOST_RCP,
                             depth_write=(scale, bias))
        _assert_same(f"depth write {scale}, {bias} func {func} seed {seed}", f.ids, f.depth, ids, depth)
        plain = O.Frame().render(W, H, 1, clip, idx, depth_func=func, threads=4, live_rcp=HOST_RCP)
        changed += int((plain.ids != f.ids).sum())
    print(f"func {func} affine {scale},{bias}: {changed} pixels won by another primitive than without depth writes")


def test_depth_write_closed_form():
    """Two screen-aligned quads at depths 0.5 (first) and 0.52 (second): with LESS and a program
    writing depth + 0.05, the second passes against the first's written 0.55 and wins, keeping
    0.57; with bias 0.01 the first keeps 0.51 and the second fails."""
    def quad(z, x0, x1):
        return np.array([[x0, -0.5, z, 1], [x1, -0.5, z, 1], [x1, 0.5, z, 1], [x0, 0.5, z, 1]], np.float32)

    clip = np.concatenate([quad(0.5, -0.5, 0.5), quad(0.52, -0.5, 0.5)])
    idx = np.array([0, 2, 1, 0, 3, 2, 4, 6, 5, 4, 7, 6], np.int32)
    for order in (idx, idx.reshape(-1, 3)[:, ::-1].reshape(-1).copy()):
        f = O.Frame().render(32, 32, 1, clip, order, threads=1)
        if (f.ids != 0xFFFFFFFF).any():
            idx = order
            break
    c = 16 * 32 + 16
    # each quad's own interpolated depth at the pixel (within rounding of 0.5 and 0.52)
    d1 = O.Frame().render(32, 32, 1, clip, idx[:6], threads=1).depth[c]
    d2 = O.Frame().render(32, 32, 1, clip, idx[6:], threads=1).depth[c]
    assert abs(d1 - 0.5) < 1e-6 and abs(d2 - 0.52) < 1e-6
    f = O.Frame().render(32, 32, 1, clip, idx, threads=1, depth_write=(1.0, 0.05))
    assert f.ids[c] in (2, 3) and f.depth[c] == np.float32(d2 * np.float32(1.0)) + np.float32(0.05)
    f = O.Frame().render(32, 32, 1, clip, idx, threads=1, depth_write=(1.0, 0.01))
    assert f.ids[c] in (0, 1) and f.depth[c] == np.float32(d1 * np.float32(1.0)) + np.float32(0.01)
    # the tagging mode's shading pass discards written depth
    f = O.Frame().render(32, 32, 1, clip, idx, threads=1, depth_write=(1.0, 0.05), tagging=True)
    assert f.ids[c] in (0, 1) and f.depth[c] == d1
