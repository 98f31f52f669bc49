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


def _random_scene(seed, kind):
    rng = np.random.default_rng(seed)
    W = int(rng.integers(17, 300))
    H = int(rng.integers(17, 260))
    n = int(rng.integers(1, 400))
    if kind == 0:
        clip = np.concatenate([rng.uniform(-1.5, 1.5, (3 * n, 2)), rng.uniform(-0.2, 1.2, (3 * n, 1)),
                               rng.uniform(0.2, 2, (3 * n, 1))], 1)
        clip[:, :3] *= clip[:, 3:4]
    elif kind == 1:
        clip = np.concatenate([rng.uniform(-2, 2, (3 * n, 2)), rng.uniform(-1, 2, (3 * n, 1)),
                               rng.uniform(-1, 2, (3 * n, 1))], 1)
    elif kind == 2:
        c = rng.uniform(-1, 1, (n, 1, 2))
        d = rng.normal(0, 3.0 / max(W, H), (n, 3, 2))
        xy = (c + d).reshape(-1, 2)
        clip = np.concatenate([xy, np.repeat(rng.uniform(0, 1, (n, 1)), 3, 0), np.ones((3 * n, 1))], 1)
    else:
        w = np.repeat(rng.uniform(0.5, 2, (n, 1)), 3, 0)
        xy = rng.uniform(-3, 3, (3 * n, 2))
        z = np.repeat(rng.choice([0.25, 0.5, 0.75], (n, 1)), 3, 0)
        clip = np.concatenate([xy * w, z * w, w], 1)
    idx = np.arange(3 * n, dtype=np.int32).reshape(-1, 3)
    flip = rng.random(n) < 0.5
    idx[flip] = idx[flip][:, ::-1]
    return W, H, clip.astype(np.float32), idx.reshape(-1)


@pytest.mark.parametrize("samples", [1, 4, 8])
def test_random_clip_space_scenes(renderer_factory, samples):
    """Random clip-space triangles (including w-crossing ones that take the
    homogeneous clipper, sub-pixel ones and equal-depth ties), drawn through
    an identity vertex program so the GPU sees exactly these clip positions."""
    r = renderer_factory()
    for it in range(24):
        W, H, clip, idx = _random_scene(1000 * samples + it, it % 4)
        f = O.Frame().render(W, H, samples, clip, idx)
        r.set_target(W, H, samples)
        r.bind_geometry(clip, idx, np.eye(4, dtype=np.float32), clip_passthrough=True)
        s = scenes.Scene("rand", W, H, samples, clip, idx, np.eye(4, dtype=np.float32), program=scenes.PROGRAM_FLAT)
        r.bind_fragment_program(s)
        st = r.draw()
        ids, depth = r.visibility()
        _compare_vis(s, f, ids, depth, f"rand s={samples} it={it} kind={it % 4} {W}x{H}")
        assert st["fragments"] == f.stats["fragments"]
        assert st["coveredSamples"] == f.stats["covered_samples"]
        if samples == 1:
            assert np.array_equal(r.coverage_class(), f.cls)
        ref, _ = f.shade(s)
        assert np.abs(r.color().astype(int) - ref.astype(int)).max() <= 1


def test_config3_scaled_visibility(renderer_factory):
    s = scenes.config3(width=480, height=270, samples=4, gx=256, gy=128)
    f, _ = O.render_scene(s)
    r, st, ids, depth = _render_gpu(renderer_factory, s)
    _compare_vis(s, f, ids, depth, "config3 scaled")
    assert st["fragments"] == f.stats["fragments"]
    ref, _ = f.shade(s)
    assert np.abs(r.color().astype(int) - ref.astype(int)).max() <= 1


def test_config2_scaled_visibility_and_color(renderer_factory):
    s = scenes.config2(width=480, height=270, samples=8, spheres=24, level=3, tex_size=256, host_textures=True)
    f, _ = O.render_scene(s)
    r, st, ids, depth = _render_gpu(renderer_factory, s)
    _compare_vis(s, f, ids, depth, "config2 scaled")
    rgba = r.color()
    ref, flags = f.shade(s)
    diff = np.abs(rgba.astype(int) - ref.astype(int)).max(-1)
    bad = (diff > 1) & (flags == 0)
    print(f"[config2 scaled] colour max diff {diff.max()} >1LSB unflagged {bad.sum()} flagged {flags.sum()}", flush=True)
    assert bad.sum() == 0


@pytest.mark.parametrize("samples", [1, 4])
def test_config5_layered_ggx_color(renderer_factory, samples):
    """Layered GGX (albedo, tangent-space normal and roughness layers, 16x
    anisotropic) against the FP64 oracle on the same visibility."""
    s = scenes.config5(width=480, height=270, samples=samples, spheres=48, level=3, tex_size=256,
                       host_textures=True)
    f, _ = O.render_scene(s)
    r, st, ids, depth = _render_gpu(renderer_factory, s)
    _compare_vis(s, f, ids, depth, f"config5 scaled s={samples}")
    rgba = r.color()
    ref, flags = f.shade(s)
    diff = np.abs(rgba.astype(int) - ref.astype(int)).max(-1)
    bad = (diff > 1) & (flags == 0)
    covered = (f.ids.reshape(s.height, s.width, samples) != 0xFFFFFFFF).any(-1)
    print(f"[config5 scaled s={samples}] covered {covered.sum()} colour max diff {diff.max()} "
          f">1LSB unflagged {bad.sum()} flagged {flags.sum()}", flush=True)
    assert covered.sum() > 1000
    assert bad.sum() == 0


def test_fbm_and_mips_match_host_definition(renderer_factory):
    r = renderer_factory()
    r.set_target(16, 16, 1)
    for size, seed in [(256, 3), (64, 12345)]:
        t = scenes.Texture(size, size, scenes.num_levels(size, size), None, seed)
        tid = r.add_texture(t)
        got = r.download_texture(tid, t)
        want = scenes.build_mips(scenes.fbm_texture(size, seed))
        assert np.array_equal(got, want), f"fbm/mips differ at {np.flatnonzero(got != want)[:8]}"


def test_draw_is_deterministic(renderer_factory):
    s = scenes.config2(width=640, height=360, samples=8, spheres=32, level=3, tex_size=128, host_textures=True)
    r = renderer_factory()
    r.load_scene(s)
    r.draw()
    a = r.visibility()
    ca = r.color()
    r.draw()
    b = r.visibility()
    cb = r.color()
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32))
    assert np.array_equal(ca, cb)


@pytest.mark.parametrize("which", ["config2", "config3"])
def test_full_size_visibility(renderer_factory, which):
    """The benchmark workloads at full size against the multithreaded oracle."""
    s = scenes.config2() if which == "config2" else scenes.config3()
    f, _ = O.render_scene(s)
    r = renderer_factory()
    r.set_target(s.width, s.height, s.samples)
    r.bind_geometry(s.vertices, s.indices, s.mvp)
    s2 = scenes.Scene(s.name, s.width, s.height, s.samples, s.vertices, s.indices, s.mvp, program=scenes.PROGRAM_FLAT)
    r.bind_fragment_program(s2)
    st = r.draw()
    ids, depth = r.visibility()
    _compare_vis(s, f, ids, depth, which + " full")
    assert st["fragments"] == f.stats["fragments"]
    assert st["assembledTriangles"] == f.stats["assembled_triangles"]
