"""The texture pipeline on the GPU against the oracle: sampler objects
(addressing modes, border colours, nearest texels, anisotropic taps) on a
textured quad whose coordinates run outside [0, 1], cube maps on a sphere at
two levels of detail, layered uploads with mip generation, and blitImage
between levels (bit for bit against the oracle's blit)."""
import dataclasses
import math

import numpy as np
import pytest

from ifrit_hip import scenes
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _quad(W=160, H=120, uv=(-1.25, 2.25)):
    lo, hi = uv
    verts = np.array([[-2, -1.5, 5, 0, 0, -1, lo, lo], [2, -1.5, 5, 0, 0, -1, hi, lo], [2, 1.5, 5, 0, 0, -1, hi, hi],
                      [-2, 1.5, 5, 0, 0, -1, lo, hi]], np.float32)
    rng = np.random.default_rng(21)
    lvl0 = rng.integers(0, 256, (16, 16, 4), dtype=np.uint8)
    lvl0[..., 3] = 255
    mvp = scenes.perspective(math.radians(50), W / H, 0.1, 100) @ scenes.look_at((0, 0, 0), (0, 0, 1), (0, 1, 0))
    s = scenes.Scene("quad", W, H, 1, verts, np.array([0, 2, 1, 0, 3, 2], np.int32), mvp.astype(np.float32),
                     program=scenes.PROGRAM_BLINN_PHONG, camera=(0, 0, 0), ambient=0.2, kd=0.8, ks=0.0,
                     lights=[((0.0, 0.0, -1.0, 0.0), (1.0, 1.0, 1.0, 0.0))],
                     textures=[scenes.Texture(16, 16, 5, scenes.build_mips(lvl0))])
    return s


def _front(s):
    """The scene with the winding that faces the camera (the back-face cull keeps it)."""
    f, _ = O.render_scene(s)
    if (f.ids == 0xFFFFFFFF).all():
        s = dataclasses.replace(s, indices=s.indices.reshape(-1, 3)[:, ::-1].reshape(-1).copy())
    return s


def _check(r, s, label):
    r.draw()
    f, _ = O.render_scene(s)
    assert np.array_equal(r.visibility()[0], f.ids)
    ref, flags = f.shade(s)
    diff = np.abs(r.color().astype(int) - ref.astype(int)).max(-1)
    covered = (f.ids != 0xFFFFFFFF).sum()
    print(f"[{label}] covered {covered} colour max diff {diff.max()} flagged {int(flags.sum())}", flush=True)
    assert covered > 1000
    assert ((diff > 1) & (flags == 0)).sum() == 0


@pytest.mark.parametrize("mode", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("filt,border", [(1, 0), (0, 1)])
def test_sampler_address_modes(renderer_factory, mode, filt, border):
    s = _quad()
    s = _front(dataclasses.replace(s, sampler_state=dict(filter=filt, address_u=mode, address_v=(mode + 2) % 5,
                                                         border=border)))
    r = renderer_factory()
    r.load_scene(s)
    _check(r, s, f"sampler mode {mode} filter {filt} border {border}")


def test_sampler_anisotropic(renderer_factory):
    s = _quad(uv=(-0.5, 1.5))
    # a tilted quad: anisotropic footprints
    s.vertices[:, 2] = np.array([3, 3, 9, 9], np.float32)
    s = _front(dataclasses.replace(s, sampler_state=dict(filter=1, address_u=2, address_v=1, anisotropy=True,
                                                         max_anisotropy=8.0)))
    r = renderer_factory()
    r.load_scene(s)
    _check(r, s, "sampler anisotropic clamp/mirror")


@pytest.mark.parametrize("lod,filt", [(0.0, 1), (1.5, 1), (0.7, 0)])
def test_cube_map_program(renderer_factory, lod, filt):
    s = scenes.sphere_scene()
    s = dataclasses.replace(s, textures=[scenes.cube_texture(32)], lod=lod,
                            sampler_state=dict(filter=filt, address_u=2, address_v=2))
    r = renderer_factory()
    r.load_scene(s)
    _check(r, s, f"cube map lod {lod} filter {filt}")


def test_layered_upload_and_mips(renderer_factory):
    r = renderer_factory()
    r.set_target(16, 16, 1)
    size = 32
    faces = [scenes.fbm_texture(size, 90 + k) for k in range(6)]
    levels = scenes.num_levels(size, size)
    tid = r.L.iftrCreateTextureLayers(r.h, size, size, levels, 6)
    r._check()
    for k, f in enumerate(faces):
        a = np.ascontiguousarray(f)
        r.L.iftrCopyHostBufferToImageLayer(r.h, tid, k, a.ctypes.data)
    r.L.iftrGenerateMipmap(r.h, tid)
    r._check()
    got = r.download_texture(tid, scenes.Texture(size, size, levels, None, None, layers=6))
    assert np.array_equal(got, scenes.layered_mips(faces))


@pytest.mark.parametrize("linear", [True, False])
def test_blit_matches_oracle(renderer_factory, linear):
    r = renderer_factory()
    r.set_target(16, 16, 1)
    rng = np.random.default_rng(8 + linear)
    src0 = rng.integers(0, 256, (64, 64, 4), dtype=np.uint8)
    src = scenes.Texture(64, 64, 7, scenes.build_mips(src0))
    dst = scenes.Texture(32, 32, 6, scenes.build_mips(np.zeros((32, 32, 4), np.uint8)))
    si, di = r.add_texture(src), r.add_texture(dst)
    cur = np.zeros((32, 32, 4), np.uint8)
    for sreg, dreg, sl in [((0, 0, 64, 64), (0, 0, 32, 32), 0), ((5, 9, 37, 50), (3, 2, 30, 19), 0),
                           ((1, 2, 31, 32), (0, 20, 32, 32), 1), ((0, 0, 7, 3), (10, 10, 31, 31), 2)]:
        level = src0 if sl == 0 else scenes.build_mips(src0)[
            sum(4 * (64 >> k) ** 2 for k in range(sl)):sum(4 * (64 >> k) ** 2 for k in range(sl + 1))
        ].reshape(64 >> sl, 64 >> sl, 4)
        r.blit(si, di, sreg, dreg, src_level=sl, linear=linear)
        cur = O.blit(level, cur, sreg, dreg, linear=linear)
        got = r.download_texture(di, dst)[:32 * 32 * 4].reshape(32, 32, 4)
        bad = int((got != cur).any(-1).sum())
        print(f"[blit linear={linear} {sreg} level {sl} -> {dreg}] mismatching texels {bad}", flush=True)
        assert bad == 0
