"""Neural material on the GPU.

* The MFMA network (BF16 and FP8 operands, FP32 accumulation) against a plain
  FP32 numpy evaluation of the same weights: a wrong operand layout, a
  permutation mismatch between layers or a missing bias gives errors of the
  order of the outputs themselves, far above the stated tolerances.
* A short training run on a small layered-GGX scene: the loss falls, and the
  neural frame approaches the FP64 oracle of the layered-GGX material (PSNR
  over covered pixels).
"""
import numpy as np
import pytest

from ifrit_hip import NEURAL_PARAMS, scenes
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _random_params(rng):
    p = np.zeros(NEURAL_PARAMS, np.float32)
    w1 = rng.uniform(-1, 1, (14, 64)) * np.sqrt(6 / 78)
    w2 = rng.uniform(-1, 1, (64, 64)) * np.sqrt(6 / 128)
    w3 = rng.uniform(-1, 1, (64, 3)) * np.sqrt(6 / 67)
    p[0:1024].reshape(16, 64)[:14] = w1
    p[1024:1088] = rng.uniform(-0.1, 0.1, 64)
    p[1088:5184] = w2.reshape(-1)
    p[5184:5248] = rng.uniform(-0.1, 0.1, 64)
    p[5248:5504].reshape(64, 4)[:, :3] = w3
    p[5504:5507] = rng.uniform(-0.1, 0.1, 3)
    return p


def _mlp_fp32(p, x):
    w1 = p[0:1024].reshape(16, 64)[:14]
    b1 = p[1024:1088]
    w2 = p[1088:5184].reshape(64, 64)
    b2 = p[5184:5248]
    w3 = p[5248:5504].reshape(64, 4)[:, :3]
    b3 = p[5504:5507]
    h1 = np.maximum(x @ w1 + b1, 0)
    h2 = np.maximum(h1 @ w2 + b2, 0)
    return h2 @ w3 + b3


@pytest.mark.parametrize("precision,tol", [(0, 0.03), (1, 0.2)])
def test_mlp_matches_fp32_reference(renderer_factory, precision, tol):
    rng = np.random.default_rng(5 + precision)
    r = renderer_factory()
    p = _random_params(rng)
    r.set_neural_params(p)
    assert np.array_equal(r.get_neural_params(), p)
    x = rng.uniform(-1, 1, (1000, 14)).astype(np.float32)  # not a multiple of 64: partial last wave
    got = r.eval_neural(x, precision)
    want = _mlp_fp32(p.astype(np.float64), x.astype(np.float64))
    err = np.abs(got - want).max()
    scale = np.abs(want).max()
    corr = np.corrcoef(got.reshape(-1), want.reshape(-1))[0, 1]
    print(f"[mlp precision={precision}] max |err| {err:.4f} max |y| {scale:.3f} corr {corr:.5f}", flush=True)
    assert err <= tol * scale
    assert corr > (0.999 if precision == 0 else 0.98)


def _psnr(a, b, mask):
    d = (a[mask][:, :3].astype(np.float64) - b[mask][:, :3].astype(np.float64)) / 255.0
    mse = (d * d).mean()
    return 10 * np.log10(1.0 / max(mse, 1e-12)), np.abs(d).max() * 255


def test_training_reduces_loss_and_approaches_the_material(renderer_factory):
    s = scenes.config5(width=320, height=180, samples=1, spheres=32, level=3, tex_size=256, host_textures=True)
    r = renderer_factory()
    r.load_scene(s)
    r.draw()
    plain = r.color()
    rep = r.train_neural(views=48, view_size=(128, 128), max_fragments=1 << 20, steps=400, batch=16384,
                         latent_size=128, seed=3)
    print(f"[neural train] {rep}", flush=True)
    assert rep["fragments"] > 10000
    assert rep["lossLast"] < 0.5 * rep["lossFirst"]
    f, _ = O.render_scene(s)
    ref, flags = f.shade(s)
    mask = (f.ids != 0xFFFFFFFF).reshape(s.height, s.width)
    ps_plain, mx_plain = _psnr(plain, ref, mask)
    out = {}
    for prec in (0, 1):
        r.bind_fragment_program(s, program=4, precision=prec)
        r.draw()
        out[prec] = _psnr(r.color(), ref, mask)
    print(f"[neural] PSNR plain {ps_plain:.1f} dB (max {mx_plain:.0f}); bf16 {out[0][0]:.1f} dB (max {out[0][1]:.0f});"
          f" fp8 {out[1][0]:.1f} dB (max {out[1][1]:.0f})", flush=True)
    assert ps_plain > 45
    assert out[0][0] > 20
    assert out[1][0] > 18


def test_validation_falls_back_below_the_bound(renderer_factory):
    """An untrained network is far from the material: validation reports a low
    PSNR and the neural program then draws the material itself."""
    s = scenes.config5(width=256, height=144, samples=1, spheres=16, level=3, tex_size=128, host_textures=True)
    r = renderer_factory()
    r.load_scene(s)
    r.draw()
    plain = r.color()
    r.set_neural_params(_random_params(np.random.default_rng(9)))
    r.bind_fragment_program(s, program=4, precision=0)
    psnr = r.validate_neural(0, 40.0)
    print(f"[neural fallback] untrained PSNR {psnr:.1f} dB", flush=True)
    assert psnr < 40.0
    r.draw()
    assert np.array_equal(r.color(), plain), "below the bound the neural program must draw the material"
    assert r.validate_neural(0, 0.0) == psnr  # a zero bound clears the fallback again
