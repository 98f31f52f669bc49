"""Benchmark: one frame of the tile rasterization pipeline on one MI355X.

Workload (the headline configuration, config 2): 1920x1080 with 8x MSAA, 256
seeded icospheres at subdivision level 4 (1,310,720 triangles, 655,872
vertices), one seeded 2048^2 RGBA8 fBm texture with a full mip chain per
sphere, 16x anisotropic filtering with quad derivatives, Blinn-Phong with 4
point lights, RGBA8 resolve.  A step is one full frame: vertex transform,
setup/clip, binning, coverage and early Z for all 16.6M samples, shading and
resolve, with the scene already resident in HBM.

Prints one JSON line.  value = fragments per second (Gfrag/s, fragment =
(pixel, triangle) pair with at least one covered sample); ms_per_step is the
frame time.  roofline covers the dominant kernel, measured with HIP events on
the renderer's stream.  Further legs: config 1 and config 3 frames, the
64-view batch (config 4), the neural material (config 5), and the reference
CPU path's own workload drawn on the GPU beside the reference itself
(cpu_baseline, timed on this host's cores).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ifrit-v2_amd"))
sys.path.insert(0, ROOT)

METRIC = "ms/frame & Gfrag/s at 1080p 8xMSAA; neural vs plain shade ms at ≥40 dB PSNR"
PEAK_HBM_GBS = 8000.0
PEAK_BF16_TFLOPS = 2500.0
PEAK_FP8_TFLOPS = 5000.0
STAGES = ["vertex", "setup", "bin", "raster", "raster_merge", "shade"]
NEURAL_FLOP_PER_FRAGMENT = 2 * (14 * 64 + 64 * 64 + 64 * 3)  # 14-64-64-3 MLP: 10,368


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def algorithmic_bytes(scene, stage: str) -> float:
    """Compulsory HBM bytes of one launch of each stage: positions in and clip positions out, indices and
    clip positions in, per-sample depth and ID out, RGBA8 out plus one texel per pixel with its mips."""
    W, H, S = scene.width, scene.height, scene.samples
    V, T = scene.vertices.shape[0], scene.num_triangles
    if stage == "vertex":
        return V * 12 + V * 16            # positions in, clip positions out
    if stage == "setup":
        return T * 12 + 3 * T * 16        # indices and clip positions in
    if stage == "raster":
        return W * H * S * 8              # per-sample depth + primitive id out
    if stage == "shade":
        tex = 4 * W * H * 4 / 3 if scene.textures else 0.0  # one RGBA8 texel per pixel with its mip chain
        return W * H * 4 + tex            # RGBA8 resolve out
    return 0.0


def frame_bytes(scene) -> float:
    """Compulsory bytes of a whole frame: B = V s_v + 12 T + W H S 8 + 4 W H + tau (tau: one RGBA8 texel
    per pixel with its mips when the frame samples a texture)."""
    W, H, S = scene.width, scene.height, scene.samples
    V, T = scene.vertices.shape[0], scene.num_triangles
    tau = 4 * W * H * 4 / 3 if scene.textures else 0.0
    return V * scene.stride * 4 + 12 * T + W * H * S * 8 + 4 * W * H + tau


# kernels of each timed stage (the raster stage's outputs are written by both
# k_raster and, for tiles whose list was cut into items, k_raster_merge)
STAGE_KERNELS = {"vertex": ["k_vertex"], "setup": ["k_setup", "k_clip"],
                 "bin": ["k_bin_small", "k_bin_big_count", "k_plan_sums", "k_plan_scan", "k_plan_write"],
                 "raster": ["k_raster", "k_raster_merge"], "shade": ["k_shade"]}


def measured_traffic(stage: str):
    """HBM bytes per launch of the stage's kernels from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (scaled by a copy of known size), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            per = json.load(f)["bytes_per_launch"]
    except (OSError, ValueError, KeyError):
        return None
    total = 0
    for k in STAGE_KERNELS.get(stage, []):
        hits = [v for name, v in per.items() if name == k]
        if not hits:
            return None
        total += hits[0]
    return total


def host_cpu() -> dict:
    """The host's CPU model, vendor and the CPUs this process may run on."""
    model = vendor = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name") and model == "unknown":
                    model = line.split(":", 1)[1].strip()
                elif line.startswith("vendor_id") and vendor == "unknown":
                    vendor = line.split(":", 1)[1].strip()
    except OSError:
        pass
    try:
        cpus = len(os.sched_getaffinity(0))
    except AttributeError:
        cpus = os.cpu_count() or 1
    return {"cpu_model": model, "cpu_vendor": vendor, "cpus": cpus, "nproc": os.cpu_count() or 1}


def time_frames(r, steps):
    """Mean ms of `steps` asynchronous frames on the renderer's stream, bracketed by synchronisation
    of that stream (the renderer runs nothing on any other)."""
    r.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        r.draw_async()
    r.synchronize()
    return (time.perf_counter() - t0) * 1000.0 / steps


def stage_times(r, frames=10):
    r.stage_timing(True)
    acc = [0.0] * len(STAGES)
    for _ in range(frames):
        r.draw()
        s = r.stats()
        for k in range(len(STAGES)):
            acc[k] += s["stageMs"][k] / frames
    r.stage_timing(False)
    return {STAGES[k]: acc[k] for k in range(len(STAGES))}


def frame_leg(scene, steps, label):
    """One configuration's frame: ms/frame, triangles/s, fragments/s, stage times, the dominant stage's
    roofline and the whole frame's roofline on its compulsory bytes."""
    from ifrit_hip import Renderer

    r = Renderer()
    r.load_scene(scene)
    for _ in range(3):
        r.draw()
    st = r.stats()
    ms = time_frames(r, steps)
    stg = stage_times(r)
    r.close()
    timed = dict(stg)
    timed["raster"] = stg["raster"] + stg["raster_merge"]
    dom = max(("raster", "shade", "setup", "bin", "vertex"), key=lambda k: timed[k])
    ab = algorithmic_bytes(scene, dom)
    fb = frame_bytes(scene)
    return {"workload": label, "width": scene.width, "height": scene.height, "samples": scene.samples,
            "triangles": scene.num_triangles, "steps": steps, "ms_per_frame": round(ms, 4),
            "triangles_per_s": round(scene.num_triangles / (ms * 1e-3), 1),
            "fragments": st["fragments"], "fragments_per_s": round(st["fragments"] / (ms * 1e-3), 1),
            "samples_per_s": round(scene.width * scene.height * scene.samples / (ms * 1e-3), 1),
            "stage_ms": {k: round(v, 4) for k, v in stg.items()},
            "roofline": {"kernel": dom, "bound": "hbm", "algorithmic_bytes": ab,
                         "achieved": round(ab / (timed[dom] * 1e-3) / 1e9, 2), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                         "frac": round(ab / (timed[dom] * 1e-3) / 1e9 / PEAK_HBM_GBS, 5),
                         "duration_ms": round(timed[dom], 4)},
            "frame_roofline": {"algorithmic_bytes": round(fb), "achieved": round(fb / (ms * 1e-3) / 1e9, 2),
                               "peak": PEAK_HBM_GBS, "unit": "GB/s",
                               "frac": round(fb / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 5)}}


def reference_workload_gpu(args):
    """The reference CPU path's own timing workload on the GPU: config-2 geometry, 1920x1080, 1 sample
    per pixel, the normal-Lambert program, FP32 colour written (as the reference writes its FP32
    colour attachment).  Device-resident frames, and blocking frames that also copy the FP32 colour
    to host memory (the drop-in call sequence)."""
    import numpy as np

    from ifrit_hip import IfritError, Renderer, scenes

    s = scenes.reference_workload()
    r = Renderer()
    r.load_scene(s)
    r.set_float_output(True)
    # the setup reproduces this host CPU's RCPPS, so the frame is the one the reference build draws here
    try:
        r.set_setup_reciprocal("host")
        rcp = "host"
    except IfritError:
        rcp = "intel"
    for _ in range(3):
        r.draw()
    st = r.stats()
    ms = time_frames(r, args.steps)
    out = np.empty((s.height, s.width, 4), np.float32)
    t = []
    for _ in range(11):
        t0 = time.perf_counter()
        r.draw()
        out[...] = r.color_f32()
        t.append((time.perf_counter() - t0) * 1000.0)
    r.close()
    return s, {"workload": "reference CPU timing workload: config-2 geometry (1,310,720 tris), 1920x1080, 1 spp, "
                           "normal-Lambert shading, FP32 colour",
               "ms_per_frame": round(ms, 4), "fragments": st["fragments"],
               "ms_per_frame_with_fp32_readback": round(sorted(t)[len(t) // 2], 4), "setup_reciprocal": rcp}


def neural_leg(args):
    """Config 5: layered-GGX material vs the neural material on the config-2 geometry at 1080p, 1 sample
    per pixel.  Shade-pass time from HIP events; the library's validation (PSNR, largest and
    99.9th-percentile per-pixel error against the material frame, which the GPU tests hold within one
    RGBA8 step of the FP64 oracle).  Returns the frames for the FP64 comparison of the cpu_baseline leg."""
    from ifrit_hip import Renderer, scenes

    scene = scenes.config5(spheres=args.spheres)
    r = Renderer()
    r.load_scene(scene)
    r.stage_timing(True)

    def shade_ms(n=10):
        acc = 0.0
        for _ in range(n):
            r.draw()
            acc += r.stats()["stageMs"][STAGES.index("shade")]
        return acc / n

    r.draw()
    plain_ms = shade_ms()
    frames = {"material": r.color()}
    ids, _ = r.visibility()
    covered = int((ids != 0xFFFFFFFF).sum())
    textures = [scenes.Texture(t.width, t.height, t.levels, r.download_texture(i, t))
                for i, t in enumerate(scene.textures)]
    rep = r.train_neural(views=args.neural_views, view_size=(256, 256), max_fragments=16 << 20,
                         steps=args.neural_steps, batch=65536, latent_size=1024, single_tap=args.neural_single_tap)
    out = {"workload": "config5: config-2 geometry, 1920x1080, 1 spp, layered GGX (3 x 2048^2 layers, 16x aniso) "
                       "vs 8-ch FP16 1024^2 latent + MLP 14-64-64-3",
           "plain_shade_ms": round(plain_ms, 4), "covered_pixels": covered,
           "train": {k: (round(v, 6) if isinstance(v, float) else v) for k, v in rep.items()},
           "psnr_bound_db": 40.0, "max_err_bound_lsb": args.neural_max_err}
    W, H = scene.width, scene.height
    hbm_bytes = W * H * (16 * 4 / 3 + 8 + 4)  # latent texel with mips, inputs, RGBA8 out per pixel
    flop = covered * NEURAL_FLOP_PER_FRAGMENT
    for prec, name in ((0, "bf16"), (1, "fp8")):
        r.bind_fragment_program(scene, program=4, precision=prec)
        r.draw()
        ms = shade_ms()
        frames[name] = r.color()
        v = r.validate_neural_ex(prec, out["psnr_bound_db"], args.neural_max_err)
        # the FP8 network runs its hidden layer (64 x 64 of the 14-64-64-3 MACs) on FP8 MFMA and the
        # input and output layers on BF16: its peak is the FLOP-weighted harmonic mean of the two
        hidden = 64 * 64 / (14 * 64 + 64 * 64 + 64 * 3)
        peak = PEAK_BF16_TFLOPS if prec == 0 else 1.0 / (hidden / PEAK_FP8_TFLOPS + (1 - hidden) / PEAK_BF16_TFLOPS)
        out[name] = {"shade_ms": round(ms, 4), "speedup_vs_plain": round(plain_ms / ms, 3),
                     "vs_material": {"psnr_db": round(v["psnrDb"], 2), "max_err_lsb": v["maxErrLsb"],
                                     "p999_err_lsb": v["p999ErrLsb"]},
                     "falls_back_to_plain": bool(v["fallback"]),
                     "roofline": {"hbm": {"algorithmic_bytes": round(hbm_bytes),
                                          "achieved": round(hbm_bytes / (ms * 1e-3) / 1e9, 2), "peak": PEAK_HBM_GBS,
                                          "unit": "GB/s", "frac": round(hbm_bytes / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 5)},
                                  "mfma": {"flop": flop, "achieved": round(flop / (ms * 1e-3) / 1e12, 2), "peak": round(peak, 1),
                                           "unit": "TFLOP/s", "frac": round(flop / (ms * 1e-3) / 1e12 / peak, 5)}}}
    r.close()
    return out, (scene, textures, frames)


def batch_leg(args):
    """Config 4: 64 views of the config-2 scene in one draw at 1 sample per pixel: 4 shadow
    cascades at 4096^2 and 6 cube faces at 1024^2 (depth only, FP32), 54 seeded cameras of
    uneven size (shaded).  Every stage is one launch over all views and the draw is one
    hipGraph replay.  The autotuner first picks the launch parameters (it checks that every
    configuration produces the same digest of depth, IDs and colour)."""
    from ifrit_hip import Renderer, scenes

    scene = scenes.config2(spheres=args.spheres, samples=1)
    views = scenes.config4_views()
    r = Renderer()
    r.load_scene(scene)
    r.set_views(views)
    r.set_graph_mode(True)
    r.draw()
    t0 = time.perf_counter()
    tune = r.autotune(frames=3)
    tune_s = time.perf_counter() - t0
    for _ in range(2):
        r.draw()
    st = r.stats()
    ms = time_frames(r, args.batch_steps)
    # per-stage device time (direct launches with events between the stages)
    r.set_graph_mode(False)
    acc = stage_times(r, 3)
    pixels = sum(v["width"] * v["height"] for v in views)
    shaded = sum(v["width"] * v["height"] for v in views if v["shade"])
    V, T = scene.vertices.shape[0], scene.num_triangles
    # compulsory bytes: vertices and indices once, depth + ID per pixel of every view, RGBA8 and one
    # texel with its mips per shaded pixel
    ab = V * 32 + T * 12 + pixels * 8 + shaded * 4 + shaded * 4 * 4 / 3
    achieved = ab / (ms * 1e-3) / 1e9
    r.close()
    return {"workload": "config4: 64 views of the config-2 scene, 1 spp: 4 cascades 4096^2 + 6 cube faces "
                        "1024^2 depth-only, 54 seeded cameras 256..1920 x 256..1080 shaded (Blinn-Phong, 16x aniso)",
            "views": len(views), "pixels": pixels, "shaded_pixels": shaded, "steps": args.batch_steps,
            "ms_per_batch": round(ms, 4), "gpixels_per_s": round(pixels / (ms * 1e-3) / 1e9, 3),
            "triangles_per_s": round(len(views) * T / (ms * 1e-3), 1),
            "fragments_per_s": round(st["fragments"] / (ms * 1e-3), 1), "fragments": st["fragments"],
            "bin_entries": st["binEntries"], "clipped_triangles": st["clippedTriangles"],
            "stage_ms": {k: round(v, 4) for k, v in acc.items()},
            "autotune": {"best": tune["best"], "default_ms": round(tune["default_ms"], 4),
                         "best_ms": round(tune["best_ms"], 4), "candidates": tune["candidates"],
                         "identical": tune["identical"], "bin_demand": tune["bin_demand"],
                         "seconds": round(tune_s, 2)},
            "roofline": {"bound": "hbm", "algorithmic_bytes": round(ab), "achieved": round(achieved, 2),
                         "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 5)}}


def _errors(img, ref, mask):
    import numpy as np

    d = np.abs(img[mask][:, :3].astype(np.int64) - ref[mask][:, :3].astype(np.int64))
    mse = ((d / 255.0) ** 2).mean()
    worst = d.max(-1)
    return {"psnr_db": round(float(10 * np.log10(1.0 / max(mse, 1e-12))), 2), "max_err_lsb": int(worst.max()),
            "p999_err_lsb": int(np.quantile(worst, 0.999))}


def cpu_baseline(ref_scene, ref_gpu, neural_frames, frames: int = 11):
    """CPU legs (the only place this script touches oracle/):
    * the reference CPU tile rasterizer (oracle/_ref, built from the reference's own sources) on its own
      timing workload -- the frame reference_workload_gpu draws on the GPU -- with its 16 workers and the
      calling thread, median of `frames` frames; the same frame's GPU times beside it;
    * the FP64 shading oracle on the config-5 frame (visibility and shading, all host threads), used as
      the checker of the neural frames: PSNR, largest and 99.9th-percentile error against it."""
    import ctypes as C

    import numpy as np

    from oracle import oracle as O

    cpu = host_cpu()
    out = {}
    if O.ref_available():
        s = ref_scene
        v = np.ascontiguousarray(s.vertices, np.float32)
        i = np.ascontiguousarray(s.indices, np.int32)
        m = np.ascontiguousarray(s.mvp, np.float32)
        # the reference library logs to stdout; keep this process's stdout to the one JSON line
        sys.stdout.flush()
        saved = os.dup(1)
        os.dup2(2, 1)
        try:
            ms = O.ref().ref_render_timed(s.width, s.height, v.ctypes.data_as(C.c_void_p), v.shape[1], v.shape[0],
                                          i.ctypes.data_as(C.c_void_p), i.size, m.ctypes.data_as(C.c_void_p), frames,
                                          None)
            C.CDLL(None).fflush(None)
        finally:
            os.dup2(saved, 1)
            os.close(saved)
        out = {"value": round(1000.0 / ms, 3), "unit": "frames/s", "ms_per_frame": round(ms, 3),
               "cores": cpu["cpus"], "threads": 17, "cpu_model": cpu["cpu_model"], "cpu_vendor": cpu["cpu_vendor"],
               "kind": "reference",
               "sample": f"median of {frames} frames of the reference CPU rasterizer (built from its own sources, 16 "
                         f"workers + the calling thread on {cpu['cpus']} host CPUs) drawing the config-2 geometry "
                         "(1,310,720 tris) at 1920x1080, 1 spp, with its timing harness's normal-Lambert shader "
                         "into FP32 colour and depth attachments; the GPU draws the same frame with the "
                         "normal-Lambert program (tests/test_gpu_reference_workload.py checks it against this build)",
               "gpu_same_workload_ms": ref_gpu["ms_per_frame"],
               "gpu_same_workload_with_readback_ms": ref_gpu["ms_per_frame_with_fp32_readback"],
               "speedup_device_resident": round(ms / ref_gpu["ms_per_frame"], 1),
               "speedup_with_fp32_readback": round(ms / ref_gpu["ms_per_frame_with_fp32_readback"], 1)}
        out["note"] = ("the reference's triangle setup uses RCPPS, whose results are vendor-specific; the GPU leg "
                       f"uses the setup reciprocal model '{ref_gpu['setup_reciprocal']}' ('host': this CPU's "
                       "table), so both legs draw the same frame")
    if neural_frames is not None:
        scene, textures, imgs = neural_frames
        import dataclasses

        t0 = time.perf_counter()
        f, _ = O.render_scene(scene)
        ref, flags = f.shade(dataclasses.replace(scene, textures=textures))
        oms = (time.perf_counter() - t0) * 1000.0
        mask = (f.ids != 0xFFFFFFFF).reshape(scene.height, scene.width)
        out["fp64_oracle_config5"] = {"ms_per_frame": round(oms, 1), "threads": cpu["nproc"], "kind": "port",
                                      "sample": "one config-5 frame: the oracle's visibility and FP64 layered-GGX "
                                                "shading with 16x anisotropic filtering",
                                      "material_vs_oracle": _errors(imgs["material"], ref, mask),
                                      "bf16_vs_oracle": _errors(imgs["bf16"], ref, mask),
                                      "fp8_vs_oracle": _errors(imgs["fp8"], ref, mask)}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--spheres", type=int, default=256)
    ap.add_argument("--no-neural", action="store_true", help="skip the neural-material leg")
    ap.add_argument("--neural-steps", type=int, default=3000)
    ap.add_argument("--neural-views", type=int, default=512)
    ap.add_argument("--neural-single-tap", type=int, default=1,
                    help="1: one trilinear latent tap per fragment, 0: the material's anisotropic taps")
    ap.add_argument("--neural-max-err", type=int, default=255,
                    help="per-pixel fallback bound of the neural validation, RGBA8 steps (255: PSNR only)")
    ap.add_argument("--no-batch", action="store_true", help="skip the multi-view batch leg")
    ap.add_argument("--no-configs", action="store_true", help="skip the config 1 and config 3 legs")
    ap.add_argument("--only", choices=["batch", "neural", "c1", "c3", "refwork"],
                    help="run one leg alone and print its JSON (for profiling; not the benchmark line)")
    ap.add_argument("--batch-steps", type=int, default=5)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a float32 copy of 256 MiB (x + 0) three times: the reference kernel for "
                         "scaling FETCH_SIZE / WRITE_SIZE counters to bytes")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch

        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.add(src, 0.0, out=dst)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "copy", "bytes_read": src.numel() * 4, "bytes_written": dst.numel() * 4}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the renderer runs on a single "
              "MI355X; multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)
    from ifrit_hip import Renderer, scenes

    if args.only:
        if args.only == "batch":
            res = batch_leg(args)
        elif args.only == "neural":
            res = neural_leg(args)[0]
        elif args.only == "c1":
            res = frame_leg(scenes.config1(), args.steps, "config1")
        elif args.only == "c3":
            res = frame_leg(scenes.config3(), args.steps, "config3")
        else:
            res = reference_workload_gpu(args)[1]
        print(json.dumps({args.only: res}), flush=True)
        return

    log("config 2 frame")
    scene = scenes.config2(spheres=args.spheres)
    r = Renderer()
    t0 = time.time()
    r.load_scene(scene)
    setup_s = time.time() - t0
    for _ in range(args.warmup):
        r.draw()
    st = r.stats()
    ms = time_frames(r, args.steps)
    # per-stage device time with HIP events on the renderer's stream
    stage_ms = stage_times(r, 10)
    r.close()
    timed = dict(stage_ms)
    timed["raster"] = stage_ms["raster"] + stage_ms["raster_merge"]
    dom = max(("raster", "shade", "setup", "bin", "vertex"), key=lambda k: timed[k])
    ab = algorithmic_bytes(scene, dom)
    achieved = ab / (timed[dom] * 1e-3) / 1e9 if timed[dom] > 0 else 0.0
    traffic = measured_traffic(dom)
    fb = frame_bytes(scene)

    frags = st["fragments"]
    value = frags / (ms * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "Gfrag/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scene generator, procedural fBm textures)",
        "config": {"workload": "config2: 1920x1080 8xMSAA, 256 icospheres L4 (1,310,720 tris), 2048^2 fBm "
                               "textures with mips, 16x aniso, Blinn-Phong 4 point lights",
                   "width": scene.width, "height": scene.height, "samples": scene.samples,
                   "triangles": scene.num_triangles, "vertices": int(scene.vertices.shape[0]),
                   "fragments": frags, "covered_samples": st["coveredSamples"],
                   "assembled_triangles": st["assembledTriangles"], "bin_entries": st["binEntries"],
                   "triangles_per_s": round(scene.num_triangles / (ms * 1e-3), 1),
                   "stage_ms": {k: round(v, 4) for k, v in stage_ms.items()},
                   "frame_roofline": {"algorithmic_bytes": round(fb), "achieved": round(fb / (ms * 1e-3) / 1e9, 2),
                                      "peak": PEAK_HBM_GBS, "unit": "GB/s",
                                      "frac": round(fb / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 5)},
                   "scene_setup_s": round(setup_s, 2)},
        "roofline": {"kernel": dom, "kernels": STAGE_KERNELS[dom], "bound": "hbm", "achieved": round(achieved, 2),
                     "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 5),
                     "traffic": traffic, "algorithmic_bytes": ab, "duration_ms": round(timed[dom], 4)},
    }
    if not args.no_configs:
        log("config 1 and config 3 frames")
        out["config1"] = frame_leg(scenes.config1(), 200, "config1: 640x480 1 spp, UV sphere 64x32 + ground quad, "
                                                         "512^2 checker trilinear, Blinn-Phong 1 directional light")
        out["config3"] = frame_leg(scenes.config3(), args.steps, "config3: 1920x1080 4xMSAA, 2048x1024 fBm "
                                                                "heightfield (4,194,304 tris) + 16 large triangles, "
                                                                "flat colour")
    if not args.no_batch:
        log("config 4 batch")
        out["batch"] = batch_leg(args)
    neural_frames = None
    if not args.no_neural:
        log("config 5 neural material")
        out["neural"], neural_frames = neural_leg(args)
    if not args.no_cpu:
        log("reference workload on the GPU, then the CPU legs")
        ref_scene, ref_gpu = reference_workload_gpu(args)
        out["reference_workload_gpu"] = ref_gpu
        out["cpu_baseline"] = cpu_baseline(ref_scene, ref_gpu, neural_frames)
        if "fp64_oracle_config5" in out["cpu_baseline"] and "neural" in out:
            o = out["cpu_baseline"]["fp64_oracle_config5"]
            out["neural"]["material_vs_fp64_oracle"] = o["material_vs_oracle"]
            out["neural"]["bf16"]["vs_fp64_oracle"] = o["bf16_vs_oracle"]
            out["neural"]["fp8"]["vs_fp64_oracle"] = o["fp8_vs_oracle"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
