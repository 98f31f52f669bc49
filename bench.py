"""Benchmark: one frame of the tile rasterization pipeline on one MI355X.

Workload (the headline configuration): 1920x1080 with 8x MSAA, 256 seeded
icospheres at subdivision level 4 (1,310,720 triangles, 655,872 vertices),
one seeded 2048^2 RGBA8 fBm texture with a full mip chain per sphere, 16x
anisotropic filtering with quad derivatives, Blinn-Phong with 4 point lights,
RGBA8 resolve.  A step is one full frame: vertex transform, setup/clip,
binning, coverage and early Z for all 16.6M samples, shading and resolve,
with the scene already resident in HBM.

Prints one JSON line.  value = fragments per second (Gfrag/s, fragment =
(pixel, triangle) pair with at least one covered sample); ms_per_step is the
frame time.  roofline covers the dominant kernel, measured with HIP events on
the renderer's stream; cpu_baseline times the reference CPU rasterizer
(compiled from its own sources) on the same geometry on this host's cores.
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
STAGES = ["vertex", "setup", "bin", "raster", "raster_merge", "shade"]


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
        tex = 4 * W * H * 4 / 3           # one RGBA8 texel per pixel with its mip chain
        return W * H * 4 + tex            # RGBA8 resolve out
    return 0.0


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


def cpu_baseline(scene, frames: int = 5):
    """Reference CPU tile rasterizer (oracle/_ref) when built, else the oracle port."""
    import ctypes as C

    import numpy as np

    from oracle import oracle as O

    ncpu = os.cpu_count() or 1
    if O.ref_available():
        v = np.ascontiguousarray(scene.vertices, np.float32)
        i = np.ascontiguousarray(scene.indices, np.int32)
        m = np.ascontiguousarray(scene.mvp, np.float32)
        # the reference library logs to stdout; keep this process's stdout to the one JSON line
        sys.stdout.flush()
        saved = os.dup(1)
        os.dup2(2, 1)
        try:
            ms = O.ref().ref_render_timed(scene.width, scene.height, v.ctypes.data_as(C.c_void_p), v.shape[1],
                                          v.shape[0], i.ctypes.data_as(C.c_void_p), i.size,
                                          m.ctypes.data_as(C.c_void_p), frames, None)
            C.CDLL(None).fflush(None)
        finally:
            os.dup2(saved, 1)
            os.close(saved)
        return {"value": round(1000.0 / ms, 3), "unit": "frames/s", "ms_per_frame": round(ms, 3), "cores": 17,
                "kind": "reference",
                "sample": f"{frames} frames (median) of the same 1.31M-triangle geometry at {scene.width}x{scene.height} "
                          "with 1 sample and untextured Lambert shading: the reference CPU path (16 workers + caller, "
                          f"host nproc={ncpu}) has no MSAA or texture sampling"}
    import numpy as np  # noqa: F811

    clip = O.vertex_transform(scene.vertices, scene.mvp)
    t = []
    for _ in range(frames):
        t0 = time.perf_counter()
        O.Frame().render(scene.width, scene.height, scene.samples, clip, scene.indices, threads=ncpu)
        t.append(time.perf_counter() - t0)
    ms = sorted(t)[len(t) // 2] * 1000
    return {"value": round(1000.0 / ms, 3), "unit": "frames/s", "ms_per_frame": round(ms, 3), "cores": ncpu,
            "kind": "port", "sample": f"{frames} frames of visibility (no shading) at the full workload size"}


def neural_leg(args):
    """Config 5: layered-GGX material vs the neural material on the config-2
    geometry at 1080p, 1 sample per pixel.  Shade-pass time from HIP events;
    PSNR and max error of the neural frames against the layered-GGX frame of
    the same visibility (that frame is within one RGBA8 step of the FP64
    oracle, which the GPU tests check)."""
    import numpy as np

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
    plain = r.color()
    ids, _ = r.visibility()
    mask = (ids != 0xFFFFFFFF).reshape(scene.height, scene.width)
    rep = r.train_neural(views=args.neural_views, view_size=(256, 256), max_fragments=16 << 20,
                         steps=args.neural_steps, batch=65536, latent_size=1024, single_tap=args.neural_single_tap)
    out = {"workload": "config5: config-2 geometry, 1920x1080, 1 spp, layered GGX (3 x 2048^2 layers, 16x aniso) "
                       "vs 8-ch FP16 1024^2 latent + MLP 14-64-64-3",
           "plain_shade_ms": round(plain_ms, 4), "train": {k: (round(v, 6) if isinstance(v, float) else v)
                                                            for k, v in rep.items()},
           "psnr_reference": "layered-GGX frame on the GPU (within 1 RGBA8 step of the FP64 oracle)"}
    out["psnr_bound_db"] = 40.0
    for prec, name in ((0, "bf16"), (1, "fp8")):
        r.bind_fragment_program(scene, program=4, precision=prec)
        r.draw()
        ms = shade_ms()
        img = r.color()
        d = (img[mask][:, :3].astype(np.float64) - plain[mask][:, :3].astype(np.float64)) / 255.0
        # the library's validation: PSNR against the material frame; below the bound the neural
        # program at this precision falls back to the material from here on
        psnr = r.validate_neural(prec, out["psnr_bound_db"])
        fallback = psnr < out["psnr_bound_db"]
        out[name] = {"shade_ms": round(ms, 4), "psnr_db": round(psnr, 2),
                     "max_err_lsb": int(round(np.abs(d).max() * 255)), "speedup_vs_plain": round(plain_ms / ms, 3),
                     "falls_back_to_plain": bool(fallback)}
    r.close()
    return out


def batch_leg(args):
    """Config 4: 64 views of the config-2 scene in one draw at 1 sample per pixel: 4 shadow
    cascades at 4096^2 and 6 cube faces at 1024^2 (depth only, FP32), 54 seeded cameras of
    uneven size (shaded).  Every stage is one launch over all views and the draw is one
    hipGraph replay.  The autotuner first picks the launch parameters (it checks that every
    configuration produces the same digest of depth, IDs and colour)."""
    import torch

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
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.batch_steps):
        r.draw_async()
    r.synchronize()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1000.0 / args.batch_steps
    if r.stats()["status"]:
        raise SystemExit("batch incomplete during timing")
    # per-stage device time (direct launches with events between the stages)
    r.set_graph_mode(False)
    r.stage_timing(True)
    acc = [0.0] * len(STAGES)
    for _ in range(3):
        r.draw()
        for k, v in enumerate(r.stats()["stageMs"][:len(STAGES)]):
            acc[k] += v / 3
    r.stage_timing(False)
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
            "stage_ms": {STAGES[k]: round(acc[k], 4) for k in range(len(STAGES))},
            "autotune": {"best": tune["best"], "default_ms": round(tune["default_ms"], 4),
                         "best_ms": round(tune["best_ms"], 4), "candidates": tune["candidates"],
                         "identical": tune["identical"], "bin_demand": tune["bin_demand"],
                         "seconds": round(tune_s, 2)},
            "roofline": {"bound": "hbm", "algorithmic_bytes": round(ab), "achieved": round(achieved, 2),
                         "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 5)}}


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
    ap.add_argument("--no-batch", action="store_true", help="skip the multi-view batch leg")
    ap.add_argument("--only", choices=["batch", "neural"],
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
    if args.only:
        print(json.dumps({args.only: batch_leg(args) if args.only == "batch" else neural_leg(args)}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the renderer runs on a single "
              "MI355X; multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    import torch

    from ifrit_hip import Renderer, scenes

    scene = scenes.config2(spheres=args.spheres)
    r = Renderer()
    t0 = time.time()
    r.load_scene(scene)
    setup_s = time.time() - t0
    for _ in range(args.warmup):
        r.draw()
    st = r.stats()

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r.draw_async()
    r.synchronize()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1000.0 / args.steps
    st_timed = r.stats()
    if st_timed["status"]:
        raise SystemExit(f"frame incomplete during timing: status {st_timed['status']:#x}")

    # per-stage device time with HIP events on the renderer's stream
    r.stage_timing(True)
    acc = [0.0] * len(STAGES)
    nt = 10
    for _ in range(nt):
        r.draw()
        s = r.stats()
        for k in range(len(STAGES)):
            acc[k] += s["stageMs"][k]
    r.stage_timing(False)
    stage_ms = {STAGES[k]: acc[k] / nt for k in range(len(STAGES))}
    timed = dict(stage_ms)
    timed["raster"] = stage_ms["raster"] + stage_ms["raster_merge"]
    dom = max(("raster", "shade", "setup", "bin", "vertex"), key=lambda k: timed[k])
    ab = algorithmic_bytes(scene, dom)
    achieved = ab / (timed[dom] * 1e-3) / 1e9 if timed[dom] > 0 else 0.0
    traffic = measured_traffic(dom)

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
                   "scene_setup_s": round(setup_s, 2)},
        "roofline": {"kernel": dom, "kernels": STAGE_KERNELS[dom], "bound": "hbm", "achieved": round(achieved, 2),
                     "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 5),
                     "traffic": traffic, "algorithmic_bytes": ab, "duration_ms": round(timed[dom], 4)},
    }
    if not args.no_batch:
        out["batch"] = batch_leg(args)
    if not args.no_neural:
        out["neural"] = neural_leg(args)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(scene)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
