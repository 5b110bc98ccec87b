"""Benchmark: path-traced frames of the 1M-triangle soup at 1920x1080, 8 spp, depth 8.

A step is one frame: RayGen, then per bounce Extend (closest-hit traversal), Shade
(integrator + wave-aggregated queue appends) and Shadow (any-hit traversal), then
Finalize, all on one MI355X with the scene already resident in HBM. `value` is the
whole-frame throughput: rays traced (extension + shadow) / wall time of the timed
frames. Traversal-only throughput (rays / wall time of the Extend and Shadow phases from
HIP events; Shadow of bounce d runs beside Extend of bounce d + 1 on a second stream) is
reported beside it, with the roofline of the Extend kernel computed from algorithmic
bytes (ray + hit records, node visits x node bytes, triangle tests x triangle bytes;
visit counts from a counting pass over the same frame).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="soup1m", choices=["soup1m", "cornell", "instanced"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-denoise", action="store_true",
                    help="skip the time-to-relMSE measurement (denoiser training + convergence runs)")
    ap.add_argument("--denoise-steps", type=int, default=4000)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--presplit", type=float, default=6.0,
                    help="BVH spatial pre-splitting budget, extra references per triangle (0 = off)")
    ap.add_argument("--copy-calibrate", action="store_true",
                    help="run a plain 1 GiB device copy (calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--denoise-only", action="store_true",
                    help="time only U-Net inference on a 1920x1080 1-spp Cornell frame (untrained weights; "
                         "--steps timed calls after --warmup)")
    return ap.parse_args()


def denoise_only(steps, warmup):
    """Inference time of the denoiser on one 1080p frame; the network's cost does not
    depend on its weights, so freshly initialised ones are used."""
    from trayracer_amd import Renderer, scenes
    rend = Renderer(0)
    rend.UpdateScene(scenes.cornell(width=1920, height=1080, spp=1))
    rend.RenderFrame()
    rend.denoiser_init(seed=0)
    for _ in range(warmup):
        rend.denoise()
    st = [rend.denoise() for _ in range(steps)]
    n = len(st)
    print(json.dumps({"workload": "U-Net denoise, 1920x1080", "steps": n, "warmup": warmup,
                      "denoise_ms": round(sum(s["total_ms"] for s in st) / n, 4),
                      "features_ms": round(sum(s["features_ms"] for s in st) / n, 4),
                      "network_ms": round(sum(s["network_ms"] for s in st) / n, 4)}), flush=True)
    rend.close()


def copy_calibrate():
    import torch
    n = 1 << 28  # 1 GiB of float32
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "launches": 3}), flush=True)


def make_scene(name):
    from trayracer_amd import scenes
    if name == "soup1m":
        return scenes.soup(), "soup-1M-tris 1920x1080 8spp depth8 (RR from depth 3)"
    if name == "cornell":
        return scenes.cornell(), "cornell-36-tris 256x256 16spp depth4"
    return scenes.instanced(), "instanced-1000x20480-tris 3840x2160 4spp depth6"


def cpu_baseline(scene, threads, spp=2):
    """Oracle (host build of the same per-vertex math) on a bounded sample of the same
    workload: the same camera and resolution at `spp` samples per pixel (8 in the GPU
    frame), same depth, about 10 s of host work on 16 threads."""
    from oracle import Oracle
    t0 = time.perf_counter()
    orc = Oracle(scene)
    build_s = time.perf_counter() - t0
    r = dict(scene["render"])
    W, H = r["width"], r["height"]
    t0 = time.perf_counter()
    out = orc.render(render=dict(width=W, height=H, spp=spp), threads=threads)
    dt = time.perf_counter() - t0
    rays = int(out["counts"].sum())
    orc.close()
    return {
        "value": rays / dt / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
        "sample": f"same scene and camera at {W}x{H}, {spp} spp, depth {r['max_depth']}: {rays} rays in {dt:.2f} s "
                  f"(oracle BVH build {build_s:.1f} s not timed)",
    }


# U-Net convolutions (input channels, output channels, level): encoder 2 per level, decoder
# [upsample | skip] + 1, then the 3-channel output layer (stored padded to 32 channels)
UNET_CONVS = [(16, 32, 0), (32, 32, 0), (32, 64, 1), (64, 64, 1), (64, 96, 2), (96, 96, 2), (96, 128, 3),
              (128, 128, 3), (224, 96, 2), (96, 96, 2), (160, 64, 1), (64, 64, 1), (96, 32, 0), (32, 32, 0),
              (32, 3, 0)]
MFMA_BF16_PEAK_TFLOPS = 2500.0


def unet_flops(width, height):
    """Algorithmic FLOPs of one inference: 2 x 9 x Cin x Cout per output pixel per conv
    (the 16 input channels as stored, 10 of them used; the output layer's 3 channels)."""
    return sum(2.0 * 9 * ci * co * (width >> lvl) * (height >> lvl) for ci, co, lvl in UNET_CONVS)


def network_roofline(rend, calls=10):
    """MFMA roofline of the network on the frame currently held (1920x1080)."""
    ms = [rend.denoise()["network_ms"] for _ in range(calls)]
    t = sum(ms) / len(ms)
    flops = unet_flops(1920, 1080)
    achieved = flops / (t * 1e-3) / 1e12
    return {"bound": "mfma", "achieved": round(achieved, 1), "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / MFMA_BF16_PEAK_TFLOPS, 4), "flops_per_inference": flops,
            "network_ms": round(t, 4), "profile": "profiles/unet_denoise_1080p.md"}


def denoise_metric(rend, steps):
    """Time to relMSE 0.01 against a 4096-spp render, plain accumulation vs learned
    denoising, on the Cornell box at 1920x1080 seen from its own (held-out) camera. The
    denoiser is first trained on the GPU from frames the tracer renders (training time
    reported separately); the instanced scene's view is reported beside it."""
    import math

    from trayracer_amd import convergence as C
    from trayracer_amd import scenes

    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    rend.denoiser_init(seed=0)
    t_pairs = C.build_training_set(rend, views=(("cornell", 6), ("inst", 3), ("soup", 1)), log=log)
    t_train, losses = C.train(rend, steps=steps, batch=8, crop=128, lr=1e-3)
    log(f"denoiser trained: {steps} steps in {t_train:.1f} s, loss {float(losses[-50:].mean()):.5f}")
    sc = scenes.cornell(width=1920, height=1080)
    res = C.time_to_relmse(rend, sc, sc["camera"], max_plain_spp=4096, log=log)
    inst = C.time_to_relmse(rend, scenes.instanced(), C.INST_CAMERA, max_plain_spp=256, max_denoise_spp=16, log=log)

    def pick(r):
        out = {k: r[k] for k in ("target", "plain_time_s", "plain_spp", "plain_how", "denoised_time_s",
                                 "denoised_spp", "speedup", "reference_spp")}
        out["plain_relmse_by_spp"] = {int(k): round(e, 5) for k, _, e in r["plain_curve"]}
        out["denoised_relmse_by_spp"] = {int(k): round(e, 5) for k, _, e, _ in r["denoised_curve"]}
        out["denoise_ms"] = round(r["denoised_curve"][-1][3], 3)
        return out
    return {"view": "cornell 1920x1080, its own camera (not a training view), depth 4",
            "relmse": "mean (x - r)^2 / (r^2 + 1e-2) vs a 4096-spp render", **pick(res),
            "train": {"pairs_s": round(t_pairs, 1), "train_s": round(t_train, 1), "steps": steps, "batch": 8,
                      "crop": 128, "final_loss": float(losses[-50:].mean())},
            "instanced_view": pick(inst), "network_roofline": network_roofline(rend)}


def main():
    args = parse()
    if args.copy_calibrate:
        copy_calibrate()
        return
    if args.denoise_only:
        denoise_only(args.steps, args.warmup)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "single-GPU build: every configuration runs on one MI355X; "
                                   "multi-GPU runs are not part of this build (--gpus must be 1)"}))
        sys.exit(2)
    from trayracer_amd import Renderer

    scene, workload = make_scene(args.workload)
    rend = Renderer(0)
    t0 = time.perf_counter()
    rend.UpdateScene(scene, presplit=args.presplit)
    build_s = time.perf_counter() - t0
    info = rend.bvh_info()
    for _ in range(args.warmup):
        rend.RenderFrame()
    stats = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        stats.append(rend.RenderFrame())
    wall = time.perf_counter() - t0
    rays = sum(s["extend_rays"] + s["shadow_rays"] for s in stats)
    ext_ms = sum(s["extend_ms"] for s in stats)
    sh_ms = sum(s["shadow_ms"] for s in stats)
    trace_ms = sum(s["trace_ms"] for s in stats)
    shade_ms = sum(s["shade_ms"] for s in stats)
    frame_ms = sum(s["frame_ms"] for s in stats)
    ext_rays = sum(s["extend_rays"] for s in stats)
    sh_rays = sum(s["shadow_rays"] for s in stats)
    launches = sum(s["extend_launches"] for s in stats)

    tc = rend.measure_traversal()
    ray_io = 32 + 16  # ray record in, hit record out
    bytes_per_ext = ray_io + tc["closest_nodes_per_ray"] * tc["node_bytes"] + tc["closest_tris_per_ray"] * tc["tri_bytes"]
    ext_bytes = bytes_per_ext * ext_rays
    achieved = ext_bytes / (ext_ms * 1e-3) / 1e9
    traffic, traffic_src = None, None
    tpath = os.path.join(ROOT, "profiles", "k_extend_sl_soup1m_traffic.json")
    if args.workload == "soup1m" and os.path.exists(tpath):
        with open(tpath) as f:
            traffic = round(json.load(f)["traffic_bytes_per_launch"])
        traffic_src = "rocprofv3 FETCH_SIZE/WRITE_SIZE, copy-calibrated (profiles/k_extend_sl_soup1m_hbm_traffic.md)"
    avg_launch_s = ext_ms / max(1, launches) * 1e-3
    roofline = {
        "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": traffic_src,
        "kernel": "k_extend" if info["two_level"] else "k_extend_sl<false>",
        "avg_launch_ms": round(ext_ms / max(1, launches), 4),
        "dram_gbs": round(traffic / avg_launch_s / 1e9, 1) if traffic else None,
        "bytes_per_ray": round(bytes_per_ext, 1), "nodes_per_ray": round(tc["closest_nodes_per_ray"], 2),
        "tris_per_ray": round(tc["closest_tris_per_ray"], 2),
    }
    dn = None
    if not args.no_denoise:
        dn = denoise_metric(rend, args.denoise_steps)
    cpu = None
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(scene, threads)

    line = {
        "metric": "Mrays/s (1080p, 1M tris); time to rel-MSE 0.01: plain vs denoised",
        "value": round(rays / wall / 1e6, 2), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded procedural scene)",
        "config": {"workload": workload,
                   "triangles": int(scene["positions"].shape[0]) * (1 if scene.get("instances") is None
                                                                    else int(scene["instances"].shape[0])),
                   "rays_per_frame": rays // args.steps},
        "trace_mrays_s": round((ext_rays + sh_rays) / (trace_ms * 1e-3) / 1e6, 1),
        "extend_mrays_s": round(ext_rays / (ext_ms * 1e-3) / 1e6, 1),
        "shadow_mrays_s": round(sh_rays / (sh_ms * 1e-3) / 1e6, 1) if sh_ms > 0 else None,
        "rates_note": "trace_mrays_s: Extend + Shadow rays over the wall time of both phases; extend_mrays_s and "
                      "shadow_mrays_s: each phase's rays over its own launch intervals, and every Shadow launch "
                      "shares the GPU with the next bounce's Extend (second stream), so these understate the "
                      "kernels' own rates",
        "gpu_frame_ms": round(frame_ms / args.steps, 3),
        "phase_ms_per_frame": {"extend": round(ext_ms / args.steps, 3), "shade": round(shade_ms / args.steps, 3),
                               "shadow": round(sh_ms / args.steps, 3), "trace_wall": round(trace_ms / args.steps, 3),
                               "note": "extend and shadow are per-launch intervals that overlap across bounces; "
                                       "trace_wall is the wall time of both phases"},
        "bvh": {"nodes": info["nodes"], "depth": info["depth"], "sah_cost": round(info["sah_cost"], 2),
                "presplit": args.presplit, "references": info["references"], "build_s": round(build_s, 2),
                "trt_build_s": round(info["build_ms"] / 1e3, 2)},
        "roofline": roofline, "cpu_baseline": cpu, "time_to_relmse": dn,
    }
    print(json.dumps(line), flush=True)
    rend.close()


if __name__ == "__main__":
    main()
