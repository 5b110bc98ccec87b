"""Benchmark: path-traced frames of the 1M-triangle soup at 1920x1080, 8 spp, depth 8.

A step is one frame: RayGen, then per bounce Extend (closest-hit traversal), Shade
(integrator + wave-aggregated queue appends) and Shadow (any-hit traversal), then
Finalize, all on one MI355X with the scene already resident in HBM. `value` is the
whole-frame throughput: rays traced (extension + shadow) / wall time of the timed
frames. Traversal-only throughput (rays / wall time of the Extend and Shadow phases from
HIP events; Shadow of bounce d runs beside Extend of bounce d + 1 on a second stream) is
reported beside it, with two rooflines of the Extend kernel (HBM: the bytes a launch
must move at least, against 8 TB/s; vector memory: the bytes every ray loads, against the
measured rate of random 48-byte gathers from the L2; counts from a counting pass over the
same frame) and Shade's HBM roofline.
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
    ap.add_argument("--denoise-steps", type=int, default=12000)
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


def host_cpu():
    """CPU model and the host threads this process may use: the CPUs of its affinity mask,
    capped by OMP_NUM_THREADS when that is set (the per-process CPU share of a shared
    machine)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    affinity = len(os.sched_getaffinity(0))
    omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return model, affinity, (min(affinity, omp) if omp > 0 else affinity)


CPU_CROP = (832, 412, 256, 256)  # the 1080p film's centre


def cpu_baseline(threads, presplit, crop_spp=640):
    """The host build of the same per-vertex code (camera, traversal, BSDF, integrator)
    on the host cores, traversing the product's own binned-SAH tree over pre-split
    references (the GPU's tree, built by the same host code), near child first: the
    Cornell configuration (256x256, 16 spp, depth 4) in full, and the 1M-triangle soup's
    frame on a fixed 256x256 window at the film's centre, the same paths as the GPU
    frame's. One spp of that window (~0.24 M rays, a few tens of ms) is timed and
    reported as well; the headline rate traces `crop_spp` samples, about 10 s of work on
    16 host threads, for a stable figure."""
    from oracle import Oracle
    from trayracer_amd import scenes
    model, affinity, _ = host_cpu()
    sc0 = scenes.cornell()
    orc = Oracle(sc0)
    orc.use_product_bvh(presplit, threads)
    t0 = time.perf_counter()
    out0 = orc.render(threads=threads)
    dt0 = time.perf_counter() - t0
    rays0 = int(out0["counts"].sum())
    orc.close()
    sc1 = scenes.soup()
    t0 = time.perf_counter()
    orc = Oracle(sc1)
    orc.use_product_bvh(presplit, threads)
    build_s = time.perf_counter() - t0
    x0, y0, w, h = CPU_CROP
    t0 = time.perf_counter()
    one = orc.render(render=dict(spp=1), crop=CPU_CROP, threads=threads)
    dt_one = time.perf_counter() - t0
    rays_one = int(one["counts"].sum())
    t0 = time.perf_counter()
    out1 = orc.render(render=dict(spp=crop_spp), crop=CPU_CROP, threads=threads)
    dt1 = time.perf_counter() - t0
    rays1 = int(out1["counts"].sum())
    orc.close()
    r = sc1["render"]
    return {
        "value": round(rays1 / dt1 / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
        "sample": f"1M-triangle soup frame ({r['width']}x{r['height']}, depth {r['max_depth']}), window {w}x{h} at "
                  f"({x0}, {y0}), {crop_spp} spp: {rays1} rays in {dt1:.2f} s on {threads} threads; the product's "
                  f"pre-split SAH tree (build {build_s:.1f} s, not timed)",
        "cpu_model": model, "affinity_cpus": affinity,
        "crop_1spp": {"rays": rays_one, "seconds": round(dt_one, 4), "mrays_s": round(rays_one / dt_one / 1e6, 3)},
        "cornell_full": {"config": "256x256, 16 spp, depth 4, RR from depth 3", "seconds": round(dt0, 3),
                         "rays": rays0, "mrays_s": round(rays0 / dt0 / 1e6, 3)},
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
    """Time to relMSE 0.01 against a converged render, plain accumulation vs learned
    denoising, on the two scenes the denoiser's inputs come from: the 1M-triangle soup
    (lit by the sky: alone it renders black) and the 1000-instance scene, at 1920x1080,
    each from a camera that is not among the training views. The denoiser is first
    trained on the GPU from frames the tracer renders of both scenes under random
    cameras (training time reported separately)."""
    from trayracer_amd import convergence as C

    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    rend.denoiser_init(seed=0)
    t_pairs = C.build_training_set(rend, log=log)
    scale = steps / sum(n for n, _ in C.TRAIN_SCHEDULE)
    schedule = tuple((max(1, int(round(n * scale))), lr) for n, lr in C.TRAIN_SCHEDULE)
    t_train, losses = C.train(rend, schedule=schedule, batch=8, crop=128)
    log(f"denoiser trained: {len(losses)} steps in {t_train:.1f} s, loss {float(losses[-50:].mean()):.5f}")
    soup = C.time_to_relmse(rend, C.make_scene("soup"), C.SOUP_CAMERA, max_plain_spp=2048, max_denoise_spp=64, log=log)
    inst = C.time_to_relmse(rend, C.make_scene("inst"), C.INST_CAMERA, max_plain_spp=256, max_denoise_spp=64, log=log)

    def pick(r, view):
        out = {"view": view}
        out.update({k: r[k] for k in ("target", "plain_time_s", "plain_spp", "plain_how", "denoised_time_s",
                                      "denoised_spp", "speedup", "reference_spp", "failure")})
        out["reference_relmse"] = round(r["reference_relmse"], 5)
        out["plain_relmse_by_spp"] = {int(k): round(e, 5) for k, _, e in r["plain_curve"]}
        out["denoised_relmse_by_spp"] = {int(k): round(e, 5) for k, _, e, _ in r["denoised_curve"]}
        out["denoise_ms"] = round(r["denoised_curve"][-1][3], 3)
        return out
    return {"relmse": "mean (x - r)^2 / (r^2 + 1e-2) vs the mean of two independent reference halves; "
                      "targets compare relMSE minus the reference's estimated own error",
            "soup": pick(soup, "1M-triangle soup under the sky, 1920x1080, camera (0.5, 0.5, -1.2), depth 8"),
            "instanced": pick(inst, "1000-instance scene, 1920x1080, camera (50, 50, -20), depth 6"),
            "train": {"pairs_s": round(t_pairs, 1), "train_s": round(t_train, 1), "steps": int(len(losses)),
                      "schedule": [[n, lr] for n, lr in schedule], "batch": 8, "crop": 128,
                      "views": f"{C.TRAIN_VIEWS} random cameras, inputs {C.TRAIN_INPUT_SPP} spp, targets 1024 spp",
                      "final_loss": float(losses[-50:].mean())},
            "network_roofline": network_roofline(rend)}


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_extend_sl_soup1m_traffic.json")
# the sources whose code the measured traffic belongs to
TRAVERSAL_SOURCES = ["trayracer_amd/csrc/kernels.hip", "trayracer_amd/csrc/device.h", "trayracer_amd/csrc/bvh_build.cpp",
                     "trayracer_amd/csrc/bvh_build.h", "trayracer_amd/csrc/api.cpp", "trayracer_amd/csrc/trt/intersect.h"]


def traversal_source_hash():
    import hashlib
    h = hashlib.sha256()
    for rel in TRAVERSAL_SOURCES:
        with open(os.path.join(ROOT, rel), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def gather_ceilings():
    """Rates of independent 48-byte random gathers (trt_measure_gather) from a buffer that
    fits a CU's L1 (16 KiB), one that fits the L2 (4 MiB), one that fits the Infinity
    Cache (128 MiB) and one in HBM (4 GiB)."""
    import ctypes

    from trayracer_amd import _lib
    lib = _lib.load()
    out = {}
    for name, size in (("l1_16KiB", 16 << 10), ("l2_4MiB", 4 << 20), ("mall_128MiB", 128 << 20),
                       ("hbm_4GiB", 4 << 30)):
        g = ctypes.c_double()
        st = lib.trt_measure_gather(0, size, 256, ctypes.byref(g))
        out[name] = round(g.value, 1) if st == 0 else None
    return out


def load_instruction_costs():
    """Time per wave-level 16-byte load instruction per CU (trt_measure_gather_lanes) for
    48-byte gathers from an L1-resident buffer: every lane its own random record with 64,
    32 and 8 active lanes, and all active lanes at one record with 64 and 8. Equal times
    with 64 and 8 lanes at one record mean that an idle lane costs the load path as much
    as an active one."""
    import ctypes

    from trayracer_amd import _lib
    lib = _lib.load()
    out = {}
    for name, active, same in (("random_64_lanes", 64, 0), ("random_32_lanes", 32, 0), ("random_8_lanes", 8, 0),
                               ("one_record_64_lanes", 64, 1), ("one_record_8_lanes", 8, 1)):
        v = ctypes.c_double()
        st = lib.trt_measure_gather_lanes(0, 16 << 10, 256, active, same, ctypes.byref(v))
        out[name] = round(v.value, 2) if st == 0 else None
    return out


def extend_roofline(tc, info, ext_rays, ext_ms, launches, workload):
    """Two ceilings for the closest-hit traversal (Extend), per launch, from HIP-event
    launch times:
    * HBM: the bytes a launch must move at least, its rays' records (32 B ray in, 16 B hit
      out) plus every distinct node and triangle record it reads once (48 B each; counted
      by the counting pass's touch bitmaps), against the 8 TB/s peak; `traffic` is the
      rocprofv3 measurement of the same kernel (FETCH_SIZE + WRITE_SIZE, copy-calibrated),
      null when the profile was taken from other kernel sources;
    * vector memory: the bytes every ray loads (ray and hit records, node visits x 48 B,
      triangle tests x 48 B) against the rate of independent random 48-byte gathers from
      an L1-resident buffer (trt_measure_gather), the fastest such loads go through the
      CUs' load path (Extend's node loads mostly hit L1 and L2)."""
    n = max(1, launches)
    avg_s = ext_ms / n * 1e-3
    per_ray = 48 + tc["closest_nodes_per_ray"] * tc["node_bytes"] + tc["closest_tris_per_ray"] * tc["tri_bytes"]
    distinct = (tc["closest_distinct_nodes"] * tc["node_bytes"] + tc["closest_distinct_tris"] * tc["tri_bytes"]) / \
        max(1, tc["launches"])
    rays_per_launch = ext_rays / n
    compulsory = 48 * rays_per_launch + distinct
    achieved = compulsory / avg_s / 1e9
    traffic, stale = None, None
    if workload == "soup1m" and os.path.exists(TRAFFIC_PROFILE):
        with open(TRAFFIC_PROFILE) as f:
            prof = json.load(f)
        stale = prof.get("kernel_source_sha16") != traversal_source_hash()
        if not stale:
            traffic = round(prof["traffic_bytes_per_launch"])
    ceil = gather_ceilings()
    loaded = per_ray * rays_per_launch / avg_s / 1e9
    vm_peak = ceil["l1_16KiB"]
    return {
        "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
        "traffic_source": "rocprofv3 FETCH_SIZE + WRITE_SIZE per launch, copy-calibrated "
                          "(profiles/k_extend_sl_soup1m_hbm_traffic.md)" if traffic else None,
        "traffic_stale": stale,
        "dram_gbs": round(traffic / avg_s / 1e9, 1) if traffic else None,
        "dram_frac": round(traffic / avg_s / 1e9 / HBM_PEAK_GBS, 4) if traffic else None,
        "kernel": "k_extend_tl<false>" if info["two_level"] else "k_extend_sl<false>",
        "avg_launch_ms": round(ext_ms / n, 4), "rays_per_launch": round(rays_per_launch),
        "algorithmic_bytes_per_launch": round(compulsory),
        "algorithmic_bytes": "48 B per ray (ray in, hit out) + 48 B per distinct node and triangle record read",
        "vector_memory": {"achieved": round(loaded, 1), "peak": vm_peak, "unit": "GB/s",
                          "frac": round(loaded / vm_peak, 4) if vm_peak else None,
                          "bytes_per_ray": round(per_ray, 1), "nodes_per_ray": round(tc["closest_nodes_per_ray"], 2),
                          "tris_per_ray": round(tc["closest_tris_per_ray"], 2),
                          "loads_16B_per_ray": round(per_ray / 16.0, 1),
                          "peak_source": "trt_measure_gather: independent random 48-byte gathers, 16 KiB buffer",
                          "gather_ceilings_gbs": ceil,
                          "load_instr_ns_per_cu": load_instruction_costs()},
        "stack_spills_per_ray": round(tc["closest_spills"] / max(1, tc["closest_rays"]), 4),
    }


def shade_roofline(scene, stats, steps, info):
    """Shade's rooflines per frame. HBM: the bytes a path's vertex must stream at least,
    its queue record in (52 B: origin, direction, throughput, path), its hit (16 B, + 4 B
    instance), and the records it writes, an extension (52 B) for each next-bounce ray and
    a shadow ray (48 B) for each shadow ray emitted, over the frame's Shade time against
    8 TB/s. Shade writes no radiance (the shadow stream adds it); its emission-list
    entries (16 B per vertex that reaches an emitter or the sky) are not counted, so the
    figure is a lower bound. Shade's launches run beside the previous bounce's shadow
    launch, whose tail holds CU slots, and their intervals include that time. The surface
    data it gathers (triangle 36 B, + 36 B normals, material id 4 B, material 80 B, light
    record and alias entry 88 B per shaded hit) is mostly served by the caches and is
    reported beside."""
    two = bool(info["two_level"])
    nrm = scene.get("normals") is not None
    ext = sum(s["extend_rays"] for s in stats) / steps
    prim = sum(s["primary_rays"] for s in stats) / steps
    shadow = sum(s["shadow_rays"] for s in stats) / steps
    ms = sum(s["shade_ms"] for s in stats) / steps
    stream = ext * (52 + 16 + (4 if two else 0)) + (ext - prim) * 52 + shadow * 48
    gathered = 36 + (36 if nrm else 0) + 4 + 80 + 88
    achieved = stream / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "streamed_bytes_per_frame": round(stream),
            "gathered_bytes_per_path": gathered,
            "gathered_gbs": round(gathered * ext / (ms * 1e-3) / 1e9, 1), "paths_per_frame": round(ext),
            "shade_ms_per_frame": round(ms, 3),
            "valu_profile": "profiles/wavefront_soup1m_counters.md (VALU load, lane use, TA/TD busy per kernel)"}


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
    # The timed frames record only the events that keep the two streams in order (every
    # event between two kernels delays the second by a few microseconds); the phase
    # breakdown comes from as many frames again with phase timing on.
    rend.SetPhaseTiming(False)
    stats = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        stats.append(rend.RenderFrame())
    wall = time.perf_counter() - t0
    rend.SetPhaseTiming(True)
    pstats = [rend.RenderFrame() for _ in range(args.steps)]
    rays = sum(s["extend_rays"] + s["shadow_rays"] for s in stats)
    frame_ms = sum(s["frame_ms"] for s in stats)
    ext_ms = sum(s["extend_ms"] for s in pstats)
    sh_ms = sum(s["shadow_ms"] for s in pstats)
    trace_ms = sum(s["trace_ms"] for s in pstats)
    shade_ms = sum(s["shade_ms"] for s in pstats)
    timed_frame_ms = sum(s["frame_ms"] for s in pstats)
    ext_rays = sum(s["extend_rays"] for s in pstats)
    sh_rays = sum(s["shadow_rays"] for s in pstats)
    launches = sum(s["extend_launches"] for s in pstats)

    tc = rend.measure_traversal()
    roofline = extend_roofline(tc, info, ext_rays, ext_ms, launches, args.workload)
    shade = shade_roofline(scene, pstats, args.steps, info)
    dn = None
    if not args.no_denoise:
        dn = denoise_metric(rend, args.denoise_steps)
    cpu = None
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or host_cpu()[2]
        cpu = cpu_baseline(threads, args.presplit)

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
        "rates_note": "trace_mrays_s: Extend + Shadow rays over trace_wall (frame minus RayGen, Shade and "
                      "Finalize; Shade runs beside the previous bounce's Shadow, so that overlap counts as Shade "
                      "time and trace_mrays_s overstates the two phases' rate by as much); extend_mrays_s and "
                      "shadow_mrays_s: each phase's rays over its own launch intervals, and every Shadow launch "
                      "shares the GPU with the next bounce's Extend (second stream), so these understate the "
                      "kernels' own rates",
        "gpu_frame_ms": round(frame_ms / args.steps, 3),
        "phase_ms_per_frame": {"extend": round(ext_ms / args.steps, 3), "shade": round(shade_ms / args.steps, 3),
                               "shadow": round(sh_ms / args.steps, 3), "trace_wall": round(trace_ms / args.steps, 3),
                               "frame": round(timed_frame_ms / args.steps, 3),
                               "note": "from as many frames again as were timed, rendered with phase timing on "
                                       "(the timed frames record only the events that order the two streams); "
                                       "extend and shadow are per-launch intervals that overlap across bounces; "
                                       "trace_wall is the frame less RayGen, Shade and Finalize"},
        "bvh": {"nodes": info["nodes"], "depth": info["depth"], "sah_cost": round(info["sah_cost"], 2),
                "presplit": args.presplit, "references": info["references"], "build_s": round(build_s, 2),
                "trt_build_s": round(info["build_ms"] / 1e3, 2)},
        "roofline": roofline, "shade_roofline": shade, "cpu_baseline": cpu, "time_to_relmse": dn,
    }
    print(json.dumps(line), flush=True)
    rend.close()


if __name__ == "__main__":
    main()
