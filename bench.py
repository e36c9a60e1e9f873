#!/usr/bin/env python3
"""Throughput of the ray-tracing path on one MI355X.

Workload (default): the two-level instanced scene at 1920x1080, 8 samples
per pixel, depth 8 - 1024 instances of four seeded prototype meshes (5,120 /
12,288 / 20,000 / 131,072 triangles), one 4096^2 RGBA8 Perlin albedo
texture, one emissive sphere - rendered with the ray_trace (hybrid
compute-only) integrator. A step is one frame of the animated scene: the
instances move as rigid bodies (linear and angular velocity, dt = 1/60), the
TLAS is rebuilt on the device, then the frame is traced. Its rays (camera,
bounce and shadow queries) over the wall time of the step give Mrays/s. Scene
data is resident in HBM before the timed region; each frame runs integrate ->
bounds -> TLAS build -> pack, then raygen -> [extend -> shade -> shadow] x
depth -> resolve, all on the device.

Prints one JSON line (see the end of main) with the roofline of the dominant
kernel (closest-hit traversal) and the CPU oracle timed on this host.

    python bench.py [--gpus 1] [--steps 5] [--warmup 1]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# extension-ray traversal kernel of each trace schedule (nearest-first)
TRACE_KERNEL = {0: "k_trace_ww<false, 16, 256>", 1: "k_trace_lane<false, false>", 2: "k_trace<false, false>"}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def bytes_per_ray(counts):
    """Algorithmic bytes per ray of the reference traversal
    (ray_trace.comp.hybrid.hlsl:862-917, :600-707): a 48 B BVHNode per
    visited node, a 128 B Object per leaf object test, a 48 B MeshInstance
    per mesh entered, 3 x 4 B indices + 3 x 32 B vertices per triangle test,
    plus 64 B of ray read and hit write."""
    rays, nodes, objs, insts, tris = [float(x) for x in counts]
    if rays == 0:
        return 0.0
    return (48 * nodes + 128 * objs + 48 * insts + 108 * tris) / rays + 64.0


def measured_traffic(args, kernel):
    """HBM bytes per launch of the extension traversal, from the rocprofv3
    FETCH_SIZE / WRITE_SIZE passes recorded in profiles/traffic.json (scaled
    by the copy-kernel calibration factors), when this run is that workload."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    default = (args.width, args.height, args.spp, args.depth, args.instances) == (1920, 1080, 8, 8, 1024)
    if not default or args.lane_trace or args.trace_schedule not in (None, 0) or not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    if t.get("kernel") != kernel:
        return None
    return 1024.0 * (t["fetch_kib_per_launch"] * t["fetch_factor"] + t["write_kib_per_launch"] * t["write_factor"])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--spp", type=int, default=8)
    ap.add_argument("--depth", type=int, default=8)
    ap.add_argument("--instances", type=int, default=1024)
    ap.add_argument("--tex", type=int, default=4096)
    ap.add_argument("--served-schedule", type=int, default=0, choices=(0, 1),
                    help="served traversal: 0 one ray per lane, 1 persistent pass walk per job")
    ap.add_argument("--graph", type=int, default=1, choices=(0, 1),
                    help="1: each chunk's bounce loop is one captured graph replay; 0: stage by stage")
    ap.add_argument("--cpu-stride", type=int, default=4,
                    help="CPU oracle renders every N-th pixel of one frame (bounded sample)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--autotune", action="store_true",
                    help="run avrt_autotune before the timed steps (default: the default tuning, which the profiles describe)")
    ap.add_argument("--lane-trace", action="store_true", help="one ray per lane instead of persistent lanes")
    ap.add_argument("--trace-schedule", type=int, default=None,
                    help="0 persistent node/leaf passes (default), 1 one ray per lane, 2 one-step persistent walk")
    ap.add_argument("--served-jobs", type=int, default=160, help="jobs of the served stream (0 skips the leg)")
    ap.add_argument("--latency-jobs", type=int, default=1000,
                    help="jobs of the Poisson run that the p50/p99 latency is taken over")
    ap.add_argument("--served-load", type=float, default=0.6,
                    help="Poisson arrival rate of the latency run, as a fraction of the measured served capacity")
    ap.add_argument("--hybrid-frames", type=int, default=16, help="frames of the hybrid orbit (0 skips the leg)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the path has no multi-GPU form here.",
              file=sys.stderr)
        sys.exit(2)

    from aviatorengine2_amd import abi, scenes
    from aviatorengine2_amd.renderer import Renderer

    t0 = time.perf_counter()
    spec = scenes.instanced(width=args.width, height=args.height, spp=args.spp, depth=args.depth,
                            n_instances=args.instances, tex_size=args.tex)
    builder = scenes.Builder()
    library = scenes.mesh_library(spec.meshes, builder)
    packed = scenes.pack(spec, builder, library)
    objs, inst, local = scenes.authoring(spec, library)
    cfg = packed.config.copy()
    log(f"scene built in {time.perf_counter() - t0:.1f}s: {len(packed.objects)} objects, "
        f"{len(packed.indices) // 3} BLAS triangles, {len(packed.blas)} BLAS nodes")

    sched = args.trace_schedule if args.trace_schedule is not None else (1 if args.lane_trace else 0)
    r = Renderer().Initialize(device=0, trace_schedule=sched)
    if args.graph == 0:
        r.SetTuning(**dict(r.GetTuning(), graph=0))
    r.UploadMeshGeometry(library[0], library[1], library[2])
    r.UploadTextures(spec.texture)
    r.SetAuthoring(objs, inst, local)
    dt = 1.0 / 60.0
    for w in range(args.warmup):
        r.SceneStep(spec.velocities, dt)
        r.RefitBLAS()
        s = r.RenderFrame(cfg, abi.MODE_RAY_TRACE)
        log(f"warmup {w}: {s['ms_total']:.1f} ms")
    tuning = None
    if args.autotune and sched == 0:
        t0 = time.perf_counter()
        tune = r.Autotune(cfg, abi.MODE_RAY_TRACE)
        tuning = {"best": tune["best"], "trials": [dict(t, ms=round(ms, 3)) for t, ms in tune["trials"]],
                  "seconds": time.perf_counter() - t0}
        log(f"autotune ({tuning['seconds']:.1f}s): {tune['best']}")

    tot = {"rays_primary": 0, "rays_bounce": 0, "rays_shadow": 0, "paths": 0, "ms_extend": 0.0,
           "ms_shade": 0.0, "ms_shadow": 0.0, "ms_other": 0.0, "launches_extend": 0, "ms_total": 0.0}
    t_start = time.perf_counter()
    refit_s = 0.0
    for k in range(args.steps):
        # the frame step: rigid motion and the device TLAS rebuild, the BLAS
        # refit of every prototype (vertices unchanged for rigid instances,
        # so the trees come back the same), then the frame
        r.SceneStep(spec.velocities, dt)
        t_r = time.perf_counter()
        r.RefitBLAS()
        refit_s += time.perf_counter() - t_r
        s = r.RenderFrame(cfg, abi.MODE_RAY_TRACE)
        for key in tot:
            tot[key] += s[key]
        log(f"step {k}: {s['ms_total']:.1f} ms, rays {s['rays_primary'] + s['rays_bounce'] + s['rays_shadow']}")
    wall = time.perf_counter() - t_start
    in_use = r.GetTuning()
    r.Shutdown()

    rays = tot["rays_primary"] + tot["rays_bounce"] + tot["rays_shadow"]
    ms_per_step = 1000.0 * wall / args.steps
    mrays = rays / wall / 1e6
    samples_per_s = tot["paths"] / wall

    roofline = None
    cpu = None
    if not args.no_cpu:
        import oracle
        threads = min(16, os.cpu_count() or 1)
        log(f"CPU oracle: every {args.cpu_stride}th pixel of one frame on {threads} threads")
        _, _, secs, cnt = oracle.render(packed, cfg, abi.MODE_RAY_TRACE, threads=threads, stats=True,
                                        pixel_stride=args.cpu_stride)
        ext, shd = cnt[:5], cnt[5:]
        cpu_rays = int(ext[0] + shd[0])
        cpu = {"value": cpu_rays / secs / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
               "sample": f"1/{args.cpu_stride} of the pixels of one {args.width}x{args.height}x{args.spp}spp frame "
                         f"({cpu_rays} rays, {secs:.1f} s)"}
        # one core on a smaller sample of the same frame
        stride1 = args.cpu_stride * 16
        _, _, secs1, cnt1 = oracle.render(packed, cfg, abi.MODE_RAY_TRACE, threads=1, stats=True,
                                          pixel_stride=stride1)
        rays1 = int(cnt1[0] + cnt1[5])
        cpu["single_core"] = {"value": rays1 / secs1 / 1e6, "unit": "Mrays/s",
                              "sample": f"1/{stride1} of the pixels ({rays1} rays, {secs1:.1f} s)"}
        log(f"CPU oracle: {cpu['value']:.2f} Mrays/s on {threads} threads, {cpu['single_core']['value']:.2f} on one")
        b_ext = bytes_per_ray(ext)
        ext_rays = tot["rays_primary"] + tot["rays_bounce"]
        launches = max(1, tot["launches_extend"])
        per_launch_bytes = b_ext * ext_rays / launches
        per_launch_ms = tot["ms_extend"] / launches
        kernel = TRACE_KERNEL[sched]
        if tuning is not None and sched == 0:
            kernel = f"k_trace_ww<false, {tuning['best']['stack_lds']}, {tuning['best']['block']}>"
        achieved = per_launch_bytes / (per_launch_ms * 1e-3) / 1e9 if per_launch_ms > 0 else 0.0
        roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(args, kernel),
                    "kernel": kernel,
                    "bytes_per_ray": b_ext, "avg_launch_ms": per_launch_ms}

    served = run_served(args, mrays) if args.served_jobs > 0 else None
    hyb = run_hybrid(args) if args.hybrid_frames > 0 else None

    out = {
        "metric": "Mrays/s, 1080p two-level BVH, served vs plain; p50/p99 job latency at fixed load",
        "value": mrays,
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "two-level instanced scene, rigid motion + device TLAS rebuild + BLAS refit per frame, plain form",
                   "width": args.width, "height": args.height, "spp": args.spp, "depth": args.depth,
                   "instances": args.instances, "texture": args.tex},
        "samples_per_s": samples_per_s,
        "rays": {"primary": tot["rays_primary"], "bounce": tot["rays_bounce"], "shadow": tot["rays_shadow"]},
        "device_ms_per_step": {k: tot[k] / args.steps for k in ("ms_total", "ms_extend", "ms_shade", "ms_shadow",
                                                                "ms_other")},
        "blas_refit_ms_per_step": 1000.0 * refit_s / args.steps,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "tuning": {"in_use": in_use, "autotune": tuning},
        "served": served,
        "hybrid": hyb,
    }
    print(json.dumps(out))


def run_served(args, plain_mrays):
    """The served, batched form on the mixed stream: 8 seeded scenes shaped
    like the sphere field (path_trace) and the icosphere triangle scene
    (ray_trace), resolution uniform over {256^2, 640x360, 1280x720,
    1920x1080}, samples log-uniform in [1, 64], depth 4..8. Throughput: every
    job queued at once, rays / wall time until the last finishes. Latency: a
    Poisson stream at `served_load` x the measured capacity (jobs/s), p50/p99
    of arrival -> result."""
    import numpy as np

    from aviatorengine2_amd import serve

    t0 = time.perf_counter()
    scene_list = serve.mixed_scenes(8)
    log(f"served: 8 scenes built in {time.perf_counter() - t0:.1f}s")
    jobs = serve.poisson_jobs(args.served_jobs, rate_hz=1e9, seed=0x10B5)
    with serve.Server(trace_schedule=args.served_schedule) as srv:
        ids = [srv.add_scene(pk) for pk, _ in scene_list]
        # warm the kernels once
        warm = serve.poisson_jobs(4, rate_hz=1e9, seed=7, max_spp=2)
        serve.run_stream(srv, ids, scene_list, warm)
        s0 = srv.stats()
        lat, wall, st, _ = serve.run_stream(srv, ids, scene_list, jobs, progress=log)
        rays = (st["rays_primary"] + st["rays_bounce"] + st["rays_shadow"]) - (
            s0["rays_primary"] + s0["rays_bounce"] + s0["rays_shadow"])
        cap_jobs = len(jobs) / wall
        mrays = rays / wall / 1e6
        samples = (st["paths"] - s0["paths"]) / wall
        log(f"served throughput: {mrays:.1f} Mrays/s, {cap_jobs:.1f} jobs/s over {wall:.1f}s, {st['rounds'] - s0['rounds']} rounds")
        rate = args.served_load * cap_jobs
        ljobs = serve.poisson_jobs(args.latency_jobs, rate_hz=rate, seed=0x10B6)
        llat, lwall, lst, _ = serve.run_stream(srv, ids, scene_list, ljobs, progress=log)
        p50, p99 = float(np.percentile(llat, 50)) * 1e3, float(np.percentile(llat, 99)) * 1e3
        log(f"served latency at {rate:.1f} jobs/s: p50 {p50:.1f} ms, p99 {p99:.1f} ms")
    return {"mrays_per_s": mrays, "vs_plain": mrays / plain_mrays if plain_mrays else None, "samples_per_s": samples,
            "jobs": len(jobs), "jobs_per_s": cap_jobs, "wall_s": wall,
            "latency": {"rate_jobs_per_s": rate, "load": args.served_load, "p50_ms": p50, "p99_ms": p99,
                        "jobs": len(ljobs), "wall_s": lwall},
            "schedule": "one ray per lane" if args.served_schedule == 0 else "persistent pass walk per job",
            "stream": "8 scenes (4 sphere fields path_trace, 4 icosphere triangle scenes ray_trace), "
                      "res {256^2,640x360,1280x720,1920x1080}, spp log-uniform [1,64], depth 4-8"}


def run_hybrid(args):
    """The hybrid frame on the icosphere scene in two-level form: ray-cast
    G-buffer at 1920x1080, deferred lighting on the 960x540 trace grid (1
    sample: NEE + bounce path, depth 5), upsample + reprojection to 1080p,
    over an orbit of `hybrid_frames` frames at 1 degree per frame around the
    scene at the scene camera's own distance."""
    import math

    from aviatorengine2_amd import hybrid, scenes
    from aviatorengine2_amd.renderer import Renderer

    spec = scenes.hybrid_scene(width=1920, height=1080, spp=1, depth=5, tex_size=1024)
    pk = scenes.pack(spec)
    c = pk.config
    src = (float(c["source_x"]), float(c["source_y"]), float(c["source_z"]))
    radius = math.hypot(src[0], src[2])
    with Renderer().Initialize() as r:
        r.LoadScene(pk)
        frames = hybrid.HybridFrames(r, c, trace_scale=0.5)
        frames.render()  # warm-up frame
        rays = 0
        dev_ms = 0.0
        t0 = time.perf_counter()
        for f in range(args.hybrid_frames):
            frames.set_camera(hybrid.orbit((0.0, 0.0, 0.0), radius, src[1], 1.0 * (f + 1)), (0.0, 0.0, 0.0))
            _, st = frames.render()
            rays += st["rays_primary"] + st["rays_bounce"] + st["rays_shadow"]
            dev_ms += st["ms_total"]
        wall = time.perf_counter() - t0
    g = args.hybrid_frames
    log(f"hybrid: {1e3 * wall / g:.2f} ms/frame wall, {dev_ms / g:.2f} ms device, {rays / wall / 1e6:.1f} Mrays/s")
    return {"ms_per_frame": 1e3 * wall / g, "device_ms_per_frame": dev_ms / g, "mrays_per_s": rays / wall / 1e6,
            "frames": g, "gbuffer": "1920x1080", "trace_grid": "960x540", "output": "1920x1080",
            "workload": "icosphere scene (81,920-tri mesh + ground mesh, quad light, glass/metal spheres, proxy box), "
                        "orbit 1 deg/frame"}


if __name__ == "__main__":
    main()
