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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--spp", type=int, default=8)
    ap.add_argument("--depth", type=int, default=8)
    ap.add_argument("--instances", type=int, default=1024)
    ap.add_argument("--tex", type=int, default=4096)
    ap.add_argument("--cpu-stride", type=int, default=8,
                    help="CPU oracle renders every N-th pixel of one frame (bounded sample)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--lane-trace", action="store_true", help="one ray per lane instead of persistent lanes")
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

    r = Renderer().Initialize(device=0, lane_trace=args.lane_trace)
    r.UploadMeshGeometry(library[0], library[1], library[2])
    r.UploadTextures(spec.texture)
    r.SetAuthoring(objs, inst, local)
    dt = 1.0 / 60.0
    for w in range(args.warmup):
        r.SceneStep(spec.velocities, dt)
        s = r.RenderFrame(cfg, abi.MODE_RAY_TRACE)
        log(f"warmup {w}: {s['ms_total']:.1f} ms")

    tot = {"rays_primary": 0, "rays_bounce": 0, "rays_shadow": 0, "paths": 0, "ms_extend": 0.0,
           "ms_shade": 0.0, "ms_shadow": 0.0, "ms_other": 0.0, "launches_extend": 0, "ms_total": 0.0}
    t_start = time.perf_counter()
    for k in range(args.steps):
        r.SceneStep(spec.velocities, dt)
        s = r.RenderFrame(cfg, abi.MODE_RAY_TRACE)
        for key in tot:
            tot[key] += s[key]
        log(f"step {k}: {s['ms_total']:.1f} ms, rays {s['rays_primary'] + s['rays_bounce'] + s['rays_shadow']}")
    wall = time.perf_counter() - t_start
    r.Shutdown()

    rays = tot["rays_primary"] + tot["rays_bounce"] + tot["rays_shadow"]
    ms_per_step = 1000.0 * wall / args.steps
    mrays = rays / wall / 1e6
    samples_per_s = tot["paths"] / wall

    roofline = None
    cpu = None
    if not args.no_cpu:
        import oracle
        threads = os.cpu_count() or 1
        log(f"CPU oracle: every {args.cpu_stride}th pixel of one frame on {threads} threads")
        _, _, secs, cnt = oracle.render(packed, cfg, abi.MODE_RAY_TRACE, threads=threads, stats=True,
                                        pixel_stride=args.cpu_stride)
        ext, shd = cnt[:5], cnt[5:]
        cpu_rays = int(ext[0] + shd[0])
        cpu = {"value": cpu_rays / secs / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
               "sample": f"1/{args.cpu_stride} of the pixels of one {args.width}x{args.height}x{args.spp}spp frame "
                         f"({cpu_rays} rays, {secs:.1f} s)"}
        b_ext = bytes_per_ray(ext)
        ext_rays = tot["rays_primary"] + tot["rays_bounce"]
        launches = max(1, tot["launches_extend"])
        per_launch_bytes = b_ext * ext_rays / launches
        per_launch_ms = tot["ms_extend"] / launches
        achieved = per_launch_bytes / (per_launch_ms * 1e-3) / 1e9 if per_launch_ms > 0 else 0.0
        roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "k_extend",
                    "bytes_per_ray": b_ext, "avg_launch_ms": per_launch_ms}

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
        "config": {"workload": "two-level instanced scene, rigid motion + device TLAS rebuild per frame, plain form",
                   "width": args.width, "height": args.height, "spp": args.spp, "depth": args.depth,
                   "instances": args.instances, "texture": args.tex},
        "samples_per_s": samples_per_s,
        "rays": {"primary": tot["rays_primary"], "bounce": tot["rays_bounce"], "shadow": tot["rays_shadow"]},
        "device_ms_per_step": {k: tot[k] / args.steps for k in ("ms_total", "ms_extend", "ms_shade", "ms_shadow",
                                                                "ms_other")},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
