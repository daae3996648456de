"""Benchmark: wavefront path tracing of the config-2 atrium at 1080p, depth 8.

One step = one 1920x1080 frame at 1 sample per pixel, depth 8, MIS + NEE,
quad BVH (ct_trace_frame).  value = (primary + extension + shadow rays traced)
/ wall time of the K timed steps, in Mrays/s.  Scene, tree and path state are
resident on the device before timing starts.

roofline: the closest-hit (extend) kernel.  Algorithmic bytes per closest-hit
ray = 28 (origin, direction, tmax) + 8 (t, id) + 64 * quad-node tests + 36 *
triangle tests; node and triangle tests are counted by an instrumented,
untimed frame.  achieved = those bytes / the extend launches' device time
(HIP events on the library stream around every extend launch of the timed
frames).

cpu_baseline: the CPU oracle (FP32 restatement of the reference loop, OpenMP
over pixels) on a horizontal strip of the same frame.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Mrays/s at 1080p depth 8; served p50/p99 ms/request at equal rel-MSE"
HBM_PEAK_GBS = 8000.0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--traversal", default="exact", choices=["exact", "fast"])
    ap.add_argument("--cpu-rows", type=int, default=24, help="rows of the frame the CPU baseline traces")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU form "
              "(image tiles or requests per GPU) is not built.", file=sys.stderr)
        return 2

    import numpy as np
    import cudatracer_amd as ct

    scene = ct.HostScene.generate(ct.SCENE_C2, args.seed)
    opts, W, H = scene.frame_opts()
    opts.traversal = ct.TRAV_FAST if args.traversal == "fast" else ct.TRAV_EXACT
    cam = scene.camera(W, H)
    st_scene = scene.stats()
    ctx = ct.Context(0)
    ctx.upload(scene)
    ctx.resize(W, H)

    # instrumented frame: node / triangle tests of the extend and connect stages
    cst, work = ctx.trace_frame_counted(cam, opts)
    ctx.reset()
    closest_rays = cst.primary_rays + cst.extension_rays
    alg_bytes_frame = 36 * closest_rays + 64 * work[0] + 36 * work[1]

    for _ in range(args.warmup):
        ctx.trace_frame(cam, opts)
    ctx.reset()

    rays = prim = ext = sh = 0
    ext_ms = shade_ms = conn_ms = other_ms = gpu_ms = 0.0
    launches = 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = ctx.trace_frame(cam, opts)
        prim += st.primary_rays
        ext += st.extension_rays
        sh += st.shadow_rays
        ext_ms += st.extend_ms
        shade_ms += st.shade_ms
        conn_ms += st.connect_ms
        other_ms += st.other_ms
        gpu_ms += st.gpu_ms
        launches += opts.max_depth
    t1 = time.perf_counter()
    wall = t1 - t0
    rays = prim + ext + sh
    mrays = rays / wall / 1e6
    ms_per_step = wall * 1e3 / args.steps

    # roofline of the extend kernel: algorithmic bytes per launch / mean duration
    closest_per_frame = (prim + ext) / args.steps
    bytes_per_frame = alg_bytes_frame * (closest_per_frame / max(closest_rays, 1))
    bytes_per_launch = bytes_per_frame / opts.max_depth
    mean_launch_ms = ext_ms / max(launches, 1)
    achieved = bytes_per_launch / (mean_launch_ms * 1e-3) / 1e9
    roofline = {
        "kernel": "k_extend (closest-hit traversal, quad BVH)",
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": None,
        "bytes_per_launch": int(bytes_per_launch),
        "mean_launch_ms": round(mean_launch_ms, 4),
        "node_tests_per_ray": round(work[0] / max(closest_rays, 1), 2),
        "tri_tests_per_ray": round(work[1] / max(closest_rays, 1), 2),
    }

    out = {
        "metric": METRIC,
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "config 2: procedural atrium (%d triangles), 1920x1080, 1 spp, depth 8, MIS+NEE, "
                        "quad BVH, %s traversal, typed BSDF (Lambert/GGX/mirror/dielectric) + 4 quad lights + sky"
                        % (st_scene["triangles"], args.traversal),
            "seed": args.seed,
            "width": W, "height": H, "spp": 1, "depth": opts.max_depth,
            "device": ctx.name(),
        },
        "rays_per_step": {"primary": prim // args.steps, "extension": ext // args.steps, "shadow": sh // args.steps},
        "samples_per_s": round(W * H * args.steps / wall, 1),
        "gpu_ms_per_step": round(gpu_ms / args.steps, 3),
        "stage_ms_per_step": {"extend": round(ext_ms / args.steps, 3), "shade": round(shade_ms / args.steps, 3),
                              "connect": round(conn_ms / args.steps, 3), "other": round(other_ms / args.steps, 3)},
        "roofline": roofline,
    }

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(scene, cam, opts, W, H, args.cpu_rows)
    ctx.close()
    print(json.dumps(out))
    return 0


def cpu_baseline(scene, cam, opts, W, H, rows) -> dict:
    """Oracle FP32 on rows [r0, r0+rows) of the benchmark frame."""
    from oracle import oracle as orc
    threads = min(16, os.cpu_count() or 1)
    osc = orc.Scene(scene.arrays())
    r0 = H // 2 - rows // 2
    t0 = time.perf_counter()
    _, st = osc.render(cam, opts, W, H, 0, 1, rows=(r0, r0 + rows), threads=threads)
    dt = time.perf_counter() - t0
    rays = st["primary_rays"] + st["extension_rays"] + st["shadow_rays"]
    return {"value": round(rays / dt / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"rows {r0}..{r0 + rows - 1} of the 1920x1080 frame ({W * rows} paths, {rays} rays, "
                      f"{dt:.1f} s on {threads} threads)"}


if __name__ == "__main__":
    sys.exit(main())
