"""Benchmark: wavefront path tracing of the config-2 atrium at 1080p, depth 8.

One step = one 1920x1080 frame at 1 sample per pixel, depth 8, MIS + NEE,
quad BVH (ct_trace_frame).  value = (primary + extension + shadow rays traced)
/ wall time of the K timed steps, in Mrays/s.  Scene, tree and path state are
resident on the device before timing starts.  The K frames run twice over the
same frame indices: with ct_frame_stats (ray counts, per-stage device times
from HIP events around every launch) and without (the wall time); both passes
render identical frames.

roofline: the traversal kernel k_trace (one launch per bounce b traces the
extension rays of bounce b and the shadow rays of bounce b - 1; with
roulette, k_tail traces and shades the bounces from rr_start on).  Algorithmic
bytes per closest-hit ray = 28 (origin, direction, tmax) + 8 (t, id) + 64 *
quad-node tests + 36 * triangle tests, per shadow ray 48 + 64 * node tests +
36 * triangle tests; node and triangle tests are counted by an instrumented,
untimed frame.  Three labelled figures:
  - per launch (the `achieved` / `frac` of the line): bytes per launch / the
    mean launch duration over the timed frames (HIP events on the lane's
    stream around every launch).  A frame's two chunks run on two lanes, so
    a launch shares the device with the other lane's kernels;
  - isolated: the same bytes / the launch durations of a frame whose chunks
    run one at a time (ct_trace_frame_serial, untimed by the wall clock):
    each launch has the device to itself;
  - device level: the frame's traversal bytes / the frame wall time (a lower
    bound on the device's traversal bandwidth: the frame also shades).

cpu_baseline: the CPU oracle (FP32 restatement of the reference loop, OpenMP
over pixels, OMP_NUM_THREADS threads or the process's CPU affinity) on whole
frames of the same workload, about 10 s of CPU work; the host's core count,
CPU model and OMP_NUM_THREADS are reported with it.

c3 (with the default config 2): the 7.24M-triangle displaced sphere at
3840x2160, depth 16, roulette from bounce 3 (the HBM-bound configuration):
rays/s, the three roofline figures, frame 0 against the oracle on a
3840x135 band through the middle, and a CPU baseline on that band.

accuracy: frame 0 of the workload on the GPU against the oracle on a band of
rows through the middle of the image (bit-identical fraction vs FP32,
rel-MSE vs FP64), and the rel-MSE vs FP64 of the mean of 16 frames.

served (unless --no-serve): render requests on the resident atrium arriving
as a Poisson stream (resolution 256x256 / 512x512 / 1920x1080 with
probability .5 / .35 / .15, spp uniform in {1, 4, 16}, depth 8, seeded
cameras).  p50 / p99 of request latency (completion - arrival) for the render
server (continuous batching) and for the plain per-view loop of the reference
(one request at a time, one ct_trace_frame call per 1-spp frame, like
launchKernels); images/s of both with all requests queued at once, and of the
plain loop with all spp of a request in one call (speedup: over the per-frame
loop; speedup_vs_multi_spp: over the one-call loop).  Every request renders
the same path samples either way (checked bit-exact on a subset), so the
comparison is at equal rel-MSE.

batched: 64 views of the box-and-sphere scene (sides log-uniform in
[128, 1024], aspect 1:1 or 4:3, spp in {1, 2, 4, 8, 16}, depth 8) in one
ct_render_batch call after ct_autotune vs the same two plain per-view loops;
images/s and the speedup over each.
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
WORKLOADS = {
    2: "config 2: procedural atrium (%d triangles), 1920x1080, 1 spp, depth 8, MIS+NEE, quad BVH, %s traversal, "
       "typed BSDF (Lambert/GGX/mirror/dielectric) + 4 quad lights + sky",
    3: "config 3: displaced icosphere + noisy ground (%d triangles), 3840x2160, 1 spp, depth 16, Russian roulette "
       "from bounce 3, quad BVH, %s traversal, Beckmann/GGX microfacets + 1 area light",
    4: "config 2 textured: procedural atrium (%d triangles), 1920x1080, 1 spp, depth 8, MIS+NEE, quad BVH, %s "
       "traversal, typed BSDF; tiled + normal-mapped floors, stone columns, 1024x512 sky texture",
}
HBM_PEAK_GBS = 8000.0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--traversal", default="exact", choices=["exact", "fast", "compact"])
    ap.add_argument("--config", type=int, default=2, choices=[2, 3, 4],
                    help="2: atrium 1080p depth 8 (the headline workload); 3: displaced sphere 4K depth 16, RR; "
                         "4: the atrium with textures, normal maps and a sky texture")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="minimum CPU-baseline work")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-build", action="store_true", help="skip the device tree-build leg")
    ap.add_argument("--no-c3", action="store_true", help="skip the 7.24M-triangle 4K leg of the default run")
    ap.add_argument("--build-only", action="store_true",
                    help="only the tree-build leg (device build vs host build of the workload's scene)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (calibrates FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--summarize", nargs="+", metavar="DB", help="reduce rocprofv3 result databases and exit")
    ap.add_argument("--block", type=int, default=0, help="threads per block of the traversal kernels (0 = default)")
    ap.add_argument("--stack-lds", type=int, default=0, help="per-lane LDS stack entries (0 = default)")
    ap.add_argument("--chunk", type=int, default=0, help="paths per wavefront chunk (0 = the library default)")
    ap.add_argument("--serve-chunk", type=int, default=0, help="paths per server chunk (0 = server default)")
    ap.add_argument("--sort", type=int, default=-1, choices=[-1, 0, 1],
                    help="block-local material sort before shading (-1 = the library default)")
    args = ap.parse_args()
    if args.summarize:
        return summarize(args.summarize)
    if args.copy_probe:
        return copy_probe(args.copy_probe)
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU form "
              "(image tiles or requests per GPU) is not built.", file=sys.stderr)
        return 2

    import numpy as np
    import cudatracer_amd as ct

    scene = ct.HostScene.generate({2: ct.SCENE_C2, 3: ct.SCENE_C3, 4: ct.SCENE_C2_TEXTURED}[args.config], args.seed)
    opts, W, H = scene.frame_opts()
    opts.traversal = getattr(ct, "TRAV_" + args.traversal.upper())
    if args.sort >= 0:
        opts.sort_by_material = args.sort
    cam = scene.camera(W, H)
    st_scene = scene.stats()
    ctx = ct.Context(0, block_size=args.block, stack_lds=args.stack_lds, chunk_paths=args.chunk)
    if args.build_only:
        out = {"workload": WORKLOADS[args.config] % (st_scene["triangles"], args.traversal), "device": ctx.name(),
               "build": build_leg(ctx, scene, wide=args.traversal == "fast", reps=max(args.steps, 1))}
        ctx.close()
        if not args.no_cpu and args.config == 2:
            out["build"]["cpu_baseline"] = cpu_build_baseline(scene)
        print(json.dumps(out))
        return 0
    ctx.upload(scene)
    ctx.resize(W, H)

    tm = timed_frames(ctx, cam, opts, args.steps, args.warmup)
    wall, ms_per_step = tm["wall"], tm["ms_per_step"]
    mrays = tm["rays"] / wall / 1e6
    roofline = trace_roofline(ctx, cam, opts, tm, args.config, args.traversal)
    prim, ext, sh = tm["primary"], tm["extension"], tm["shadow"]

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
            "workload": WORKLOADS[args.config] % (st_scene["triangles"], args.traversal),
            "seed": args.seed,
            "width": W, "height": H, "spp": 1, "depth": opts.max_depth, "rr_start": opts.rr_start,
            "sort_by_material": int(opts.sort_by_material),
            "device": ctx.name(),
        },
        "rays_per_step": {"primary": prim // args.steps, "extension": ext // args.steps, "shadow": sh // args.steps},
        "samples_per_s": round(W * H * args.steps / wall, 1),
        "instrumented_ms_per_step": round(tm["instrumented_ms_per_step"], 3),
        "gpu_ms_per_step": round(tm["gpu_ms"] / args.steps, 3),
        "stage_ms_per_step": {k: round(v / args.steps, 3) for k, v in tm["stage_ms"].items()},
        "roofline": roofline,
    }

    out["display"] = display_leg(ctx, W, H)
    if not args.no_cpu:
        out["accuracy"] = accuracy_strip(ctx, scene, cam, opts, W, H, strip=128 if args.config == 3 else 32)
    if not args.no_serve and args.config == 2:
        out["served"] = serve_leg(ct, ctx, scene, opts, args.seed, chunk=args.serve_chunk)
        out["batched"] = batched_leg(ct, ctx, args.seed)
    if not args.no_build:
        # replaces the scene with its device-built twin (identical trees and records)
        out["build"] = build_leg(ctx, scene, wide=args.traversal == "fast")
    if args.config == 2 and not args.no_c3:
        out["c3"] = c3_leg(ct, ctx, args)  # replaces the scene
    ctx.close()
    if not args.no_cpu and not args.no_build and args.config == 2:
        out["build"]["cpu_baseline"] = cpu_build_baseline(scene)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(scene, cam, opts, W, H, args.cpu_seconds,
                                           strip=16 if args.config == 3 else 1)
    print(json.dumps(out))
    return 0


def timed_frames(ctx, cam, opts, steps: int, warmup: int) -> dict:
    """warmup frames, then `steps` frames twice over the same frame indices
    (ct_reset before each pass, so both passes render identical frames with
    identical ray counts): first with ct_frame_stats, which records HIP
    events around every launch and returns the frames' ray counts and
    per-stage device times; then without, timed by the wall clock (each frame
    ends with the stream synchronisation of ct_trace_frame).  The events cost
    a frame about 4% on C2, so the throughput comes from the second pass."""
    for _ in range(warmup):
        ctx.trace_frame(cam, opts, stats=False)
    ctx.reset()
    acc = {"primary": 0, "extension": 0, "shadow": 0, "launches": 0, "chunks": 0, "gpu_ms": 0.0,
           "stage_ms": {"trace": 0.0, "shade": 0.0, "trace_last_shadow": 0.0, "tail": 0.0, "other": 0.0}}
    t0 = time.perf_counter()
    for _ in range(steps):
        st = ctx.trace_frame(cam, opts)
        acc["primary"] += st.primary_rays
        acc["extension"] += st.extension_rays
        acc["shadow"] += st.shadow_rays
        acc["launches"] += st.extend_launches
        acc["chunks"] += st.chunks
        acc["gpu_ms"] += st.gpu_ms
        acc["stage_ms"]["trace"] += st.extend_ms
        acc["stage_ms"]["shade"] += st.shade_ms
        acc["stage_ms"]["trace_last_shadow"] += st.connect_ms
        acc["stage_ms"]["tail"] += st.tail_ms
        acc["stage_ms"]["other"] += st.other_ms
    acc["instrumented_ms_per_step"] = (time.perf_counter() - t0) * 1e3 / steps
    ctx.reset()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.trace_frame(cam, opts, stats=False)
    acc["wall"] = time.perf_counter() - t0
    acc["steps"] = steps
    acc["ms_per_step"] = acc["wall"] * 1e3 / steps
    acc["rays"] = acc["primary"] + acc["extension"] + acc["shadow"]
    return acc


def trace_roofline(ctx, cam, opts, tm: dict, config: int, traversal: str) -> dict:
    """The traversal's algorithmic bytes (from an instrumented frame's node
    and triangle test counts) against three times: the summed durations of
    the launches that trace the timed frames' rays, the same launches in a
    serial frame (each alone on the device), and the frame wall time.  With
    roulette, the bounces from rr_start on are traced inside k_tail, which
    also shades them: its whole duration counts as traversal time, so the
    first two figures are lower bounds there."""
    cst, work = ctx.trace_frame_counted(cam, opts)  # the same frame: identical counts (deterministic)
    closest = cst.primary_rays + cst.extension_rays
    bytes_frame = (36 * closest + 64 * work[0] + 36 * work[1]) + (48 * cst.shadow_rays + 64 * work[2] + 36 * work[3])
    ser = ctx.trace_frame_serial(cam, opts)
    ctx.reset()
    steps = tm["steps"]
    launches = tm["launches"] / steps
    trace_ms = (tm["stage_ms"]["trace"] + tm["stage_ms"]["trace_last_shadow"]) / steps
    tail_ms = tm["stage_ms"]["tail"] / steps
    tail_launches = tm["chunks"] / steps if tail_ms > 0 else 0  # one k_tail launch per chunk
    bytes_per_launch = bytes_frame / max(launches + tail_launches, 1)
    mean_launch_ms = (trace_ms + tail_ms) / max(launches + tail_launches, 1)
    achieved = bytes_per_launch / (mean_launch_ms * 1e-3) / 1e9
    iso_ms = ser.extend_ms + ser.connect_ms + ser.tail_ms
    iso = bytes_frame / (iso_ms * 1e-3) / 1e9
    dev = bytes_frame / (tm["ms_per_step"] * 1e-3) / 1e9
    out = {
        "kernel": "k_trace (closest hit of bounce b + any hit of bounce b-1, quad BVH)" +
                  ("; k_tail from bounce rr_start" if tail_ms > 0 else ""),
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": measured_traffic(config, traversal),
        "label": "per launch: algorithmic traversal bytes per traversal launch / the launches' mean duration "
                 "over the timed frames (two chunks in flight, so launches share the device)" +
                 ("; the launches are k_trace's and one k_tail per chunk, whose whole duration counts (it also "
                  "shades): a lower bound" if tail_ms > 0 else ""),
        "bytes_per_launch": int(bytes_per_launch),
        "mean_launch_ms": round(mean_launch_ms, 4),
        "launches_per_frame": round(launches + tail_launches, 2),
        "node_tests_per_ray": round(work[0] / max(closest, 1), 2),
        "tri_tests_per_ray": round(work[1] / max(closest, 1), 2),
        "shadow_node_tests_per_ray": round(work[2] / max(cst.shadow_rays, 1), 2),
        "shadow_tri_tests_per_ray": round(work[3] / max(cst.shadow_rays, 1), 2),
        "isolated": {"label": "the same bytes over the traversal launches of a frame whose chunks run one at a "
                              "time (each launch alone on the device)",
                     "achieved": round(iso, 1), "frac": round(iso / HBM_PEAK_GBS, 4),
                     "trace_ms_per_frame": round(iso_ms, 4), "launches": int(ser.extend_launches)},
        "device_level": {"label": "the frame's traversal bytes over the frame wall time (the frame also shades)",
                         "achieved": round(dev, 1), "frac": round(dev / HBM_PEAK_GBS, 4),
                         "bytes_per_frame": int(bytes_frame)},
    }
    if tail_ms > 0:
        out["k_trace_mean_launch_ms"] = round(trace_ms / max(launches, 1), 4)
        out["k_tail_ms_per_frame"] = round(tail_ms, 4)
    return out


def c3_leg(ct, ctx, args, steps: int = 10, warmup: int = 3) -> dict:
    """The 7.24M-triangle displaced sphere at 4K, depth 16, roulette from
    bounce 3, on the same context (trees built on the device: byte-identical
    to the host build)."""
    scene = ct.HostScene.generate(ct.SCENE_C3, args.seed, build=False)
    ctx.upload_build(scene)
    opts, W, H = scene.frame_opts()
    opts.traversal = getattr(ct, "TRAV_" + args.traversal.upper())
    cam = scene.camera(W, H)
    ctx.resize(W, H)
    tm = timed_frames(ctx, cam, opts, steps, warmup)
    out = {"workload": WORKLOADS[3] % (scene.stats()["triangles"], args.traversal),
           "value": round(tm["rays"] / tm["wall"] / 1e6, 2), "unit": "Mrays/s", "steps": steps,
           "ms_per_step": round(tm["ms_per_step"], 3),
           "instrumented_ms_per_step": round(tm["instrumented_ms_per_step"], 3),
           "rays_per_step": {k: tm[k] // steps for k in ("primary", "extension", "shadow")},
           "stage_ms_per_step": {k: round(v / steps, 3) for k, v in tm["stage_ms"].items()},
           "roofline": trace_roofline(ctx, cam, opts, tm, 3, args.traversal)}
    if not args.no_cpu:
        trees = ctx.read_trees()
        arrays = scene.arrays()
        arrays.update(trees)
        out["accuracy"] = accuracy_strip(ctx, None, cam, opts, W, H, strip=16, frames=1, arrays=arrays)
        out["cpu_baseline"] = cpu_baseline(None, cam, opts, W, H, min(args.cpu_seconds, 5.0), strip=16,
                                           arrays=arrays)
    return out


def accuracy_strip(ctx, scene, cam, opts, W, H, strip: int, frames: int = 16, arrays=None) -> dict:
    """Frame 0 of the workload on the GPU against the CPU oracle on a band of
    H / strip rows: bit-identical fraction against the FP32 restatement and
    rel-MSE (mean of (x - ref)^2 / (ref^2 + 1e-2)) against the FP64 one; then
    the same rel-MSE for the mean of `frames` frames (a converged image: a
    sample whose path turns at a texel or sign boundary in FP32 but not in
    FP64 carries a single-sample error of order one, which averages out)."""
    import numpy as np
    from oracle import oracle as orc
    threads = _cpu_threads()
    rows = H // strip
    r0 = (H - rows) // 2  # a band through the middle of the image
    ctx.resize(W, H)
    ctx.trace_frame(cam, opts, stats=False)
    acc, _ = ctx.read_accum()
    osc = _oracle_scene(orc, scene, arrays)
    ref32 = np.zeros((H, W, 4), np.float32)
    ref64 = np.zeros((H, W, 4), np.float64)
    osc.render(cam, opts, W, H, 0, 1, rows=(r0, r0 + rows), threads=threads, accum=ref32)
    osc.render(cam, opts, W, H, 0, 1, rows=(r0, r0 + rows), precision=64, threads=threads, accum=ref64)
    g = acc[r0:r0 + rows]
    same = float(np.mean(g.view(np.uint32) == ref32[r0:r0 + rows].view(np.uint32)))
    x, r = g[..., :3].astype(np.float64), ref64[r0:r0 + rows, :, :3]
    rel = float(np.mean((x - r) ** 2 / (r ** 2 + 1e-2)))
    res = {"rows": f"{r0}..{r0 + rows - 1} of {H}", "frame": 0, "bit_identical_vs_fp32_oracle": round(same, 6),
           "rel_mse_vs_fp64_oracle": float(f"{rel:.3e}"), "nonzero_pixels": int(np.sum(np.any(x > 0, axis=2)))}
    if frames <= 1:
        return res
    # converged: the mean of `frames` accumulated frames against the FP64
    # oracle's mean of the same samples
    for _ in range(frames - 1):
        ctx.trace_frame(cam, opts, stats=False)
    acc, samples = ctx.read_accum()
    assert samples == frames
    ref64[:] = 0
    osc.render(cam, opts, W, H, 0, frames, rows=(r0, r0 + rows), precision=64, threads=threads, accum=ref64)
    xc = acc[r0:r0 + rows, :, :3].astype(np.float64) / frames
    rc = ref64[r0:r0 + rows, :, :3] / frames
    relc = float(np.mean((xc - rc) ** 2 / (rc ** 2 + 1e-2)))
    res["converged"] = {"frames": frames, "rel_mse_vs_fp64_oracle": float(f"{relc:.3e}")}
    return res


def _oracle_scene(orc, scene, arrays=None):
    """The oracle's twin of a scene; given arrays with trees (a device build
    read back), it traverses those instead of rebuilding."""
    if arrays is None:
        return orc.Scene(scene.arrays())
    osc = orc.Scene(arrays, build=False)
    osc.set_trees(arrays["bvh"], arrays["mbvh"], arrays["prim_indices"])
    return osc


def _cpu_threads() -> int:
    """The CPU baseline's threads: OMP_NUM_THREADS if set, else the CPUs this
    process may run on."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def _host_info() -> dict:
    model = ""
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = None
    return {"nproc": os.cpu_count(), "affinity_cpus": aff, "cpu_model": model,
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS")}


def measured_traffic(config: int, traversal: str):
    """HBM bytes per k_trace launch of this workload from the committed
    rocprofv3 summary (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by
    what the counter reads for a plain copy of known size), or None."""
    import re
    f = ROOT / "profiles" / f"k_trace_c{config}.md"
    if not f.exists():
        return None
    m = re.search(r"^traffic_bytes_per_launch %s: (\d+)" % traversal, f.read_text(), re.M)
    return int(m.group(1)) if m else None


def _pct(x, q):
    import numpy as np
    return round(float(np.percentile(np.asarray(x), q)), 3)


def _plain_view(ctx, cam, w: int, h: int, spp: int, opts) -> None:
    """The reference's per-view loop: spp frames of 1 sample, one call each."""
    _fresh(ctx, w, h)
    o = opts.copy()
    o.spp = 1
    for _ in range(spp):
        ctx.trace_frame(cam, o, stats=False)


def _fresh(ctx, w: int, h: int) -> None:
    """Empty accumulator of w x h (reallocated only when the size changes)."""
    if (ctx.width, ctx.height) != (w, h):
        ctx.resize(w, h)
    else:
        ctx.reset()


def serve_requests(ct, scene, base_opts, seed: int, n: int):
    """The served workload: n requests (camera, w, h, spp, opts) on the atrium."""
    import numpy as np
    rng = np.random.default_rng(seed)
    sizes = [(256, 256), (512, 512), (1920, 1080)]
    reqs = []
    for _ in range(n):
        w, h = sizes[rng.choice(3, p=[0.5, 0.35, 0.15])]
        spp = int(rng.choice([1, 4, 16]))
        pos = np.array([-85.0, 14.0, -80.0]) + rng.uniform(-10, 10, 3)
        pos[1] = min(max(pos[1], 3.0), 27.0)
        tgt = np.array([10.0, 9.0, 15.0]) + rng.uniform(-20, 20, 3)
        o = base_opts.copy()
        o.spp = spp
        reqs.append((ct.look_at(pos, tgt, 70.0), w, h, spp, o))
    return reqs


def serve_leg(ct, ctx, scene, opts, seed: int, n: int = 120, rate: float = 40.0, chunk: int = 0) -> dict:
    import numpy as np
    reqs = serve_requests(ct, scene, opts, seed + 1, n)
    arrivals = np.cumsum(np.random.default_rng(seed + 2).exponential(1.0 / rate, n))
    # plain per-view loop of the reference (launchKernels: one frame of 1 spp
    # per call, Kernel.cu:355-372): one request at a time, spp calls, read back
    service, plain_out = [], {}
    for i, (cam, w, h, spp, o) in enumerate(reqs):
        t0 = time.perf_counter()
        _plain_view(ctx, cam, w, h, spp, o)
        acc, _ = ctx.read_accum()
        service.append(time.perf_counter() - t0)
        if w == 256 and len(plain_out) < 8:
            plain_out[i] = acc
    # the same with all spp samples of a request in one call
    t0 = time.perf_counter()
    for cam, w, h, spp, o in reqs:
        _fresh(ctx, w, h)
        ctx.trace_frame(cam, o, stats=False)
        ctx.read_accum()
    multi_ips = n / (time.perf_counter() - t0)
    finish, plain_lat = 0.0, []
    for a, sv in zip(arrivals, service):  # FIFO single server on the measured service times
        finish = max(finish, a) + sv
        plain_lat.append((finish - a) * 1e3)
    plain_ips = n / sum(service)
    srv = ct.Server(ctx, chunk_paths=chunk)
    # Poisson stream
    ids = []
    t0 = time.perf_counter()
    for (cam, w, h, spp, o), a in zip(reqs, arrivals):
        dt = a - (time.perf_counter() - t0)
        if dt > 0:
            time.sleep(dt)
        ids.append(srv.submit(cam, w, h, spp, o))
    lat, same = [], 0
    for i, rid in enumerate(ids):
        res, out = srv.wait(rid)
        lat.append(res.done_ms - res.submit_ms)
        if i in plain_out:
            same += int(np.array_equal(out.view(np.uint32), plain_out[i].view(np.uint32)))
    # saturated: everything queued at once
    c0 = srv.clock_ms()
    ids = [srv.submit(cam, w, h, spp, o) for (cam, w, h, spp, o) in reqs]
    last = max(srv.wait(rid)[0].done_ms for rid in ids)
    srv.close()
    served_ips = n / ((last - c0) * 1e-3)
    return {"requests": n, "rate_per_s": rate, "chunk_paths": chunk or "server default",
            "p50_ms": _pct(lat, 50), "p99_ms": _pct(lat, 99),
            "plain_p50_ms": _pct(plain_lat, 50), "plain_p99_ms": _pct(plain_lat, 99),
            "images_per_s": round(served_ips, 2), "plain_images_per_s": round(plain_ips, 2),
            "speedup": round(served_ips / plain_ips, 3), "plain_multi_spp_images_per_s": round(multi_ips, 2),
            "speedup_vs_multi_spp": round(served_ips / multi_ips, 3),
            "bit_exact_vs_plain": f"{same}/{len(plain_out)}"}


def batched_leg(ct, ctx, seed: int, n_views: int = 64) -> dict:
    import numpy as np
    scene = ct.HostScene.generate(ct.SCENE_C1, seed)
    opts, _, _ = scene.frame_opts()
    opts.max_depth = 8
    rng = np.random.default_rng(seed + 3)
    views = []
    for _ in range(n_views):
        w = int(round(np.exp(rng.uniform(np.log(128), np.log(1024)))))
        h = w if rng.random() < 0.5 else (w * 3) // 4
        spp = int(rng.choice([1, 2, 4, 8, 16]))
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        pos = d * 42.0
        views.append((ct.look_at(pos, rng.uniform(-5, 5, 3), 70.0), w, h, spp, 0))
    ctx.upload(scene)
    tune = ctx.autotune(views[0][0], opts, 256, 256)
    ctx.render_batch(views[:4], opts)  # warm (graph capture, buffers)
    t0 = time.perf_counter()
    outs, st = ctx.render_batch(views, opts)
    t_batch = time.perf_counter() - t0
    # plain per-view loop of the reference (1 spp per call), then the same
    # with all spp of a view in one call
    t_plain, t_multi, same = 0.0, 0.0, 0
    for i, (cam, w, h, spp, fb) in enumerate(views):
        t0 = time.perf_counter()
        _plain_view(ctx, cam, w, h, spp, opts)
        acc, _ = ctx.read_accum()
        t_plain += time.perf_counter() - t0
        if i < 8:
            same += int(np.array_equal(acc.view(np.uint32), outs[i].view(np.uint32)))
        o = opts.copy()
        o.spp = spp
        t0 = time.perf_counter()
        _fresh(ctx, w, h)
        ctx.trace_frame(cam, o, stats=False)
        ctx.read_accum()
        t_multi += time.perf_counter() - t0
    return {"views": n_views, "path_samples": int(sum(w * h * spp for (_, w, h, spp, _) in views)),
            "images_per_s": round(n_views / t_batch, 2), "plain_images_per_s": round(n_views / t_plain, 2),
            "speedup": round(t_plain / t_batch, 3), "plain_multi_spp_images_per_s": round(n_views / t_multi, 2),
            "speedup_vs_multi_spp": round(t_multi / t_batch, 3), "chunks": int(st.chunks),
            "autotuned": {"block": tune.block_size, "stack_lds": tune.stack_lds, "chunk_paths": tune.chunk_paths},
            "bit_exact_vs_plain": f"{same}/8"}


def _kname(name: str) -> str:
    return name.replace("(anonymous namespace)::", "").split("(")[0]


def summarize(paths: list) -> int:
    """Reduce rocprofv3 result databases (rocpd SQLite) to per-kernel
    duration statistics and per-dispatch counter averages."""
    import sqlite3
    for p in paths:
        db = sqlite3.connect(p)
        print(f"# {p}")
        rows = db.execute("select name, count(*), sum(end-start), avg(end-start), min(end-start), max(end-start) "
                          "from kernels group by name order by sum(end-start) desc").fetchall()
        tot = sum(r[2] for r in rows) or 1
        print("| kernel | calls | total ms | avg us | min us | max us | % |")
        print("|---|---|---|---|---|---|---|")
        for r in rows:
            print(f"| {_kname(r[0])} | {r[1]} | {r[2] / 1e6:.3f} | {r[3] / 1e3:.1f} | {r[4] / 1e3:.1f} | "
                  f"{r[5] / 1e3:.1f} | {100 * r[2] / tot:.1f} |")
        # a counter can come as one row per hardware instance (XCD, SE):
        # hardware counters are summed per dispatch (GRBM_GUI_ACTIVE, a clock
        # every instance sees, takes the largest), derived ones averaged
        pmc = db.execute(
            "select k.name, d.counter_name, avg(d.v), count(*) from (select dispatch_id, counter_name, "
            "case when counter_name like 'GRBM%' then max(counter_value) "
            "when counter_name glob '[A-Z][A-Z]*_*' then sum(counter_value) else avg(counter_value) end v "
            "from pmc_events group by dispatch_id, counter_name) d "
            "join kernels k on k.dispatch_id = d.dispatch_id group by k.name, d.counter_name order by k.name").fetchall()
        if pmc:
            print("| kernel | counter | mean per dispatch | dispatches |")
            print("|---|---|---|---|")
            for r in pmc:
                print(f"| {_kname(r[0])} | {r[1]} | {r[2]:.1f} | {r[3]} |")
    return 0


def copy_probe(mib: int) -> int:
    """A plain copy kernel of known size: reads and writes mib MiB, 5 times."""
    import torch
    n = mib * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n * 4, "dispatches": 5}))
    return 0


def display_leg(ctx, W, H, reps: int = 10) -> dict:
    """draw_framebuffer's replacement for a viewer: tonemap the accumulator
    and copy the image to (pageable) host memory, as RGBA32F (the reference's
    surface format) and as RGBA8."""
    import numpy as np
    f = np.empty((H, W, 4), np.float32)
    b = np.empty((H, W, 4), np.uint8)
    res = {"pixels": W * H}
    for name, fn, nbytes in (("rgba32f", lambda: ctx.read_image(), f.nbytes), ("rgba8", lambda: ctx.read_image_rgba8(),
                                                                              b.nbytes)):
        fn()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        ms = (time.perf_counter() - t0) * 1e3 / reps
        res[name] = {"ms_per_image": round(ms, 3), "bytes": nbytes, "gb_per_s": round(nbytes / (ms * 1e-3) / 1e9, 2)}
    return res


def build_leg(ctx, scene, wide: bool, reps: int = 3) -> dict:
    """Device tree build (ct_scene_upload_build) of the workload's scene
    against the host path it replaces: the host BVHTree + MBVHTree builds
    (single thread) followed by ct_scene_upload.  Best of reps device builds;
    the trees read back must equal the host build byte for byte."""
    import numpy as np
    hs = scene.stats()
    t0 = time.perf_counter()
    ctx.upload(scene)
    host_upload_ms = (time.perf_counter() - t0) * 1e3
    best = None
    for _ in range(reps):
        st = ctx.upload_build(scene, wide=wide)
        if best is None or st["device_ms"] < best["device_ms"]:
            best = st
    a = scene.arrays()
    t = ctx.read_trees()
    same = all(np.array_equal(t[k], a[k]) for k in ("bvh", "mbvh", "prim_indices"))
    host_build_ms = (hs["bvh_build_us"] + hs["mbvh_build_us"]) / 1e3
    return {"triangles": hs["triangles"], "bvh_nodes": best["n_bvh_nodes"], "mbvh_nodes": best["n_mbvh_nodes"],
            "levels": best["levels"], "long_swap_chains": best["long_chains"], "jump_rounds": best["jump_rounds"],
            "device_build_ms": round(best["device_ms"], 2),
            "upload_build_ms": round(best["total_ms"], 1),
            "host_build_ms": round(host_build_ms, 1),
            "host_upload_ms": round(host_upload_ms, 1),
            "host_path_ms": round(host_build_ms + host_upload_ms, 1),
            "speedup_build": round(host_build_ms / max(best["device_ms"], 1e-3), 1),
            "speedup_scene_swap": round((host_build_ms + host_upload_ms) / max(best["total_ms"], 1e-3), 1),
            "wide_tree": bool(wide), "identical_to_host": bool(same),
            "note": "host_upload_ms includes the host collapse of the wide tree; upload_build_ms includes "
                    "it only with wide_tree"}


def cpu_build_baseline(scene) -> dict:
    """The oracle's restatement of the reference builder (30 full plane
    scans per node, BVHNode.cu:121-191, then the quad collapse) on one host
    thread, for the same scene."""
    from oracle import oracle as orc
    osc = orc.Scene(scene.arrays(), build=False)
    t0 = time.perf_counter()
    if orc.lib.or_scene_build(osc._h) != 0:
        raise RuntimeError("oracle build failed")
    dt = time.perf_counter() - t0
    return {"value": round(dt * 1e3, 1), "unit": "ms per build", "cores": 1, "kind": "port",
            "sample": f"one build of the {scene.stats()['triangles']}-triangle scene"}


def cpu_baseline(scene, cam, opts, W, H, min_seconds: float, max_frames: int = 256, strip: int = 1,
                 arrays=None) -> dict:
    """Oracle FP32 on the benchmark frames (frame indices 0, 1, ...) until at
    least min_seconds of CPU work, at most max_frames frames.  strip > 1
    renders a band of H / strip rows through the middle of each frame only
    (a bounded sample of the 4K workload)."""
    import numpy as np
    from oracle import oracle as orc
    threads = _cpu_threads()
    osc = _oracle_scene(orc, scene, arrays)
    rows = H // strip
    r0 = (H - rows) // 2  # a band through the middle of the image
    acc = np.zeros((H, W, 4), np.float32)
    rays = frames = 0
    t0 = time.perf_counter()
    while frames < max_frames:
        _, st = osc.render(cam, opts, W, H, frames, 1, rows=(r0, r0 + rows), threads=threads, accum=acc)
        rays += st["primary_rays"] + st["extension_rays"] + st["shadow_rays"]
        frames += 1
        if time.perf_counter() - t0 >= min_seconds:
            break
    dt = time.perf_counter() - t0
    what = f"full {W}x{H} frames" if strip == 1 else f"{W}x{rows} strips (rows {r0}..{r0 + rows - 1}) of {W}x{H} frames"
    return {"value": round(rays / dt / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"{frames} {what} at 1 spp (frame indices 0..{frames - 1}), {rays} rays, "
                      f"{dt:.1f} s on {threads} OpenMP threads; oracle built -O3 -march=x86-64-v3 without fast-math",
            "host": _host_info()}


if __name__ == "__main__":
    sys.exit(main())
