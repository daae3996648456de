"""GPU parity: the HIP kernels (through the C ABI) against the CPU oracle.

Bars (exact traversal): closest-hit primitive ids and hit distances
bit-identical; shadow (any-hit) flags identical; per-pixel FP32 radiance
within 1e-4 relative (observed: bit-identical) and relative MSE against the
FP64 oracle <= 1e-3.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import ray_set, surface_rays

pytestmark = pytest.mark.gpu

N_RAYS = 1 << 20


def _upload(gpu_ctx, scene):
    gpu_ctx.upload(scene)


@pytest.mark.parametrize("config,accel", [(1, 0), (1, 1), (2, 1)])
def test_closest_hit_bit_exact(ct, orc, gpu_ctx, config, accel):
    from conftest import get_scene
    scene, arrays, osc = get_scene(ct, orc, config)
    _upload(gpu_ctx, scene)
    o, d = ray_set(scene, arrays, N_RAYS, seed=100 + config)
    opts, _, _ = scene.frame_opts()
    opts.accel = accel
    opts.traversal = ct.TRAV_EXACT
    t_gpu, p_gpu = gpu_ctx.trace_rays(o, d, opts)
    t_ref, p_ref = osc.trace(o, d, accel, 32, opts.triangle_epsilon, threads=16)
    t_ref32 = t_ref.astype(np.float32)
    mismatch = np.flatnonzero(p_gpu != p_ref)
    print(f"config {config} accel {accel}: rays {N_RAYS}, hits {int(np.sum(p_ref >= 0))}, "
          f"id mismatches {mismatch.size}")
    assert mismatch.size == 0, f"first mismatches {mismatch[:8]} gpu {p_gpu[mismatch[:8]]} ref {p_ref[mismatch[:8]]}"
    assert np.array_equal(t_gpu.view(np.uint32), t_ref32.view(np.uint32))


def test_closest_hit_matches_brute_force_c1(ct, orc, gpu_ctx, c1):
    """Against exhaustive search: same t everywhere, ids equal except exact ties."""
    scene, arrays, osc = c1
    _upload(gpu_ctx, scene)
    o, d = ray_set(scene, arrays, 1 << 14, seed=5)
    opts, _, _ = scene.frame_opts()
    opts.accel = ct.ACCEL_BVH4
    t_gpu, p_gpu = gpu_ctx.trace_rays(o, d, opts)
    t_bf, p_bf = osc.trace(o, d, 2, 32, opts.triangle_epsilon, threads=16)
    assert np.array_equal(t_gpu, t_bf.astype(np.float32))
    ties = np.flatnonzero(p_gpu != p_bf)
    print("exact ties (same t, different id):", ties.size)
    assert ties.size <= 4


@pytest.mark.parametrize("config", [1, 2])
@pytest.mark.parametrize("mode", ["fast", "compact"])
def test_fast_traversal_closest_hit(ct, orc, gpu_ctx, config, mode):
    """The wide compressed tree (CT_TRAV_FAST) and the compact quad tree
    (CT_TRAV_COMPACT, 64-byte nodes with 8-bit child boxes) against the
    oracle's exact traversal: the same triangle records are tested with the
    same arithmetic, so t is bit-identical wherever both find the same nearest
    surface; ids differ only on exact t ties.  The few rays whose t differs
    are ones where the exact traversal's tagged tmin (the reference's low-bit
    slot tag) culls a box within ulps of the current hit; there the
    difference is a few ulps."""
    from conftest import get_scene
    scene, arrays, osc = get_scene(ct, orc, config)
    _upload(gpu_ctx, scene)
    o, d = ray_set(scene, arrays, N_RAYS, seed=200 + config)
    opts, _, _ = scene.frame_opts()
    opts.accel = ct.ACCEL_BVH4
    opts.traversal = ct.TRAV_FAST if mode == "fast" else ct.TRAV_COMPACT
    t_gpu, p_gpu = gpu_ctx.trace_rays(o, d, opts)
    t_ref, p_ref = osc.trace(o, d, 1, 32, opts.triangle_epsilon, threads=16)
    t_ref32 = t_ref.astype(np.float32)
    same_t = t_gpu.view(np.uint32) == t_ref32.view(np.uint32)
    diff_t = np.flatnonzero(~same_t)
    ties = np.flatnonzero(same_t & (p_gpu != p_ref))
    print(f"config {config} {mode}: rays {N_RAYS}, t differs {diff_t.size}, id ties {ties.size}")
    assert diff_t.size <= N_RAYS // 100000
    if diff_t.size:
        rel = np.abs(t_gpu[diff_t] - t_ref32[diff_t]) / t_ref32[diff_t]
        assert np.all(rel <= 1e-6), rel.max()
    assert ties.size <= N_RAYS // 10000


@pytest.mark.parametrize("config", [1, 2])
@pytest.mark.parametrize("mode", ["fast", "compact"])
def test_fast_traversal_any_hit(ct, orc, gpu_ctx, config, mode):
    """Any-hit over the wide tree and over the compact quad tree equals the
    exact any-hit (order-free)."""
    from conftest import get_scene
    scene, arrays, osc = get_scene(ct, orc, config)
    _upload(gpu_ctx, scene)
    n = 1 << 18
    o, d = ray_set(scene, arrays, n, seed=300 + config)
    rng = np.random.default_rng(config)
    tmax = rng.uniform(0.0, 60.0, n).astype(np.float32)
    opts, _, _ = scene.frame_opts()
    opts.accel = ct.ACCEL_BVH4
    opts.traversal = ct.TRAV_FAST if mode == "fast" else ct.TRAV_COMPACT
    occ_gpu = gpu_ctx.trace_shadow_rays(o, d, tmax, opts)
    occ_ref = osc.trace_shadow(o, d, tmax, 1, 32, opts.triangle_epsilon, threads=16)
    bad = int(np.sum(occ_gpu != occ_ref))
    print(f"config {config} {mode}: occluded {int(occ_ref.sum())} of {n}, differ {bad}")
    assert bad <= n // 100000


def test_compact_frame_c2(ct, gpu_ctx, c2):
    """A C2 frame (depth 8, NEE, typed BSDFs) over the compact quad tree
    equals the exact traversal's frame except at the few pixels whose path
    met a t-tie or a ulp-level t difference; the compact traversal needs the
    quad tree (refused for the binary one)."""
    scene, _, _ = c2
    _upload(gpu_ctx, scene)
    opts, w, h = scene.frame_opts()
    cam = scene.camera(w, h)
    gpu_ctx.resize(w, h)
    gpu_ctx.trace_frame(cam, opts)
    ref = gpu_ctx.read_accum()[0].copy()
    o = opts.copy()
    o.traversal = ct.TRAV_COMPACT
    gpu_ctx.reset()
    gpu_ctx.trace_frame(cam, o)
    got = gpu_ctx.read_accum()[0]
    same = np.all(got.view(np.uint32) == ref.view(np.uint32), axis=-1)
    print(f"C2 compact frame: {int(same.sum())} of {same.size} pixels bit-identical to exact")
    assert same.mean() >= 0.999
    rel = float(np.mean((got[..., :3] - ref[..., :3]) ** 2) / np.mean(ref[..., :3] ** 2))
    assert rel < 1e-3, rel
    o.accel = ct.ACCEL_BVH2
    with pytest.raises(ct.CTError):
        gpu_ctx.trace_frame(cam, o)


def test_c2_full_frame_same_under_every_schedule(ct, gpu_ctx, c2):
    """The headline frame (config 2, 1920x1080, depth 8) in full: a band of
    eight rows through its middle bit-identical to the FP32 oracle, and the
    whole frame bit-identical, with equal primary, extension and shadow ray
    counts, under every way the library schedules it: two chunks on two
    lanes with and without timing events, chunks one at a time
    (ct_trace_frame_serial), the instrumented counting frame
    (ct_trace_frame_counted), shading without the block-local material sort,
    and a context whose chunks hold 2^18 paths (eight chunks, four per
    lane)."""
    scene, _, osc = c2
    _upload(gpu_ctx, scene)
    opts, w, h = scene.frame_opts()
    assert (w, h, opts.max_depth, opts.sort_by_material) == (1920, 1080, 8, 1)
    cam = scene.camera(w, h)
    gpu_ctx.resize(w, h)

    def counts(st):
        return (st.primary_rays, st.extension_rays, st.shadow_rays)

    gpu_ctx.reset()
    st0 = gpu_ctx.trace_frame(cam, opts)
    ref = gpu_ctx.read_accum()[0].copy()
    assert st0.primary_rays == w * h and st0.chunks == 2
    rows, r0 = 8, (h - 8) // 2
    band = np.zeros((h, w, 4), np.float32)
    osc.render(cam, opts, w, h, 0, 1, rows=(r0, r0 + rows), threads=16, accum=band)
    assert np.array_equal(ref[r0:r0 + rows].view(np.uint32), band[r0:r0 + rows].view(np.uint32))
    print(f"C2 full frame: rays {counts(st0)}; rows {r0}..{r0 + rows - 1} bit-identical to the oracle", flush=True)

    def same(label, ctx, st):
        acc, s = ctx.read_accum()
        diff = int(np.sum(acc.view(np.uint32) != ref.view(np.uint32)))
        print(f"C2 full frame, {label}: rays {counts(st)}, words differing {diff}", flush=True)
        assert s == 1 and diff == 0 and counts(st) == counts(st0)

    gpu_ctx.reset()
    gpu_ctx.trace_frame(cam, opts, stats=False)
    assert np.array_equal(gpu_ctx.read_accum()[0].view(np.uint32), ref.view(np.uint32))
    gpu_ctx.reset()
    same("chunks one at a time", gpu_ctx, gpu_ctx.trace_frame_serial(cam, opts))
    gpu_ctx.reset()
    same("counting frame", gpu_ctx, gpu_ctx.trace_frame_counted(cam, opts)[0])
    o = opts.copy()
    o.sort_by_material = 0
    gpu_ctx.reset()
    same("unsorted shading", gpu_ctx, gpu_ctx.trace_frame(cam, o))
    eight = ct.Context(0, chunk_paths=1 << 18)
    try:
        eight.upload(scene)
        eight.resize(w, h)
        st = eight.trace_frame(cam, opts)
        assert st.chunks == 8
        same("8 chunks", eight, st)
    finally:
        eight.close()


@pytest.mark.parametrize("config,accel", [(1, 0), (1, 1), (2, 1)])
def test_shadow_any_hit(ct, orc, gpu_ctx, config, accel):
    from conftest import get_scene
    scene, arrays, osc = get_scene(ct, orc, config)
    _upload(gpu_ctx, scene)
    n = 1 << 18
    o, _ = surface_rays(arrays, n, seed=11)
    rng = np.random.default_rng(12)
    # targets: random points on light triangles or random surface points
    lidx = arrays["light_indices"]
    tri = lidx[rng.integers(0, lidx.size, n)]
    V = arrays["vertices"].astype(np.float64)
    I = arrays["indices"]
    r1, r2 = rng.random(n), rng.random(n)
    f = r1 + r2 > 1
    r1[f], r2[f] = 1 - r1[f], 1 - r2[f]
    tgt = (1 - r1 - r2)[:, None] * V[I[tri, 0]] + r1[:, None] * V[I[tri, 1]] + r2[:, None] * V[I[tri, 2]]
    dv = tgt - o
    dist = np.linalg.norm(dv, axis=1)
    d = (dv / dist[:, None]).astype(np.float32)
    tmax = (dist - 2e-3).astype(np.float32)
    opts, _, _ = scene.frame_opts()
    opts.accel = accel
    occ_gpu = gpu_ctx.trace_shadow_rays(o, d, tmax, opts)
    occ_ref = osc.trace_shadow(o, d, tmax, accel, 32, opts.triangle_epsilon, threads=16)
    print(f"config {config} accel {accel}: occluded {int(occ_ref.sum())} of {n}")
    assert np.array_equal(occ_gpu, occ_ref)


def _relmse(x, ref):
    return float(np.mean((x - ref) ** 2 / (ref ** 2 + 1e-2)))


def test_radiance_c1_16_frames(ct, orc, gpu_ctx, c1):
    """Config 1: 256x256, 1 spp x 16 frames, depth 4, NEE, binary tree."""
    scene, arrays, osc = c1
    _upload(gpu_ctx, scene)
    opts, w, h = scene.frame_opts()
    cam = scene.camera(w, h)
    gpu_ctx.resize(w, h)
    for _ in range(16):
        gpu_ctx.trace_frame(cam, opts)
    acc, samples = gpu_ctx.read_accum()
    assert samples == 16
    ref = np.zeros((h, w, 4), np.float32)
    for f in range(16):
        osc.render(cam, opts, w, h, f, 1, threads=16, accum=ref)
    diff = np.abs(acc - ref)
    rel = diff / np.maximum(np.abs(ref), 1e-30)
    bad = int(np.sum((rel > 1e-4) & (diff > 1e-7)))
    exact = int(np.sum(acc.view(np.uint32) == ref.view(np.uint32)))
    print(f"C1 16 frames: pixels over 1e-4 rel {bad}, bit-identical values {exact}/{acc.size}")
    assert bad == 0
    ref64, _ = osc.render(cam, opts, w, h, 0, 16, precision=64, threads=16)
    img = acc[..., :3] / acc[..., 3:]
    img64 = ref64[..., :3] / ref64[..., 3:]
    r = _relmse(img, img64)
    print(f"C1 rel-MSE vs FP64 oracle: {r:.3e}")
    assert r <= 1e-3


def test_radiance_c2_typed_materials(ct, orc, gpu_ctx, c2):
    """Config 2 scene (GGX, mirror, dielectric, Lambert, sky) at reduced size."""
    scene, arrays, osc = c2
    _upload(gpu_ctx, scene)
    opts, _, _ = scene.frame_opts()
    w, h = 192, 108
    opts.spp = 2
    cam = scene.camera(w, h)
    gpu_ctx.resize(w, h)
    gpu_ctx.trace_frame(cam, opts)
    acc, _ = gpu_ctx.read_accum()
    ref, _ = osc.render(cam, opts, w, h, 0, 2, threads=16)
    diff = np.abs(acc - ref)
    rel = diff / np.maximum(np.abs(ref), 1e-30)
    bad = int(np.sum((rel > 1e-4) & (diff > 1e-7)))
    print(f"C2 typed: pixels over 1e-4 rel {bad} of {acc.size}")
    assert bad == 0
    ref64, _ = osc.render(cam, opts, w, h, 0, 2, precision=64, threads=16)
    r = _relmse(acc[..., :3] / acc[..., 3:], ref64[..., :3] / ref64[..., 3:])
    print(f"C2 rel-MSE vs FP64 oracle: {r:.3e}")
    assert r <= 1e-3


def test_batch_equals_single_views(ct, orc, gpu_ctx, c1):
    scene, arrays, osc = c1
    _upload(gpu_ctx, scene)
    opts, _, _ = scene.frame_opts()
    opts.max_depth = 8
    views = []
    rng = np.random.default_rng(3)
    for i, (w, h, spp) in enumerate([(96, 72, 2), (64, 64, 1), (160, 120, 4), (33, 17, 3)]):
        ang = rng.uniform(0, 2 * np.pi)
        pos = [30 * np.cos(ang), rng.uniform(-20, 20), 30 * np.sin(ang)]
        cam = ct.look_at(pos, [0, 0, 0], 70.0)
        views.append((cam, w, h, spp, 5 * i))
    outs, st = gpu_ctx.render_batch(views, opts)
    for (cam, w, h, spp, fb), out in zip(views, outs):
        gpu_ctx.resize(w, h)
        o2 = opts.copy()
        o2.spp = 1
        acc = np.zeros((h, w, 4), np.float32)
        # single-view reference: ct_trace_frame with frame indices fb .. fb+spp-1
        gpu_ctx.write_accum(acc, fb)
        o2.spp = spp
        gpu_ctx.trace_frame(cam, o2)
        single, _ = gpu_ctx.read_accum()
        assert np.array_equal(out, single), f"view {w}x{h} spp {spp} differs"
    print("batch stats", st.primary_rays, st.extension_rays, st.shadow_rays, st.chunks)


def test_batch_chunked_graph_replay(ct, orc, gpu_ctx, c1):
    """Many chunks through the captured graph equal one big chunk."""
    scene, arrays, osc = c1
    _upload(gpu_ctx, scene)
    opts, _, _ = scene.frame_opts()
    cam = scene.camera(128, 128)
    views = [(cam, 128, 128, 8, 0)]
    outs_small = None
    small = ct.Context(0, chunk_paths=1 << 14)
    try:
        small.upload(scene)
        outs_small, st = small.render_batch(views, opts)
        assert st.chunks == 8
    finally:
        small.close()
    outs_big, _ = gpu_ctx.render_batch(views, opts)
    assert np.array_equal(outs_small[0], outs_big[0])
    ref, _ = osc.render(cam, opts, 128, 128, 0, 8, threads=16)
    assert np.array_equal(outs_big[0], ref)


def test_odd_sizes_split_pixels_vs_oracle(ct, orc, c1):
    """Views whose height is not a multiple of the pixel-tile height (a short
    last strip), one-pixel-wide and three-pixel-high views, and a chunk size
    that splits a pixel's samples over two chunks: every accumulator equals
    the oracle's bit for bit (ids, counts and sums)."""
    scene, arrays, osc = c1
    opts, _, _ = scene.frame_opts()
    sizes = [(37, 23, 3), (33, 17, 2), (5, 3, 1), (1, 7, 2), (64, 6, 5)]
    views = [(scene.camera(w, h), w, h, spp, 3 * i) for i, (w, h, spp) in enumerate(sizes)]
    ctx = ct.Context(0, chunk_paths=2048)
    try:
        ctx.upload(scene)
        outs, st = ctx.render_batch(views, opts)
    finally:
        ctx.close()
    assert st.chunks >= 2
    for (cam, w, h, spp, fb), out in zip(views, outs):
        ref, _ = osc.render(cam, opts, w, h, fb, spp, threads=16)
        assert np.array_equal(out, ref), f"view {w}x{h} spp {spp}"


@pytest.mark.parametrize("config", [1, 2])
def test_material_sort_same_image(ct, orc, gpu_ctx, config):
    """Shading in block-local material order (sort_by_material) changes which
    lane shades a path, not the path: accumulators are bit-identical."""
    from conftest import get_scene
    scene, arrays, osc = get_scene(ct, orc, config)
    _upload(gpu_ctx, scene)
    opts, w, h = scene.frame_opts()
    w, h = min(w, 480), min(h, 270)
    cam = scene.camera(w, h)
    gpu_ctx.resize(w, h)
    out = []
    for sort in (0, 1):
        o = opts.copy()
        o.sort_by_material = sort
        o.spp = 2
        gpu_ctx.reset()
        gpu_ctx.trace_frame(cam, o)
        out.append(gpu_ctx.read_accum()[0])
    assert np.array_equal(out[0].view(np.uint32), out[1].view(np.uint32))


def test_radiance_beckmann_roulette(ct, orc, gpu_ctx):
    """The box-and-sphere scene with every surface a Beckmann microfacet
    (roughness uniform in [0.02, 0.5], BeckmannSample11's Newton/bisection
    inversion, Microfacet.cuh:141-230), typed BSDFs, depth 16, Russian
    roulette from bounce 1, 2 samples per pixel: per pixel within 1e-4 of
    the FP32 oracle (observed: bit-identical) and rel-MSE <= 1e-3 vs FP64."""
    import ctypes as C_
    scene = ct.HostScene.generate(ct.SCENE_C1, 7)
    arrays = scene.arrays()
    mats = arrays["materials"]
    rng = np.random.default_rng(5)
    gpu_ctx.upload(scene)
    n_beck = 0
    for i in range(len(mats)):
        if mats[i].type == ct.MAT_LIGHT:
            continue
        mats[i].type = ct.MAT_BECKMANN
        mats[i].roughness = float(rng.uniform(0.02, 0.5))
        m = ct.Material()
        C_.memmove(C_.byref(m), C_.byref(mats[i]), C_.sizeof(m))
        gpu_ctx.material_update(i, m)
        n_beck += 1
    assert n_beck >= 3
    osc = orc.Scene(arrays)
    opts, _, _ = scene.frame_opts()
    opts.bsdf_mode = ct.BSDF_TYPED
    opts.max_depth = 16
    opts.rr_start = 1
    opts.spp = 2
    w = h = 128
    cam = scene.camera(w, h)
    gpu_ctx.resize(w, h)
    st = gpu_ctx.trace_frame(cam, opts)
    acc, _ = gpu_ctx.read_accum()
    ref, ost = osc.render(cam, opts, w, h, 0, 2, threads=16)
    assert (st.primary_rays, st.extension_rays, st.shadow_rays) == (
        ost["primary_rays"], ost["extension_rays"], ost["shadow_rays"])
    assert st.tail_ms > 0.0  # bounces 5-15 ran in the roulette-tail launch
    diff = np.abs(acc - ref)
    bad = int(np.sum((diff / np.maximum(np.abs(ref), 1e-30) > 1e-4) & (diff > 1e-7)))
    same = float(np.mean(acc.view(np.uint32) == ref.view(np.uint32)))
    print(f"Beckmann + roulette: values over 1e-4 {bad}, bit-identical fraction {same}, "
          f"extension rays {st.extension_rays} for {st.primary_rays} paths")
    assert bad == 0
    ref64, _ = osc.render(cam, opts, w, h, 0, 2, precision=64, threads=16)
    r = _relmse(acc[..., :3] / acc[..., 3:], ref64[..., :3] / ref64[..., 3:])
    print(f"rel-MSE vs FP64 {r:.3e}")
    assert r <= 1e-3


def test_batch_depth8_vs_oracle(ct, orc, gpu_ctx, c1):
    """ct_render_batch at the batched workload's depth (8) against the FP32
    oracle, views of uneven size and spp with frame bases: bit-identical."""
    scene, arrays, osc = c1
    gpu_ctx.upload(scene)
    opts, _, _ = scene.frame_opts()
    opts.max_depth = 8
    rng = np.random.default_rng(8)
    views = []
    for i, (w, h, spp) in enumerate([(128, 96, 4), (160, 160, 2), (96, 72, 16), (200, 150, 1)]):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        cam = ct.look_at(d * 42.0, rng.uniform(-5, 5, 3), 70.0)
        views.append((cam, w, h, spp, 3 * i))
    outs, st = gpu_ctx.render_batch(views, opts)
    for (cam, w, h, spp, fb), out in zip(views, outs):
        ref, _ = osc.render(cam, opts, w, h, fb, spp, threads=16)
        assert np.array_equal(out.view(np.uint32), ref.view(np.uint32)), (w, h, spp)


@pytest.mark.parametrize("mode", ["exact", "compact", "fast"])
def test_roulette_tail_each_quad_traversal(ct, orc, gpu_ctx, c2, mode):
    """The roulette tail (k_tail) over each traversal of the quad tree's
    family on the atrium (depth 12, roulette from bounce 2, 320x180, 2 spp):
    with the exact traversal the frame is bit-identical to the FP32 oracle
    and its ray counts are the oracle's; the compact quad tree and the wide
    tree give the exact traversal's frame except at the few pixels whose path
    met a t-tie or a ulp-level t difference."""
    scene, _, osc = c2
    _upload(gpu_ctx, scene)
    opts, _, _ = scene.frame_opts()
    opts.accel = ct.ACCEL_BVH4
    opts.max_depth = 12
    opts.rr_start = 2
    opts.spp = 2
    w, h = 320, 180
    cam = scene.camera(w, h)
    gpu_ctx.resize(w, h)
    st = gpu_ctx.trace_frame(cam, opts)
    exact = gpu_ctx.read_accum()[0].copy()
    assert st.tail_ms > 0.0  # bounces 2-11 ran in the roulette-tail launch
    if mode == "exact":
        ref, ost = osc.render(cam, opts, w, h, 0, 2, threads=16)
        assert np.array_equal(exact.view(np.uint32), ref.view(np.uint32))
        assert (st.primary_rays, st.extension_rays, st.shadow_rays) == (
            ost["primary_rays"], ost["extension_rays"], ost["shadow_rays"])
        return
    o = opts.copy()
    o.traversal = ct.TRAV_FAST if mode == "fast" else ct.TRAV_COMPACT
    gpu_ctx.reset()
    st2 = gpu_ctx.trace_frame(cam, o)
    got = gpu_ctx.read_accum()[0]
    assert st2.tail_ms > 0.0
    same = np.all(got.view(np.uint32) == exact.view(np.uint32), axis=-1)
    print(f"roulette tail, {mode} traversal: {int(same.sum())} of {same.size} pixels bit-identical to exact")
    assert same.mean() >= 0.999
    rel = float(np.mean((got[..., :3] - exact[..., :3]) ** 2) / np.mean(exact[..., :3] ** 2))
    assert rel < 1e-3, rel


@pytest.mark.parametrize("rr_start", [0, 2])
def test_roulette_tail_batch_vs_oracle(ct, orc, gpu_ctx, c1, rr_start):
    """The roulette tail (k_tail: a lane pair per path from bounce rr_start
    on) under ct_render_batch, whose chunks mix views of uneven size and spp:
    bit-identical to the FP32 oracle, with the ray counts equal to the
    oracle's (the batch runs untimed, from its captured graph; the tail's
    launch time is checked in test_radiance_beckmann_roulette).  rr_start 0
    puts the tail right after the primary rays' trace, so k_tail shades every
    bounce."""
    scene, arrays, osc = c1
    gpu_ctx.upload(scene)
    opts, _, _ = scene.frame_opts()
    opts.max_depth = 10
    opts.rr_start = rr_start
    rng = np.random.default_rng(10 + rr_start)
    views = []
    for i, (w, h, spp) in enumerate([(120, 90, 3), (64, 64, 8), (150, 100, 1)]):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        cam = ct.look_at(d * 42.0, rng.uniform(-5, 5, 3), 70.0)
        views.append((cam, w, h, spp, 5 * i))
    outs, st = gpu_ctx.render_batch(views, opts)
    rays = [0, 0, 0]
    for (cam, w, h, spp, fb), out in zip(views, outs):
        ref, ost = osc.render(cam, opts, w, h, fb, spp, threads=16)
        assert np.array_equal(out.view(np.uint32), ref.view(np.uint32)), (w, h, spp)
        rays[0] += ost["primary_rays"]
        rays[1] += ost["extension_rays"]
        rays[2] += ost["shadow_rays"]
    assert (st.primary_rays, st.extension_rays, st.shadow_rays) == tuple(rays)


@pytest.mark.parametrize("config,size", [(1, (256, 256)), (2, (480, 270))])
def test_ray_and_work_counts_equal_oracle(ct, orc, gpu_ctx, config, size):
    """The counts behind the benchmark's numbers: primary, extension and
    shadow rays of a frame (ct_frame_stats, from the device queue counters)
    and the closest-hit / any-hit node and triangle tests
    (ct_trace_frame_counted) equal the oracle's counts of the same frame
    exactly: the kernels visit the children of both kinds of rays in the
    reference's order (MBVHNode.cuh:126-205), as the oracle does."""
    from conftest import get_scene
    scene, arrays, osc = get_scene(ct, orc, config)
    gpu_ctx.upload(scene)
    opts, _, _ = scene.frame_opts()
    w, h = size
    cam = scene.camera(w, h)
    gpu_ctx.resize(w, h)
    st, work = gpu_ctx.trace_frame_counted(cam, opts)
    _, ost = osc.render(cam, opts, w, h, 0, 1, threads=16)
    got = {"primary_rays": st.primary_rays, "extension_rays": st.extension_rays, "shadow_rays": st.shadow_rays,
           "node_tests_closest": work[0], "tri_tests_closest": work[1], "node_tests_shadow": work[2],
           "tri_tests_shadow": work[3]}
    want = {k: ost[k] for k in got}
    print(f"config {config}: {got}")
    assert got == want
    # the plain frame counts the same rays
    gpu_ctx.reset()
    st2 = gpu_ctx.trace_frame(cam, opts)
    assert (st2.primary_rays, st2.extension_rays, st2.shadow_rays) == (
        st.primary_rays, st.extension_rays, st.shadow_rays)
