"""GPU tests of the runtime around the kernels: the render server, checkpoint
and resume of the accumulator, material edits, the autotuner, the tonemap
read-out and error returns.  Results are compared with ct_trace_frame of the
same samples (bit-exact: the path samples do not depend on how they are
batched) or with the CPU oracle.
"""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def _plain(ctx, cam, w, h, opts, spp, frame_base=0):
    """ct_trace_frame of spp samples from frame index frame_base."""
    ctx.resize(w, h)
    if frame_base:
        ctx.write_accum(np.zeros((h, w, 4), np.float32), frame_base)
    o = opts.copy()
    o.spp = spp
    ctx.trace_frame(cam, o)
    acc, _ = ctx.read_accum()
    return acc


def test_server_requests_equal_plain_frames(ct, gpu_ctx, c2):
    """Requests of mixed size and spp submitted together (so they share
    wavefront chunks) give the accumulators of separate plain frames."""
    scene, _, _ = c2
    gpu_ctx.upload(scene)
    opts, _, _ = scene.frame_opts()
    reqs = [(ct.look_at((-85 + 3 * i, 14, -80), (10, 9 - i, 15), 70.0), w, h, spp)
            for i, (w, h, spp) in enumerate([(64, 48, 4), (160, 90, 1), (96, 96, 16), (33, 17, 3), (128, 128, 2)])]
    expect = [_plain(gpu_ctx, cam, w, h, opts, spp) for (cam, w, h, spp) in reqs]
    srv = ct.Server(gpu_ctx, chunk_paths=1 << 14)
    try:
        ids = []
        for cam, w, h, spp in reqs:
            o = opts.copy()
            o.spp = spp
            ids.append(srv.submit(cam, w, h, spp, o))
        for rid, e, (_, w, h, spp) in zip(ids, expect, reqs):
            res, out = srv.wait(rid)
            assert res.status == 0 and (res.width, res.height, res.spp) == (w, h, spp)
            assert res.done_ms >= res.admit_ms >= res.submit_ms
            assert np.array_equal(_bits(out), _bits(e)), (w, h, spp)
    finally:
        srv.close()


def test_server_cancel_and_admission(ct, gpu_ctx, c1):
    scene, _, _ = c1
    gpu_ctx.upload(scene)
    opts, w, h = scene.frame_opts()
    cam = scene.camera(w, h)
    srv = ct.Server(gpu_ctx, mem_budget=64 << 20, chunk_paths=1 << 12)
    try:
        with pytest.raises(ct.CTError) as e:  # 4096^2 float4 accumulator > 64 MiB
            srv.submit(cam, 4096, 4096, 1, opts)
        assert e.value.code == ct.CT_ERR_ADMISSION
        big = srv.submit(cam, 1024, 1024, 16, opts)
        small = srv.submit(cam, 64, 64, 1, opts)
        srv.cancel(big)
        res, _ = srv.wait(big)
        assert res.status in (0, ct.CT_ERR_CANCELLED)
        res, out = srv.wait(small)
        assert res.status == 0
        assert np.array_equal(_bits(out), _bits(_plain(gpu_ctx, cam, 64, 64, opts, 1)))
    finally:
        srv.close()


def test_checkpoint_resume(ct, gpu_ctx, c1):
    """4 frames, read back, restore into a fresh context, 4 more frames ==
    8 frames in one go (frame indices continue from the sample count)."""
    scene, _, _ = c1
    gpu_ctx.upload(scene)
    opts, w, h = scene.frame_opts()
    cam = scene.camera(w, h)
    gpu_ctx.resize(w, h)
    for _ in range(8):
        gpu_ctx.trace_frame(cam, opts)
    full, n = gpu_ctx.read_accum()
    assert n == 8
    gpu_ctx.reset()
    for _ in range(4):
        gpu_ctx.trace_frame(cam, opts)
    half, n = gpu_ctx.read_accum()
    other = ct.Context(0)
    try:
        other.upload(scene)
        other.resize(w, h)
        other.write_accum(half, n)
        for _ in range(4):
            other.trace_frame(cam, opts)
        resumed, n2 = other.read_accum()
    finally:
        other.close()
    assert n2 == 8
    assert np.array_equal(_bits(resumed), _bits(full))


def test_material_update_matches_oracle(ct, orc, gpu_ctx):
    """Editing a material on the device (Application::updateMaterials) gives
    the oracle's render of the edited scene."""
    from oracle import oracle as orm
    scene = ct.HostScene.generate(ct.SCENE_C1, 7)
    opts, _, _ = scene.frame_opts()
    w = h = 64
    cam = scene.camera(w, h)
    gpu_ctx.upload(scene)
    gpu_ctx.resize(w, h)
    arrays = scene.arrays()
    mats = arrays["materials"]
    idx = int(arrays["material_idx"][-1])  # the sphere's material
    m = ct.Material()
    C_ = __import__("ctypes")
    C_.memmove(C_.byref(m), C_.byref(mats[idx]), C_.sizeof(m))
    m.albedo[0], m.albedo[1], m.albedo[2] = 0.9, 0.1, 0.2
    gpu_ctx.material_update(idx, m)
    gpu_ctx.trace_frame(cam, opts)
    acc, _ = gpu_ctx.read_accum()
    C_.memmove(C_.byref(mats[idx]), C_.byref(m), C_.sizeof(m))
    osc = orm.Scene(arrays)
    ref, _ = osc.render(cam, opts, w, h, 0, 1, threads=16)
    assert np.array_equal(_bits(acc), _bits(ref))


def test_autotune_keeps_results(ct, gpu_ctx, c2):
    scene, _, _ = c2
    gpu_ctx.upload(scene)
    opts, _, _ = scene.frame_opts()
    w, h = 320, 180
    cam = scene.camera(w, h)
    before = _plain(gpu_ctx, cam, w, h, opts, 1)
    r = gpu_ctx.autotune(cam, opts, 256, 144)
    print("autotune:", r.block_size, r.stack_lds, r.chunk_paths, r.best_ms, r.trials)
    assert r.trials > 1 and r.best_ms > 0
    assert r.block_size in (64, 128, 256, 512, 1024) and r.stack_lds > 0
    after = _plain(gpu_ctx, cam, w, h, opts, 1)
    assert np.array_equal(_bits(before), _bits(after))


def test_read_image_tonemap(ct, gpu_ctx, c1):
    """draw_framebuffer (Kernel.cu:258-270): rgb / samples, then ^(1/2.2)."""
    scene, _, _ = c1
    gpu_ctx.upload(scene)
    opts, w, h = scene.frame_opts()
    cam = scene.camera(w, h)
    gpu_ctx.resize(w, h)
    for _ in range(3):
        gpu_ctx.trace_frame(cam, opts)
    acc, n = gpu_ctx.read_accum()
    img = gpu_ctx.read_image()
    ref = np.power(acc[..., :3].astype(np.float64) / n, 1.0 / 2.2)
    assert np.allclose(img[..., :3], ref, rtol=2e-6, atol=1e-7)
    # the 8-bit image: the same values, clamped, * 255 + 0.5, truncated (float32 arithmetic)
    img8 = gpu_ctx.read_image_rgba8()
    q = np.clip(img[..., :3], np.float32(0), np.float32(1)).astype(np.float32)
    want = (q * np.float32(255.0) + np.float32(0.5)).astype(np.uint8)
    assert np.array_equal(img8[..., :3], want) and np.all(img8[..., 3] == 255)
    assert img8[..., :3].max() > 0
    # repeated reads reuse the image buffer and give the same bytes
    assert np.array_equal(gpu_ctx.read_image_rgba8(), img8)
    assert np.array_equal(gpu_ctx.read_image(), img)


def test_error_returns(ct, c1):
    ctx = ct.Context(0)
    try:
        scene, _, _ = c1
        opts, w, h = scene.frame_opts()
        cam = scene.camera(w, h)
        ctx.resize(w, h)
        with pytest.raises(ct.CTError) as e:
            ctx.trace_frame(cam, opts)
        assert e.value.code == ct.CT_ERR_NO_SCENE
        ctx.upload(scene)
        bad = opts.copy()
        bad.max_depth = 0
        with pytest.raises(ct.CTError) as e:
            ctx.trace_frame(cam, bad)
        assert e.value.code == ct.CT_ERR_INVALID_ARG
        bad = opts.copy()
        bad.accel = 7
        with pytest.raises(ct.CTError) as e:
            ctx.trace_frame(cam, bad)
        assert e.value.code == ct.CT_ERR_INVALID_ARG
        with pytest.raises(ct.CTError):
            ctx.resize(0, 10)
        ctx.resize(w, h)
        ctx.trace_frame(cam, opts)  # still usable after the errors
    finally:
        ctx.close()


def _uv_sphere_obj(n_lat: int = 12, n_lon: int = 24) -> str:
    """A UV sphere of radius 1 as OBJ text (positions and normals, quads and
    triangles at the poles)."""
    lines, idx = [], {}
    for i in range(n_lat + 1):
        th = np.pi * i / n_lat
        for j in range(n_lon):
            ph = 2 * np.pi * j / n_lon
            p = (np.sin(th) * np.cos(ph), np.cos(th), np.sin(th) * np.sin(ph))
            idx[(i, j)] = len(idx) + 1
            lines.append("v %.6f %.6f %.6f" % p)
            lines.append("vn %.6f %.6f %.6f" % p)
    for i in range(n_lat):
        for j in range(n_lon):
            a, b = idx[(i, j)], idx[(i, (j + 1) % n_lon)]
            c, d = idx[(i + 1, (j + 1) % n_lon)], idx[(i + 1, j)]
            if i == 0:  # a and b are the pole
                lines.append(f"f {a}//{a} {c}//{c} {d}//{d}")
            elif i == n_lat - 1:  # c and d are the pole
                lines.append(f"f {a}//{a} {b}//{b} {c}//{c}")
            else:
                lines.append(f"f {a}//{a} {b}//{b} {c}//{c} {d}//{d}")
    return "\n".join(lines) + "\n"


def test_obj_scene_matches_oracle(ct, orc, gpu_ctx, tmp_path):
    """A scene assembled like the reference's demo (main.cpp: an OBJ sphere as
    the light, OBJ spheres as objects, TriangleList::loadModel with scale and
    translation) renders on the GPU exactly as the FP32 oracle renders it."""
    from oracle import oracle as orm
    f = tmp_path / "sphere.obj"
    f.write_text(_uv_sphere_obj())
    s = ct.HostScene.empty()
    light = s.add_material((20.0, 20.0, 20.0), mtype=ct.MAT_LIGHT)
    red = s.add_material((0.8, 0.2, 0.2))
    white = s.add_material((0.7, 0.7, 0.7))

    def tr(x, y, z):
        m = np.eye(4, dtype=np.float32)
        m[:3, 3] = (x, y, z)
        return m
    s.load_model(f, 1.0, tr(0, 40, 0), light)
    s.load_model(f, 10.0, tr(-12, 10, 0), red)
    s.load_model(f, 10.0, tr(12, 10, 0), white)
    s.load_model(f, 200.0, tr(0, -200, 0), white)  # ground: a huge sphere
    s.build()
    arrays = s.arrays()
    assert arrays["light_indices"].size == 12 * 24 * 2 - 2 * 24  # the light sphere's triangles
    opts = ct.default_opts()
    opts.max_depth = 4
    opts.dist_epsilon = 1e-3
    w, h = 128, 96
    cam = ct.look_at((0, 15, 60), (0, 10, 0), 60.0)
    gpu_ctx.upload(s)
    gpu_ctx.resize(w, h)
    for _ in range(2):
        gpu_ctx.trace_frame(cam, opts)
    acc, _ = gpu_ctx.read_accum()
    ref = np.zeros((h, w, 4), np.float32)
    osc = orm.Scene(arrays)
    for fr in range(2):
        osc.render(cam, opts, w, h, fr, 1, threads=16, accum=ref)
    assert np.any(acc[..., :3] > 0)
    assert np.array_equal(_bits(acc), _bits(ref))
