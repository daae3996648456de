"""The render server at the served workload's sizes (resident atrium, depth 8,
256x256 / 512x512 / 1920x1080 requests with 1, 4 or 16 samples) against the
FP32 oracle and against the same request rendered alone; every request
option honoured; admission by accumulator memory.

A request's path samples are keyed by (pixel, frame_base + sample, bounce)
and summed in sample order, so its result does not depend on what else is
in flight: alone == in the mix, bit for bit."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def _alone(ctx, cam, w, h, opts, spp, frame_base=0):
    """The request alone: ct_trace_frame of spp samples from frame_base."""
    ctx.resize(w, h)
    if frame_base:
        ctx.write_accum(np.zeros((h, w, 4), np.float32), frame_base)
    o = opts.copy()
    o.spp = spp
    ctx.trace_frame(cam, o)
    acc, _ = ctx.read_accum()
    return acc


def _requests(ct, opts):
    rng = np.random.default_rng(11)
    sizes = [(256, 256, 16), (256, 256, 1), (512, 512, 4), (1920, 1080, 1), (256, 256, 4), (512, 512, 1),
             (1920, 1080, 4), (256, 256, 1)]
    reqs = []
    for i, (w, h, spp) in enumerate(sizes):
        pos = np.array([-85.0, 14.0, -80.0]) + rng.uniform(-10, 10, 3)
        pos[1] = min(max(pos[1], 3.0), 27.0)
        tgt = np.array([10.0, 9.0, 15.0]) + rng.uniform(-20, 20, 3)
        o = opts.copy()
        o.spp = spp
        reqs.append((ct.look_at(pos, tgt, 70.0), w, h, spp, o, 7 * i))
    return reqs


def test_served_sizes_vs_oracle_and_alone(ct, orc, gpu_ctx, c2):
    """Eight requests submitted together (so they share chunks, 2M paths
    each, with the server's default lanes): each 256x256 result equals the
    FP32 oracle's render bit for bit; each larger one on a band of 24 rows;
    every result equals the request rendered alone."""
    scene, arrays, osc = c2
    gpu_ctx.upload(scene)
    opts, _, _ = scene.frame_opts()
    assert opts.max_depth == 8
    reqs = _requests(ct, opts)
    srv = ct.Server(gpu_ctx, chunk_paths=1 << 21)
    try:
        ids = [srv.submit(cam, w, h, spp, o, frame_base=fb) for (cam, w, h, spp, o, fb) in reqs]
        outs = []
        for rid, (_, w, h, spp, _, _) in zip(ids, reqs):
            res, out = srv.wait(rid)
            assert res.status == 0 and (res.width, res.height, res.spp) == (w, h, spp)
            assert res.rays > 0 and res.done_ms >= res.admit_ms >= res.submit_ms
            outs.append(out)
    finally:
        srv.close()
    for (cam, w, h, spp, o, fb), out in zip(reqs, outs):
        print(f"request {w}x{h} spp {spp} frame_base {fb}", flush=True)
        alone = _alone(gpu_ctx, cam, w, h, opts, spp, fb)
        assert np.array_equal(_bits(out), _bits(alone)), (w, h, spp)
        ref = np.zeros((h, w, 4), np.float32)
        if w == 256:
            rows = (0, h)
        else:
            r0 = (h - 24) // 2
            rows = (r0, r0 + 24)
        osc.render(cam, o, w, h, fb, spp, rows=rows, threads=16, accum=ref)
        band = slice(rows[0], rows[1])
        assert np.array_equal(_bits(out[band]), _bits(ref[band])), (w, h, spp)
        assert np.all(out[band][..., 3] == spp)


def test_server_honours_every_option(ct, gpu_ctx, c2):
    """Requests that differ from the first only in light_pdf, in indirect,
    in an epsilon or in depth render with their own options (each equals its
    own plain render), in the same stream as the others."""
    scene, _, _ = c2
    gpu_ctx.upload(scene)
    opts, _, _ = scene.frame_opts()
    cam = ct.look_at((-80.0, 12.0, -75.0), (5.0, 8.0, 10.0), 70.0)
    w, h = 160, 96
    variants = []
    assert opts.light_pdf == ct.LIGHTPDF_REFERENCE and opts.indirect == 1 and opts.rr_start == -1
    for field, value in (("spp", 2), ("light_pdf", ct.LIGHTPDF_UNIFORM), ("indirect", 0),
                         ("normal_epsilon", 4 * opts.normal_epsilon), ("max_depth", 3), ("rr_start", 1)):
        o = opts.copy()
        o.spp = 2
        setattr(o, field, value)
        variants.append(o)
    srv = ct.Server(gpu_ctx, chunk_paths=1 << 15)
    try:
        ids = [srv.submit(cam, w, h, 2, o) for o in variants]
        outs = [srv.wait(rid)[1] for rid in ids]
    finally:
        srv.close()
    base = None
    for o, out in zip(variants, outs):
        alone = _alone(gpu_ctx, cam, w, h, o, 2)
        assert np.array_equal(_bits(out), _bits(alone))
        if base is None:
            base = out
        else:
            assert not np.array_equal(_bits(out), _bits(base))  # the option took effect


def test_served_roulette_tail_vs_oracle(ct, orc, gpu_ctx, c2):
    """Requests with Russian roulette from bounce 0 and from bounce 2 (so
    each server chunk ends in the launch that runs every path's remaining
    bounces), submitted together with a plain one: each result equals the
    FP32 oracle's render bit for bit."""
    scene, _, osc = c2
    gpu_ctx.upload(scene)
    opts, _, _ = scene.frame_opts()
    cam = ct.look_at((-82.0, 10.0, -70.0), (0.0, 10.0, 12.0), 70.0)
    w, h, spp = 128, 80, 2
    variants = []
    for rr in (0, 2, -1):
        o = opts.copy()
        o.spp = spp
        o.rr_start = rr
        variants.append(o)
    srv = ct.Server(gpu_ctx, chunk_paths=1 << 13)
    try:
        ids = [srv.submit(cam, w, h, spp, o, frame_base=3) for o in variants]
        outs = [srv.wait(rid)[1] for rid in ids]
    finally:
        srv.close()
    for o, out in zip(variants, outs):
        print(f"request {w}x{h} spp {spp} rr_start {o.rr_start}", flush=True)
        ref = np.zeros((h, w, 4), np.float32)
        osc.render(cam, o, w, h, 3, spp, threads=16, accum=ref)
        assert np.array_equal(_bits(out), _bits(ref)), o.rr_start
        assert np.all(out[..., 3] == spp)


def test_server_refuses_bad_requests(ct, gpu_ctx, c1):
    scene, _, _ = c1
    gpu_ctx.upload(scene)
    opts, w, h = scene.frame_opts()
    cam = scene.camera(w, h)
    srv = ct.Server(gpu_ctx, mem_budget=64 << 20, chunk_paths=1 << 12)
    try:
        bad = opts.copy()
        bad.max_depth = 0
        with pytest.raises(ct.CTError) as e:
            srv.submit(cam, 64, 64, 1, bad)
        assert e.value.code == ct.CT_ERR_INVALID_ARG and "max_depth" in str(e.value)
        flat = ct.Camera()
        with pytest.raises(ct.CTError) as e:  # zero forward vector, zero plane distance
            srv.submit(flat, 64, 64, 1, opts)
        assert e.value.code == ct.CT_ERR_INVALID_ARG and "camera" in str(e.value)
        # the arena is the budget less the shared queues (2 x 4096 slots of
        # 288 B): a 2048x2048 accumulator (64 MiB) can never fit
        with pytest.raises(ct.CTError) as e:
            srv.submit(cam, 2048, 2048, 1, opts)
        assert e.value.code == ct.CT_ERR_ADMISSION
        # 1900x2000 (58 MiB) fits the arena alone: admitted once the arena is free
        a = srv.submit(cam, 1900, 2000, 1, opts)
        b = srv.submit(cam, 64, 64, 1, opts)
        ra, _ = srv.wait(a)
        rb, out = srv.wait(b)
        assert ra.status == 0 and rb.status == 0
    finally:
        srv.close()
    # the context is the caller's again once the server is gone
    assert np.array_equal(_bits(out), _bits(_alone(gpu_ctx, cam, 64, 64, opts, 1)))
