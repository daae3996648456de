"""The traversal stack's bound on user-supplied trees (the reference's stack
of 32 has no guard, MBVHNode.cuh:128,167; here LDS + a spill region up to
128 entries, then CT_ERR_STACK_OVERFLOW).

A chain-shaped quad tree: node i holds node i + 1 in slot 0 and one-triangle
leaves in slots 1-3, every slot with the same box.  Equal entry distances
tagged with the slot (MBVHNode.cuh:91-94) put slot 0 on top, so each level
leaves three entries behind: a chain of L levels needs 3 L + 1 entries.
The stack model of tests/test_stack_model.py checks the address arithmetic
of both cases on the host."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from conftest import ArrayView

pytestmark = pytest.mark.gpu


def chain_scene(ct, levels: int, seed: int = 0) -> dict:
    rng = np.random.default_rng(seed)
    n_tris = 3 * (levels - 1) + 4
    verts = np.zeros((3 * n_tris, 3), np.float32)
    for t in range(n_tris):
        c = rng.uniform(-8, 8, 3)
        verts[3 * t:3 * t + 3] = c + rng.uniform(-0.6, 0.6, (3, 3))
    nrm = np.zeros_like(verts)
    cn = np.zeros((n_tris, 3), np.float32)
    for t in range(n_tris):
        p = verts[3 * t:3 * t + 3].astype(np.float64)
        n = np.cross(p[1] - p[0], p[2] - p[0])
        n /= np.linalg.norm(n)
        nrm[3 * t:3 * t + 3] = n
        cn[t] = n
    nodes = np.zeros((levels, 32), np.uint32)
    f = nodes.view(np.float32)
    leaf = 0
    for i in range(levels):
        f[i, 0:4], f[i, 4:8] = -10.0, 10.0    # x min / max of the four slots
        f[i, 8:12], f[i, 12:16] = -10.0, 10.0
        f[i, 16:20], f[i, 20:24] = -10.0, 10.0
        for s in range(4):
            if s == 0 and i < levels - 1:
                nodes[i, 24 + s] = i + 1
                nodes[i, 28 + s] = np.uint32(0xFFFFFFFF)  # count -1: interior
            else:
                nodes[i, 24 + s] = leaf
                nodes[i, 28 + s] = 1
                leaf += 1
    assert leaf == n_tris
    m = ct.Material()
    m.albedo[:] = [0.7, 0.7, 0.7]
    m.type = ct.MAT_LAMBERTIAN
    m.roughness = 1.0
    m.diffuse_tex = m.normal_tex = m.mask_tex = m.displace_tex = -1
    return {
        "vertices": verts, "normals": nrm, "texcoords": np.zeros((verts.shape[0], 2), np.float32),
        "indices": np.arange(3 * n_tris, dtype=np.uint32).reshape(-1, 3), "center_normals": cn,
        "material_idx": np.zeros(n_tris, np.uint32), "light_indices": np.zeros(0, np.uint32),
        "materials": (ct.Material * 1)(m), "prim_indices": np.arange(n_tris, dtype=np.uint32),
        "bvh": np.zeros(0, np.uint8), "mbvh": nodes.view(np.uint8).ravel(), "sky_enabled": 1,
        "sky": np.array([1.0, 1.0, 1.0], np.float32), "textures": [], "sky_texture": -1,
    }


def chain_rays(n: int, seed: int):
    rng = np.random.default_rng(seed)
    o = np.stack([np.full(n, -20.0), rng.uniform(-5, 5, n), rng.uniform(-5, 5, n)], 1).astype(np.float32)
    d = np.stack([np.ones(n), rng.uniform(-0.3, 0.3, n), rng.uniform(-0.3, 0.3, n)], 1)
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    return o, d


def test_deep_chain_within_bound_matches_oracle(ct, orc, gpu_ctx):
    """42 levels: 127 stack entries, 111 of them in the spill region (LDS
    depth 16); ids and t bit-exact against the oracle."""
    arrays = chain_scene(ct, 42, seed=1)
    av = ArrayView(ct, arrays)
    gpu_ctx.upload_view(av.view)
    opts = ct.default_opts()
    o, d = chain_rays(4096, 2)
    t_gpu, p_gpu = gpu_ctx.trace_rays(o, d, opts)
    osc = orc.Scene(arrays, build=False)
    osc.set_trees(arrays["bvh"], arrays["mbvh"], arrays["prim_indices"])
    t_ref, p_ref = osc.trace(o, d, 1, 32, opts.triangle_epsilon, threads=16)
    print(f"hits {int(np.sum(p_ref >= 0))} of {o.shape[0]}")
    assert np.array_equal(p_gpu, p_ref)
    assert np.array_equal(t_gpu.view(np.uint32), t_ref.astype(np.float32).view(np.uint32))
    occ = gpu_ctx.trace_shadow_rays(o, d, np.full(o.shape[0], 40.0, np.float32), opts)
    occ_ref = osc.trace_shadow(o, d, np.full(o.shape[0], 40.0, np.float32), 1, 32, opts.triangle_epsilon)
    assert np.array_equal(occ, occ_ref)


def test_deep_chain_frame_with_roulette_tail(ct, orc, gpu_ctx):
    """A frame over the 42-level chain (127 stack entries, 111 of them
    spilled) with roulette from bounce 0, so every bounce after the primary
    rays is traced by the roulette-tail launch's per-lane traversal through
    the spill region: bit-identical to the FP32 oracle, with the oracle's
    ray counts."""
    arrays = chain_scene(ct, 42, seed=1)
    av = ArrayView(ct, arrays)
    gpu_ctx.upload_view(av.view)
    osc = orc.Scene(arrays, build=False)
    osc.set_trees(arrays["bvh"], arrays["mbvh"], arrays["prim_indices"])
    opts = ct.default_opts()
    opts.max_depth = 6
    opts.rr_start = 0
    opts.spp = 2
    cam = ct.look_at((-30.0, 0.5, 0.3), (0.0, 0.0, 0.0), 20.0)
    w = h = 48
    gpu_ctx.resize(w, h)
    st = gpu_ctx.trace_frame(cam, opts)
    acc, _ = gpu_ctx.read_accum()
    ref, ost = osc.render(cam, opts, w, h, 0, 2, threads=16)
    print(f"chain frame: primary {st.primary_rays}, extension {st.extension_rays}, shadow {st.shadow_rays}")
    assert st.tail_ms > 0.0
    assert st.extension_rays > 0
    assert (st.primary_rays, st.extension_rays, st.shadow_rays) == (
        ost["primary_rays"], ost["extension_rays"], ost["shadow_rays"])
    assert np.array_equal(acc.view(np.uint32), ref.view(np.uint32))


def test_stack_overflow_reported_and_cleared(ct, gpu_ctx, c1):
    """50 levels need 151 entries: ct_trace_rays, ct_trace_rays_device and
    ct_trace_frame return CT_ERR_STACK_OVERFLOW (the kernels end normally,
    the overflowing rays stop), and the context works afterwards."""
    from conftest import DeviceArray
    arrays = chain_scene(ct, 50, seed=3)
    av = ArrayView(ct, arrays)
    gpu_ctx.upload_view(av.view)
    opts = ct.default_opts()
    o, d = chain_rays(1024, 4)
    with pytest.raises(ct.CTError) as e:
        gpu_ctx.trace_rays(o, d, opts)
    assert e.value.code == ct.CT_ERR_STACK_OVERFLOW
    # the device-pointer entry point reports it too, and clears the flag
    n = o.shape[0]
    h_o = np.zeros((n, 4), np.float32)
    h_d = np.zeros((n, 4), np.float32)
    h_o[:, :3], h_d[:, :3] = o, d
    d_o, d_d = DeviceArray(h_o), DeviceArray(h_d)
    hits = DeviceArray(np.zeros((n, 2), np.float32))
    with pytest.raises(ct.CTError) as e:
        gpu_ctx.trace_rays_device(d_o.addr, d_d.addr, n, hits.addr, opts)
    assert e.value.code == ct.CT_ERR_STACK_OVERFLOW
    cam = ct.look_at((-30.0, 0.5, 0.3), (0.0, 0.0, 0.0), 20.0)
    gpu_ctx.resize(32, 32)
    with pytest.raises(ct.CTError) as e:
        gpu_ctx.trace_frame(cam, opts)
    assert e.value.code == ct.CT_ERR_STACK_OVERFLOW
    # a well-formed scene traces without an error on the same context
    scene, _, _ = c1
    gpu_ctx.upload(scene)
    o1, w, h = scene.frame_opts()
    gpu_ctx.resize(64, 64)
    gpu_ctx.trace_frame(scene.camera(64, 64), o1)
    t, p = gpu_ctx.trace_rays(o[:8], d[:8], o1)
    assert t.shape == (8,)
    ms = gpu_ctx.trace_rays_device(d_o.addr, d_d.addr, 64, hits.addr, o1)
    assert ms > 0
