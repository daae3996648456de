"""GPU parity of the quad-tree traversal on rays aimed at the slab test's
special cases, against the oracle (the reference's arithmetic: (y < x) ? y : x
min/max, the 5-comparator network, traverseMBVH's push order).

Ray origins sit exactly on box planes of the tree, so entry or exit distances
are 0, and 0 * inf gives NaN where a direction component is 0. Direction
components are +0 and -0. Origins have +0 and -0 coordinates. Some directions
are not normalised. The builder pads every triangle box (min - 0.001,
max + 0.01), so built trees never hold a plane of -0.0; one test therefore
edits a built tree so that planes near 0 become +0 or -0, uploads it as a
scene view and gives the oracle the same tree. That scene also holds
duplicated triangles, so exact t ties make the id depend on the visit order.
The bar is the parity bar: t and ids bit-identical, shadow flags identical.
"""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N_EDGE = 1 << 17


def _plane_values(arrays):
    """Per axis, the bmin/bmax values of the quad tree's occupied slots."""
    raw = arrays["mbvh"]
    f = raw.view(np.float32).reshape(-1, 32)
    counts = raw.view(np.int32).reshape(-1, 32)[:, 28:32]
    occupied = counts != 0
    out = []
    for a in range(3):
        rows = f[:, 8 * a:8 * a + 8].reshape(-1, 2, 4)  # bmin row, bmax row
        out.append(rows[np.repeat(occupied[:, None, :], 2, axis=1)])
    return out


def _edge_rays(arrays, n, seed):
    rng = np.random.default_rng(seed)
    planes = _plane_values(arrays)
    V = arrays["vertices"]
    lo, hi = V.min(0), V.max(0)
    o = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    for a in range(3):
        on = rng.random(n) < 0.6
        o[on, a] = rng.choice(planes[a], int(on.sum()))
    z = rng.random((n, 3)) < 0.05
    o[z] = np.where(rng.random(int(z.sum())) < 0.5, np.float32(0.0), np.float32(-0.0))
    d = rng.normal(size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    k = rng.random((n, 3)) < 0.15
    d[k] = np.where(rng.random(int(k.sum())) < 0.5, np.float32(0.0), np.float32(-0.0))
    scale = rng.choice(np.array([1.0, 1.0, 1.0, 3.0, 1e-3], np.float32), n)
    d *= scale[:, None]
    dead = np.all(d == 0, axis=1)
    d[dead] = np.array([0.0, -1.0, 0.0], np.float32)
    return o, d


def _signed_zero_scene(ct):
    """Floor tiles in y = 0 whose vertices mix y = +0 and -0, two copies of a
    wall in x = 0 (exact ties), and scattered boxes' faces on integer planes."""
    rng = np.random.default_rng(17)
    tris = []
    for i in range(-6, 6):
        for j in range(-6, 6):
            a = np.array([[i, 0, j], [i + 1, 0, j], [i, 0, j + 1]], np.float32)
            b = np.array([[i + 1, 0, j], [i + 1, 0, j + 1], [i, 0, j + 1]], np.float32)
            for p in (a, b):
                p[:, 1] = np.where(rng.random(3) < 0.5, np.float32(0.0), np.float32(-0.0))
                tris.append(p)
    wall = [np.array([[0, 0, -3], [0, 3, -3], [0, 0, 3]], np.float32),
            np.array([[0, 3, -3], [0, 3, 3], [0, 0, 3]], np.float32)]
    for w in wall:
        w[:, 0] = np.where(rng.random(3) < 0.5, np.float32(0.0), np.float32(-0.0))
    tris += wall + [w.copy() for w in wall]
    for _ in range(40):
        c = rng.integers(-5, 5, 3).astype(np.float32)
        s = np.float32(rng.integers(1, 3))
        for ax in range(3):
            for side in (0.0, s):
                u, v = [q for q in range(3) if q != ax]
                q0 = c.copy()
                q0[ax] += side
                q1, q2, q3 = q0.copy(), q0.copy(), q0.copy()
                q1[u] += s
                q2[v] += s
                q3[u] += s
                q3[v] += s
                tris += [np.stack([q0, q1, q2]), np.stack([q1, q3, q2])]
    sc = ct.HostScene.empty()
    m = sc.add_material((0.6, 0.5, 0.4))
    for p in tris:
        nrm = np.cross(p[1] - p[0], p[2] - p[0])
        nrm = nrm / np.linalg.norm(nrm)
        sc.add_triangle(p, np.tile(nrm, 3), m)
    sc.build()
    return sc


def _signed_zero_trees(arrays, seed):
    """Copies of the scene's arrays whose tree planes within the padding of 0
    (|v| < 0.02) are replaced by +0 or -0 at random."""
    rng = np.random.default_rng(seed)
    a = dict(arrays)
    m = arrays["mbvh"].copy().view(np.float32).reshape(-1, 32)
    planes = m[:, :24]
    near = np.abs(planes) < 0.02
    planes[near] = np.where(rng.random(int(near.sum())) < 0.5, np.float32(0.0), np.float32(-0.0))
    b = arrays["bvh"].copy().view(np.float32).reshape(-1, 8)
    for col in (0, 1, 2, 4, 5, 6):
        v = b[:, col]
        near = np.abs(v) < 0.02
        v[near] = np.where(rng.random(int(near.sum())) < 0.5, np.float32(0.0), np.float32(-0.0))
        b[:, col] = v
    a["mbvh"] = m.view(np.uint8).reshape(-1)
    a["bvh"] = b.view(np.uint8).reshape(-1)
    return a


def _check_closest(ct, orc, gpu_ctx, scene, arrays, o, d, accel, label, osc=None):
    osc = osc or orc.Scene(arrays)
    opts, _, _ = scene.frame_opts()
    opts.accel = accel
    opts.traversal = ct.TRAV_EXACT
    t_gpu, p_gpu = gpu_ctx.trace_rays(o, d, opts)
    t_ref, p_ref = osc.trace(o, d, accel, 32, opts.triangle_epsilon, threads=16)
    bad = np.flatnonzero((p_gpu != p_ref) | (t_gpu.view(np.uint32) != t_ref.astype(np.float32).view(np.uint32)))
    print(f"{label} accel {accel}: rays {o.shape[0]}, hits {int(np.sum(p_ref >= 0))}, mismatches {bad.size}")
    assert bad.size == 0, f"first {bad[:4]}: o {o[bad[:4]]} d {d[bad[:4]]} gpu {p_gpu[bad[:4]]} ref {p_ref[bad[:4]]}"
    return osc, opts


@pytest.mark.parametrize("config", [1, 2])
def test_edge_rays_closest_hit(ct, orc, gpu_ctx, config):
    from conftest import get_scene
    scene, arrays, _ = get_scene(ct, orc, config)
    gpu_ctx.upload(scene)
    o, d = _edge_rays(arrays, N_EDGE, 40 + config)
    _check_closest(ct, orc, gpu_ctx, scene, arrays, o, d, ct.ACCEL_BVH4, f"config {config}")


@pytest.mark.parametrize("accel", [0, 1])
def test_signed_zero_planes_and_ties(ct, orc, gpu_ctx, accel):
    from conftest import ArrayView
    scene = _signed_zero_scene(ct)
    arrays = _signed_zero_trees(scene.arrays(), 80 + accel)
    f = arrays["mbvh"].view(np.uint32).reshape(-1, 32)[:, :24]
    assert np.any(f == 0x80000000) and np.any(f == 0), "the tree should have +0 and -0 planes"
    view = ArrayView(ct, arrays)
    gpu_ctx.upload_view(view.view)
    osc = orc.Scene(arrays, build=False)
    osc.set_trees(arrays["bvh"], arrays["mbvh"], arrays["prim_indices"])
    o, d = _edge_rays(arrays, N_EDGE, 50 + accel)
    # a fifth of the coordinates +0 or -0: there a -0 plane gives plane - o = -0
    k = np.random.default_rng(90 + accel).random((N_EDGE, 3)) < 0.2
    o[k] = np.where(np.random.default_rng(91).random(int(k.sum())) < 0.5, np.float32(0.0), np.float32(-0.0))
    _, opts = _check_closest(ct, orc, gpu_ctx, scene, arrays, o, d, accel, "signed-zero scene", osc)
    rng = np.random.default_rng(60 + accel)
    tmax = rng.uniform(0.0, 12.0, o.shape[0]).astype(np.float32)
    occ_gpu = gpu_ctx.trace_shadow_rays(o, d, tmax, opts)
    occ_ref = osc.trace_shadow(o, d, tmax, accel, 32, opts.triangle_epsilon, threads=16)
    print(f"signed-zero scene accel {accel}: occluded {int(occ_ref.sum())}, differ {int(np.sum(occ_gpu != occ_ref))}")
    assert np.array_equal(occ_gpu, occ_ref)


def test_edge_rays_any_hit(ct, orc, gpu_ctx, c2):
    scene, arrays, osc = c2
    gpu_ctx.upload(scene)
    o, d = _edge_rays(arrays, N_EDGE, 70)
    tmax = np.random.default_rng(71).uniform(0.0, 80.0, o.shape[0]).astype(np.float32)
    opts, _, _ = scene.frame_opts()
    opts.accel = ct.ACCEL_BVH4
    opts.traversal = ct.TRAV_EXACT
    occ_gpu = gpu_ctx.trace_shadow_rays(o, d, tmax, opts)
    occ_ref = osc.trace_shadow(o, d, tmax, 1, 32, opts.triangle_epsilon, threads=16)
    assert np.array_equal(occ_gpu, occ_ref)


def _edit_empty_slots(arrays, what, seed):
    """Copies of the arrays whose quad-tree empty slots (count 0) get a child
    other than 0xFFFFFFFF ("child") or a NaN plane ("nan"); neither changes
    what a traversal finds, but both turn off the fast node test (the upload
    flags kEmptyEntry and kPlaneNaN), so rays take the exact test."""
    rng = np.random.default_rng(seed)
    a = dict(arrays)
    m = arrays["mbvh"].copy()
    u = m.view(np.uint32).reshape(-1, 32)
    f = m.view(np.float32).reshape(-1, 32)
    empty = np.argwhere(u[:, 28:32] == 0)
    assert empty.shape[0] > 0, "the tree should have empty slots"
    pick = empty[rng.choice(empty.shape[0], min(20, empty.shape[0]), replace=False)]
    for node, slot in pick:
        if what == "child":
            u[node, 24 + slot] = 7
        else:
            f[node, 8 * int(rng.integers(0, 3)) + slot] = np.nan
    a["mbvh"] = m
    return a


@pytest.mark.parametrize("what", ["child", "nan"])
def test_user_tree_flags_take_exact_test(ct, orc, gpu_ctx, c2, what):
    from conftest import ArrayView, ray_set
    scene, arrays, _ = c2
    edited = _edit_empty_slots(arrays, what, 5)
    view = ArrayView(ct, edited)
    gpu_ctx.upload_view(view.view)
    osc = orc.Scene(edited, build=False)
    osc.set_trees(edited["bvh"], edited["mbvh"], edited["prim_indices"])
    o, d = ray_set(scene, edited, 1 << 17, seed=95)
    _check_closest(ct, orc, gpu_ctx, scene, edited, o, d, ct.ACCEL_BVH4, f"empty slots ({what})", osc)
    opts, _, _ = scene.frame_opts()
    opts.accel = ct.ACCEL_BVH4
    tmax = np.random.default_rng(96).uniform(0.0, 80.0, o.shape[0]).astype(np.float32)
    occ_gpu = gpu_ctx.trace_shadow_rays(o, d, tmax, opts)
    occ_ref = osc.trace_shadow(o, d, tmax, 1, 32, opts.triangle_epsilon, threads=16)
    assert np.array_equal(occ_gpu, occ_ref)
