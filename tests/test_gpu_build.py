"""GPU tests of the device tree build (ct_scene_upload_build).

The device build must reproduce the host build (BVHTree / MBVHTree, itself
checked against the oracle's 30-plane rescan and the golden tree hashes on
the CPU) byte for byte: binary nodes, quad nodes and primitive order, and
the leaf-ordered triangle records, so frames rendered from either upload are
bit-identical.  Scenes cover the configurations (large nodes split by the
chunked kernels), small and degenerate inputs (leaf root, identical
triangles that form one large SAH leaf, duplicates) and skewed sizes.
"""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def _assert_same_trees(ctx, arrays):
    t = ctx.read_trees()
    assert np.array_equal(t["prim_indices"], arrays["prim_indices"])
    assert t["bvh"].size == arrays["bvh"].size
    assert np.array_equal(t["bvh"], arrays["bvh"])
    assert t["mbvh"].size == arrays["mbvh"].size
    assert np.array_equal(t["mbvh"], arrays["mbvh"])


def _soup(ct, tris: np.ndarray):
    """Host scene of triangles (n, 3, 3) with flat normals, one Lambert material."""
    s = ct.HostScene.empty()
    m = s.add_material((0.6, 0.5, 0.4))
    for p in tris.astype(np.float32):
        n = np.cross(p[1] - p[0], p[2] - p[0])
        ln = np.linalg.norm(n)
        n = n / ln if ln > 0 else np.array([0.0, 1.0, 0.0])
        s.add_triangle(p, np.tile(n, 3), m)
    s.build()
    return s


def _random_tris(rng, n, spread=100.0, size=1.0):
    c = rng.uniform(-spread, spread, (n, 1, 3))
    return c + rng.normal(0.0, size, (n, 3, 3))


@pytest.mark.parametrize("config", [1, 2])
def test_device_build_equals_host_build(ct, gpu_ctx, config, c1, c2):
    scene, arrays, _ = c1 if config == 1 else c2
    st = gpu_ctx.upload_build(scene)
    _assert_same_trees(gpu_ctx, arrays)
    hs = scene.stats()
    assert st["n_bvh_nodes"] == hs["bvh_nodes"] and st["n_mbvh_nodes"] == hs["mbvh_nodes"]
    assert st["max_depth"] == hs["max_depth"]
    assert st["device_ms"] > 0.0
    if config == 2:
        # this scene has partition swap chains longer than the in-kernel chase,
        # so the pointer-jumping path is part of what the equality covers
        assert st["long_chains"] > 0 and st["jump_rounds"] > 0


def test_read_trees_of_host_upload(ct, gpu_ctx, c1):
    """ct_scene_read_trees returns what ct_scene_upload was given."""
    scene, arrays, _ = c1
    gpu_ctx.upload(scene)
    _assert_same_trees(gpu_ctx, arrays)


_CASES = ["one", "two", "three", "four", "five", "soup64", "soup2049", "soup5000", "identical3000", "duplicates",
          "clusters", "skewed", "far_first"]


@pytest.mark.parametrize("case", _CASES)
def test_device_build_small_and_degenerate(ct, gpu_ctx, case):
    rng = np.random.default_rng(1000 + _CASES.index(case))
    base = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    if case in ("one", "two", "three", "four", "five"):
        n = ["one", "two", "three", "four", "five"].index(case) + 1
        tris = _random_tris(rng, n, 10.0)
    elif case.startswith("soup"):
        tris = _random_tris(rng, int(case[4:]))
    elif case == "identical3000":
        tris = np.repeat(base[None], 3000, axis=0)  # one SAH leaf of 3000 (a large node that does not split)
    elif case == "duplicates":
        tris = np.concatenate([np.repeat(_random_tris(rng, 40), 60, axis=0), _random_tris(rng, 500)])
    elif case == "clusters":
        centres = rng.uniform(-1e3, 1e3, (12, 1, 1, 3))
        tris = (centres + rng.normal(0, 1.0, (12, 400, 3, 3)) * rng.uniform(0.01, 20.0, (12, 1, 1, 1))).reshape(-1, 3, 3)
    elif case == "far_first":
        # triangle 0 alone far away, 6000 near the origin: the root split puts
        # slot 0 on the right and every later slot on the left, so the swap
        # chain of the last slot runs through all 6000 slots
        tris = np.concatenate([base[None] + np.array([5000.0, 0.0, 0.0]), _random_tris(rng, 6000, 10.0, 0.3)])
    else:  # skewed: geometrically growing positions and sizes along x
        k = np.arange(120)
        x = 1.35 ** k
        tris = np.stack([np.stack([x, np.zeros_like(x), np.zeros_like(x)], 1),
                         np.stack([x * 1.1, np.ones_like(x), np.zeros_like(x)], 1),
                         np.stack([x, np.zeros_like(x), np.ones_like(x)], 1)], 1)
    s = _soup(ct, tris)
    arrays = s.arrays()
    st = gpu_ctx.upload_build(s)
    _assert_same_trees(gpu_ctx, arrays)
    assert st["max_depth"] == s.stats()["max_depth"]
    if case == "far_first":
        assert st["long_chains"] > 0


def test_frames_identical_after_device_build(ct, gpu_ctx, c2):
    """Triangle records and trees from the device build render the same bits
    as the host upload (exact traversal, quad and binary trees)."""
    scene, _, _ = c2
    opts, _, _ = scene.frame_opts()
    w, h = 240, 136
    cam = scene.camera(w, h)
    out = {}
    for how in ("host", "device"):
        if how == "host":
            gpu_ctx.upload(scene)
        else:
            gpu_ctx.upload_build(scene)
        for accel in (ct.ACCEL_BVH4, ct.ACCEL_BVH2):
            o = opts.copy()
            o.accel = accel
            gpu_ctx.resize(w, h)
            gpu_ctx.trace_frame(cam, o)
            out[(how, accel)] = gpu_ctx.read_accum()[0]
    for accel in (ct.ACCEL_BVH4, ct.ACCEL_BVH2):
        assert np.array_equal(_bits(out[("host", accel)]), _bits(out[("device", accel)]))


def test_wide_tree_after_device_build(ct, gpu_ctx, c1):
    """CT_BUILD_WIDE gives the wide tree of the host upload; without it the
    fast traversal is refused."""
    scene, _, _ = c1
    opts, w, h = scene.frame_opts()
    cam = scene.camera(w, h)
    o = opts.copy()
    o.traversal = ct.TRAV_FAST
    gpu_ctx.upload(scene)
    gpu_ctx.resize(w, h)
    gpu_ctx.trace_frame(cam, o)
    ref = gpu_ctx.read_accum()[0]
    gpu_ctx.upload_build(scene, wide=True)
    gpu_ctx.resize(w, h)
    gpu_ctx.trace_frame(cam, o)
    assert np.array_equal(_bits(gpu_ctx.read_accum()[0]), _bits(ref))
    gpu_ctx.upload_build(scene)
    with pytest.raises(ct.CTError):
        gpu_ctx.trace_frame(cam, o)
