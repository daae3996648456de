"""CPU tests: pin the oracle (golden vectors, closed forms, brute force, a
second builder implementation) and check the host side of the product
(builders, generators, C-ABI exports, error paths).  No GPU needed."""
from __future__ import annotations

import ctypes as C
import hashlib
import json
import math
import re
from pathlib import Path

import numpy as np
import pytest

from conftest import ROOT, camera_rays, get_scene, ray_set, surface_rays

GOLDEN = Path(__file__).resolve().parent / "golden"


def _ulps(a: np.ndarray, ref: np.ndarray) -> np.ndarray:
    ref32 = ref.astype(np.float32)
    spacing = np.spacing(np.abs(ref32)).astype(np.float64)
    return np.abs(a.astype(np.float64) - ref) / spacing


# ------------------------------------------------------------------- RNG --
def test_rng_known_answers(orc):
    g = json.loads((GOLDEN / "rng.json").read_text())
    for s, h in g["wang_hash"]:
        assert orc.lib.or_wang_hash(s) == h
    for s, seq in g["xorshift"]:
        x = s
        for v in seq:
            x = orc.lib.or_xorshift(x)
            assert x == v


def test_rng_float_per_seed_uniform(orc):
    """The shading stage draws at most five numbers per fresh WangHash seed
    (Kernel.cu:144-165); each draw position is uniform on [0, 1].  (A long
    stream from one seed is not: the reference's generator shifts left three
    times, so low bits never mix upwards.)"""
    draws = np.zeros((5, 20000), np.float32)
    for k in range(20000):
        x = orc.lib.or_wang_hash((k * 16789 + 3 * 1791 + 2 * 720898027) & 0xFFFFFFFF)
        for j in range(5):
            x = orc.lib.or_xorshift(x)
            draws[j, k] = np.float32(x) * np.float32(2.3283064365387e-10)
    assert draws.min() >= 0.0 and draws.max() <= 1.0
    for j in range(5):
        hist, _ = np.histogram(draws[j], 10, (0, 1))
        chi2 = float(np.sum((hist - 2000.0) ** 2 / 2000.0))
        assert chi2 < 30.0, (j, chi2)  # 9 dof, p ~ 4e-4


# ------------------------------------------------------ transcendentals --
@pytest.mark.parametrize("name,fn,lo,hi,max_ulp", [
    ("or_ct_sinf", math.sin, -7.0, 7.0, 2.0),
    ("or_ct_cosf", math.cos, -7.0, 7.0, 2.0),
    ("or_ct_expf", math.exp, -80.0, 80.0, 2.0),
    ("or_ct_acosf", math.acos, -1.0, 1.0, 2.0),
])
def test_transcendentals_vs_libm(orc, name, fn, lo, hi, max_ulp):
    rng = np.random.default_rng(1)
    xs = rng.uniform(lo, hi, 20000).astype(np.float32)
    f = getattr(orc.lib, name)
    got = np.array([f(float(x)) for x in xs], np.float32)
    ref = np.array([fn(float(x)) for x in xs])
    ok = np.abs(ref) > 1e-6  # relative ulps away from zeros of sin/cos
    assert _ulps(got[ok], ref[ok]).max() <= max_ulp


def test_log_pow_vs_libm(orc):
    rng = np.random.default_rng(2)
    xs = np.exp(rng.uniform(-80, 80, 20000)).astype(np.float32)
    got = np.array([orc.lib.or_ct_logf(float(x)) for x in xs], np.float32)
    ref = np.array([math.log(float(x)) for x in xs])
    ok = np.abs(ref) > 1e-3
    assert _ulps(got[ok], ref[ok]).max() <= 2.0
    # pow(x, y) = exp(y log x) for the ranges BeckmannSample11 uses
    xs = rng.uniform(1e-6, 1.0, 5000).astype(np.float32)
    ys = rng.uniform(0.2, 1.2, 5000).astype(np.float32)
    got = np.array([orc.lib.or_ct_powf(float(x), float(y)) for x, y in zip(xs, ys)])
    ref = np.array([float(x) ** float(y) for x, y in zip(xs, ys)])
    assert np.max(np.abs(got - ref) / ref) < 2e-6
    assert orc.lib.or_ct_logf(0.0) == -math.inf
    assert math.isnan(orc.lib.or_ct_logf(-1.0))
    assert orc.lib.or_ct_expf(-200.0) == 0.0 and orc.lib.or_ct_expf(200.0) == math.inf


def test_erf_erfinv(orc):
    # A&S 7.1.26 has |error| <= 1.5e-7; Giles' erfinv inverts it to ~1e-6
    xs = np.linspace(-3, 3, 601)
    got = np.array([orc.lib.or_ct_erff(float(x)) for x in xs])
    ref = np.array([math.erf(x) for x in xs])
    assert np.max(np.abs(got - ref)) < 5e-7
    ys = np.linspace(-0.99, 0.99, 397)
    inv = np.array([orc.lib.or_ct_erfinvf(float(y)) for y in ys])
    back = np.array([math.erf(v) for v in inv])
    assert np.max(np.abs(back - ys)) < 2e-6


# ---------------------------------------------------------------- builder --
@pytest.mark.parametrize("config", [1, 2])
def test_builder_two_implementations_agree(ct, orc, config):
    """The product's binned builder and the oracle's 30-plane rescan build
    byte-identical binary trees, quad trees and primitive orders."""
    scene, arrays, osc = get_scene(ct, orc, config)
    bvh, mbvh, prims = osc.trees()
    assert np.array_equal(bvh, arrays["bvh"])
    assert np.array_equal(mbvh, arrays["mbvh"])
    assert np.array_equal(prims, arrays["prim_indices"])


@pytest.mark.parametrize("config", [1, 2])
def test_builder_golden_hashes(ct, orc, config):
    g = json.loads((GOLDEN / "trees.json").read_text())[f"c{config}_seed7"]
    scene, arrays, _ = get_scene(ct, orc, config)
    assert hashlib.sha256(arrays["bvh"].tobytes()).hexdigest() == g["bvh_sha256"]
    assert hashlib.sha256(arrays["mbvh"].tobytes()).hexdigest() == g["mbvh_sha256"]
    assert hashlib.sha256(arrays["prim_indices"].tobytes()).hexdigest() == g["prims_sha256"]


def _bvh_records(raw: np.ndarray):
    rec = np.frombuffer(raw.tobytes(), dtype=[("min", "<f4", 3), ("lf", "<i4"), ("max", "<f4", 3), ("cnt", "<i4")])
    return rec


@pytest.mark.parametrize("config", [1, 2])
def test_builder_invariants(ct, orc, config):
    scene, arrays, _ = get_scene(ct, orc, config)
    nodes = _bvh_records(arrays["bvh"])
    N = arrays["indices"].shape[0]
    V = arrays["vertices"]
    I = arrays["indices"]
    seen = np.zeros(N, np.int32)
    stack = [(0, 1)]
    max_depth = 0
    while stack:
        i, depth = stack.pop()
        max_depth = max(max_depth, depth)
        n = nodes[i]
        if n["cnt"] >= 0:
            lf, c = int(n["lf"]), int(n["cnt"])
            assert c > 0
            prims = arrays["prim_indices"][lf:lf + c]
            seen[prims] += 1
            # every triangle inside its leaf box
            pts = V[I[prims]].reshape(-1, 3)
            assert np.all(pts >= n["min"] - 1e-6) and np.all(pts <= n["max"] + 1e-6)
            continue
        for ch in (int(n["lf"]), int(n["lf"]) + 1):
            c = nodes[ch]
            assert np.all(c["min"] >= n["min"]) and np.all(c["max"] <= n["max"])  # child inside parent
            stack.append((ch, depth + 1))
    assert np.all(seen == 1)  # every primitive in exactly one leaf
    assert max_depth <= 63


def _decode_wide(w):
    """Decoded child boxes (n, 8, 3) lo / hi of wide nodes, as the kernel
    computes them: fmaf(q, 2^k, p) rounded once to float32."""
    p = w[:, 0:3].view(np.float32).astype(np.float64)
    k = np.stack([((w[:, 3] >> (8 * a)) & 0xFF).astype(np.int64) - 127 for a in range(3)], 1)
    step = np.ldexp(1.0, k)
    lo = np.zeros((len(w), 8, 3), np.float32)
    hi = np.zeros((len(w), 8, 3), np.float32)
    for a in range(3):
        for c in range(8):
            ql = (w[:, 4 + 4 * a + c // 4] >> (8 * (c % 4))) & 0xFF
            qh = (w[:, 6 + 4 * a + c // 4] >> (8 * (c % 4))) & 0xFF
            # q * 2^k + p is exact in float64 for these magnitudes; one rounding = fmaf
            lo[:, c, a] = (ql * step[:, a] + p[:, a]).astype(np.float32)
            hi[:, c, a] = (qh * step[:, a] + p[:, a]).astype(np.float32)
    return lo, hi


@pytest.mark.parametrize("config", [1, 2])
def test_wide_nodes_contain_their_geometry(ct, orc, config):
    """The fast mode's wide tree: every decoded child box (as the kernel
    decodes it) contains all vertices beneath it, every primitive is reached
    through exactly one leaf, leaves are the binary tree's leaves, arity <= 8,
    and the worst-case stack fits the traversal's 128 entries."""
    scene, arrays, _ = get_scene(ct, orc, config)
    w, bound = ct.build_wide_nodes(scene)
    assert 0 < bound <= 128
    n = len(w)
    valid = (w[:, 3] >> 24) & 0xFF
    arity = np.array([bin(int(v)).count("1") for v in valid])
    assert np.all(arity >= 1) and np.all(arity <= 8)
    assert np.all(valid == (1 << arity) - 1)  # children packed from slot 0
    bvh = _bvh_records(arrays["bvh"])
    leaf_count = {int(r["lf"]): int(r["cnt"]) for r in bvh if r["cnt"] >= 0}
    V = arrays["vertices"]
    I = arrays["indices"]
    prims = arrays["prim_indices"]
    lo, hi = _decode_wide(w)
    # vertex bounds beneath each node, children before parents (higher index)
    vmin = np.full((n, 3), np.inf)
    vmax = np.full((n, 3), -np.inf)
    seen = np.zeros(len(prims), np.int32)
    for i in range(n - 1, -1, -1):
        for c in range(int(arity[i])):
            e = int(w[i, 16 + c])
            if e & 0x80000000:
                first = e & 0x7FFFFFFF
                cnt = leaf_count[first]
                seen[first:first + cnt] += 1
                pts = V[I[prims[first:first + cnt]]].reshape(-1, 3)
                bmin, bmax = pts.min(0), pts.max(0)
            else:
                assert i < e < n
                bmin, bmax = vmin[e], vmax[e]
            assert np.all(lo[i, c] <= bmin) and np.all(hi[i, c] >= bmax), (i, c)
            vmin[i] = np.minimum(vmin[i], bmin)
            vmax[i] = np.maximum(vmax[i], bmax)
    assert np.all(seen == 1)


# -------------------------------------------------------------- traversal --
@pytest.mark.parametrize("config,accel", [(1, 0), (1, 1), (2, 1)])
def test_traversal_equals_brute_force(ct, orc, config, accel):
    scene, arrays, osc = get_scene(ct, orc, config)
    o, d = ray_set(scene, arrays, 1 << 12, seed=config)
    t, p = osc.trace(o, d, accel)
    tb, pb = osc.trace(o, d, 2)
    assert np.array_equal(t, tb)  # the closest distance is exact
    ties = np.flatnonzero(p != pb)
    assert ties.size <= 2  # ids differ only on exact t ties
    # any-hit agrees with brute force (order-free)
    tmax = np.where(pb >= 0, tb * 0.5, 50.0).astype(np.float32)
    occ = osc.trace_shadow(o, d, tmax, accel)
    occb = osc.trace_shadow(o, d, tmax, 2)
    assert np.array_equal(occ, occb)


def _single_triangle_scene(ct, orc, p, extra=()):
    s = ct.HostScene.empty()
    m = s.add_material([0.5, 0.5, 0.5])
    n = np.cross(np.subtract(p[1], p[0]), np.subtract(p[2], p[0]))
    n = n / np.linalg.norm(n)
    for tri in (p,) + tuple(extra):
        s.add_triangle(np.array(tri, np.float32), np.array([n, n, n], np.float32), m)
    s.build()
    return s, orc.Scene(s.arrays())


def test_moller_trumbore_cases(ct, orc):
    """Closed-form hits on the triangle (0,0,-5),(10,0,-5),(0,10,-5)."""
    s, osc = _single_triangle_scene(ct, orc, [[0, 0, -5], [10, 0, -5], [0, 10, -5]])
    o = np.array([[2, 2, 0], [9, 9, 0], [0, 0, 0], [2, 2, 0], [2, 2, -4.9995], [2, 2, 0]], np.float32)
    d = np.array([[0, 0, -1], [0, 0, -1], [0, 0, -1], [1, 0, 0], [0, 0, -1], [0, 0, 1]], np.float32)
    t, p = osc.trace(o, d, 2)
    assert t[0] == 5.0 and p[0] == 0               # interior hit at t = 5
    assert p[1] == -1                               # u + v > 1: miss
    assert t[2] == 5.0 and p[2] == 0                # vertex hit (u = v = 0 inclusive)
    assert p[3] == -1                               # parallel ray: |a| < eps
    assert p[4] == -1                               # t below the 1e-3 epsilon
    assert p[5] == -1                               # behind the origin
    for accel in (0, 1):
        t2, p2 = osc.trace(o, d, accel)
        assert np.array_equal(t2, t) and np.array_equal(p2, p)


def test_leaf_root_scene_is_visible(ct, orc):
    """A scene whose binary root is a leaf (< 4 triangles): the reference's
    quad collapse hides it; here both trees see it."""
    s, osc = _single_triangle_scene(ct, orc, [[0, 0, -5], [10, 0, -5], [0, 10, -5]])
    st = s.stats()
    assert st["bvh_nodes"] == 2 and st["mbvh_nodes"] == 1
    o = np.array([[2, 2, 0]], np.float32)
    d = np.array([[0, 0, -1]], np.float32)
    for accel in (0, 1):
        t, p = osc.trace(o, d, accel)
        assert p[0] == 0 and t[0] == 5.0


# ---------------------------------------------------- closed-form radiance --
def _icosphere(sub: int, radius: float):
    t = (1 + 5 ** 0.5) / 2
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6],
         [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10],
         [8, 6, 7], [9, 8, 1]]
    tris = np.array([[v[a], v[b], v[c]] for a, b, c in f])
    for _ in range(sub):
        a, b, c = tris[:, 0], tris[:, 1], tris[:, 2]
        ab, bc, ca = [x / np.linalg.norm(x, axis=1, keepdims=True) for x in (a + b, b + c, c + a)]
        tris = np.concatenate([np.stack(t_, 1) for t_ in ((a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca))])
    return (tris * radius).astype(np.float32)


def test_white_furnace_convex_lambert(ct, orc):
    """Lambert sphere (albedo rho) under a constant sky E, no lights: each
    path is camera -> sphere -> sky, so a pixel is exactly E (miss) or
    rho * E (cosine sampling weight f cos / pdf = rho)."""
    s = ct.HostScene.empty()
    rho = 0.6
    m = s.add_material([rho, rho, rho])
    for tri in _icosphere(3, 10.0):
        n = np.cross(tri[1] - tri[0], tri[2] - tri[0])
        n = n / np.linalg.norm(n)
        if np.dot(n, tri.sum(0)) < 0:
            n = -n
        s.add_triangle(tri, np.array([n, n, n], np.float32), m)
    E = 2.0
    s.set_sky(True, [E, E, E])
    s.build()
    osc = orc.Scene(s.arrays())
    import cudatracer_amd as ctm
    cam = ctm.look_at([0, 0, 40], [0, 0, 0], 40.0)
    o = ctm.default_opts()
    o.max_depth = 4
    for frame in range(4):  # one sample per pixel per frame (jitter mixes samples otherwise)
        acc, st = osc.render(cam, o, 64, 64, frame, 1, precision=64)
        img = acc[..., :3]
        is_bg = np.isclose(img, E, rtol=1e-9)
        is_sph = np.isclose(img, rho * E, rtol=1e-6)
        assert np.all(is_bg | is_sph)
        assert 0.1 < is_sph.mean() < 0.9


def _polygon_irradiance(p, n, verts):
    """Lambert's formula: E / L = 1/2 * sum_i Theta_i (n . Gamma_i)."""
    u = verts - p
    u = u / np.linalg.norm(u, axis=1, keepdims=True)
    s = 0.0
    for i in range(len(u)):
        a, b = u[i], u[(i + 1) % len(u)]
        th = math.acos(max(-1.0, min(1.0, float(np.dot(a, b)))))
        g = np.cross(a, b)
        g = g / np.linalg.norm(g)
        s += th * float(np.dot(n, g))
    return 0.5 * abs(s)


@pytest.mark.parametrize("shape,light_pdf,expect", [("tri", 0, 1.0), ("quad", 1, 1.0), ("quad", 0, None)])
def test_direct_light_matches_polygon_irradiance(ct, orc, shape, light_pdf, expect):
    """Lambert floor under an area light (NEE + MIS + light hits, depth 2,
    black sky): E[L] = rho/pi * Le * E/L from Lambert's polygon formula.
    One light triangle: the reference's light pdf is exact.  Two triangles:
    exact with the uniform-selection pdf; the reference's x count pdf is
    biased low (about 1/count^2 where the light pdf dominates)."""
    s = ct.HostScene.empty()
    Le = 10.0
    lm = s.add_material([Le, Le, Le], mtype=ct.MAT_LIGHT)
    fm = s.add_material([0.5, 0.5, 0.5])
    h = 4.0
    if shape == "tri":
        lv = np.array([[-1, h, -1], [1, h, -1], [0, h, 1]], np.float32)
        tris = [lv]
    else:
        lv = np.array([[-1, h, -1], [1, h, -1], [1, h, 1], [-1, h, 1]], np.float32)
        tris = [lv[[0, 1, 2]], lv[[0, 2, 3]]]
    dn = np.array([[0, -1, 0]] * 3, np.float32)
    for t_ in tris:
        s.add_triangle(t_, dn, lm)
    up = np.array([[0, 1, 0]] * 3, np.float32)
    fv = np.array([[-20, 0, -20], [20, 0, -20], [20, 0, 20], [-20, 0, 20]], np.float32)
    s.add_triangle(fv[[0, 1, 2]], up, fm)
    s.add_triangle(fv[[0, 2, 3]], up, fm)
    s.build()
    osc = orc.Scene(s.arrays())
    import cudatracer_amd as ctm
    cam = ctm.look_at([0.0, 3.0, 6.0], [0.0, 0.0, 0.0], 50.0)
    o = ctm.default_opts()
    o.max_depth = 2
    o.dist_epsilon = 1e-3
    o.light_pdf = light_pdf
    w = hgt = 16
    acc, _ = osc.render(cam, o, w, hgt, 0, 2048, precision=64, threads=8)
    img = acc[..., 0] / acc[..., 3]
    pos = np.array(cam.position, np.float64)
    wv = np.array(cam.forward, np.float64)
    u = np.cross(wv, [0, 1, 0]); u /= np.linalg.norm(u)
    v = np.cross(u, wv); v /= np.linalg.norm(v)
    ref = np.zeros((hgt, w))
    mask = np.zeros((hgt, w), bool)
    sub = (np.arange(6) + 0.5) / 6 - 0.5  # 6x6 sub-pixel grid: the render jitters over the pixel
    for y in range(hgt):
        for x in range(w):
            vals = []
            for jy in sub:
                for jx in sub:
                    # Kernel.cu:230-231: pixel x samples [x - 0.5, x + 0.5]
                    sx = 2 * (x + jx) / w - 1
                    sy = 1 - 2 * (y + jy) / hgt
                    dd = wv + u * cam.plane_distance * sx + v * cam.plane_distance * sy
                    dd /= np.linalg.norm(dd)
                    if dd[1] >= -1e-3:
                        break
                    p = pos + (-pos[1] / dd[1]) * dd
                    if abs(p[0]) > 18 or abs(p[2]) > 18:
                        break
                    vals.append(_polygon_irradiance(p, np.array([0, 1, 0.0]), lv.astype(np.float64)))
            if len(vals) == sub.size ** 2:
                ref[y, x] = 0.5 / math.pi * Le * float(np.mean(vals))
                mask[y, x] = True
    assert mask.sum() > 50
    ratio = img[mask].sum() / ref[mask].sum()  # pixel footprints average out over the image
    if expect is None:
        assert ratio < 0.5
    else:
        assert abs(ratio - expect) < 0.02, ratio


# ------------------------------------------------------------- microfacet --
@pytest.mark.parametrize("mtype", [4, 5])
@pytest.mark.parametrize("alpha", [0.1, 0.3, 0.7])
def test_microfacet_D_normalised(orc, mtype, alpha):
    """int D(wh) cos(theta_h) d(omega_h) = 1 over the hemisphere."""
    nt, nphi = 4000, 8
    th = (np.arange(nt) + 0.5) / nt * (np.pi / 2)
    total = 0.0
    for t_ in th:
        wh = (C.c_double * 3)(math.sin(t_), 0.0, math.cos(t_))
        dval = orc.lib.or_mf_D(mtype, alpha, wh)
        total += dval * math.cos(t_) * math.sin(t_)
    total *= (np.pi / 2) / nt * 2 * np.pi
    assert abs(total - 1.0) < 5e-3


@pytest.mark.parametrize("mtype", [4, 5])
def test_vndf_sampling_chi2(orc, mtype):
    """Histogram of sampled visible normals vs D_wo(wm) = G1(wo) max(0, wo.wm) D(wm) / cos(theta_o)."""
    alpha = 0.4
    wo = np.array([math.sin(0.8), 0.0, math.cos(0.8)], np.float32)
    rng = np.random.default_rng(4)
    n = 40000
    cos_t = np.empty(n)
    wm = (C.c_float * 3)()
    wof = (C.c_float * 3)(*wo)
    for i, (r1, r2) in enumerate(rng.random((n, 2)).astype(np.float32)):
        orc.lib.or_mf_sample_vndf32(mtype, alpha, wof, float(r1), float(r2), wm)
        cos_t[i] = wm[2]
    nb = 20
    edges = np.linspace(0, np.pi / 2, nb + 1)
    theta = np.arccos(np.clip(cos_t, -1, 1))
    hist, _ = np.histogram(theta, edges)
    # expected bin masses by quadrature of the visible-normal density
    lam = orc.lib.or_mf_lambda(mtype, alpha, (C.c_double * 3)(*wo.astype(float)))
    g1 = 1 / (1 + lam)
    exp = np.zeros(nb)
    for b in range(nb):
        ts = np.linspace(edges[b], edges[b + 1], 24)
        ph = np.linspace(-np.pi, np.pi, 96, endpoint=False)
        acc = 0.0
        for t_ in (ts[:-1] + ts[1:]) / 2:
            for p_ in ph:
                m = np.array([math.sin(t_) * math.cos(p_), math.sin(t_) * math.sin(p_), math.cos(t_)])
                dv = orc.lib.or_mf_D(mtype, alpha, (C.c_double * 3)(*m))
                acc += g1 * max(0.0, float(np.dot(wo, m))) * dv / wo[2] * math.sin(t_)
        exp[b] = acc * (ts[1] - ts[0]) * (2 * np.pi / 96)
    assert abs(exp.sum() - 1.0) < 0.02
    e = exp / exp.sum() * n
    keep = e > 20
    chi2 = float(np.sum((hist[keep] - e[keep]) ** 2 / e[keep]))
    assert chi2 < 3 * keep.sum() + 30, chi2


def test_microfacet_reciprocity_and_energy(orc):
    rng = np.random.default_rng(9)
    for mtype in (4, 5):
        for _ in range(200):
            a = rng.normal(size=3); a[2] = abs(a[2]) + 0.05; a /= np.linalg.norm(a)
            b = rng.normal(size=3); b[2] = abs(b[2]) + 0.05; b /= np.linalg.norm(b)
            f1 = (C.c_float * 3)(); f2 = (C.c_float * 3)(); p1 = C.c_float(); p2 = C.c_float()
            orc.lib.or_bsdf_local(mtype, 0.3, (C.c_float * 3)(*a), (C.c_float * 3)(*b), 0, 0, None, f1, C.byref(p1))
            orc.lib.or_bsdf_local(mtype, 0.3, (C.c_float * 3)(*b), (C.c_float * 3)(*a), 0, 0, None, f2, C.byref(p2))
            assert abs(f1[0] - f2[0]) <= 1e-5 * max(1.0, abs(f1[0]))
        # sampled weight f cos / pdf <= 1 for a white conductor (no energy gain)
        wo = (C.c_float * 3)(0.3, 0.1, math.sqrt(1 - 0.1))
        for r1, r2 in rng.random((500, 2)):
            wi = (C.c_float * 3)(); f = (C.c_float * 3)(); p = C.c_float()
            orc.lib.or_bsdf_local(mtype, 0.3, wo, None, float(r1), float(r2), wi, f, C.byref(p))
            if p.value > 0 and wi[2] > 0:
                assert f[0] * wi[2] / p.value <= 1.0 + 1e-4


def test_fresnel_dielectric_closed_form(orc):
    for eta in (1.5, 1 / 1.5, 1.33):
        r0 = ((eta - 1) / (eta + 1)) ** 2
        assert abs(orc.lib.or_fresnel_dielectric(1.0, eta) - r0) < 1e-12
        for ci in np.linspace(0.01, 1, 50):
            st2 = (1 - ci * ci) / eta ** 2
            if st2 >= 1:
                assert orc.lib.or_fresnel_dielectric(ci, eta) == 1.0  # total internal reflection
                continue
            ct_ = math.sqrt(1 - st2)
            rs = ((ci - eta * ct_) / (ci + eta * ct_)) ** 2
            rp = ((eta * ci - ct_) / (eta * ci + ct_)) ** 2
            assert abs(orc.lib.or_fresnel_dielectric(ci, eta) - 0.5 * (rs + rp)) < 1e-12


# ----------------------------------------------------- scene conventions --
def test_config_lights_first_and_flat_normals(ct, orc):
    """Q1 is inert (light triangles are primitives 0..L-1) and Q8 is inert
    (flat per-face normals) on the generated configurations."""
    for config in (1, 2):
        scene, arrays, _ = get_scene(ct, orc, config)
        L = arrays["light_indices"]
        assert np.array_equal(L, np.arange(L.size))
        nrm = arrays["normals"][arrays["indices"]]
        assert np.allclose(nrm[:, 0], nrm[:, 1]) and np.allclose(nrm[:, 0], nrm[:, 2])


def test_generated_sizes(ct, orc):
    s1, a1, _ = get_scene(ct, orc, 1)
    assert a1["indices"].shape[0] == 12 + 20480 + 2
    s2, a2, _ = get_scene(ct, orc, 2)
    assert 250_000 < a2["indices"].shape[0] < 270_000
    # the reference's absolute determinant epsilon culls nothing: min area >> 5e-4
    for a in (a1, a2):
        V = a["vertices"].astype(np.float64)[a["indices"]]
        area = 0.5 * np.linalg.norm(np.cross(V[:, 1] - V[:, 0], V[:, 2] - V[:, 0]), axis=1)
        assert area.min() > 0.05


def test_oracle_fp32_vs_fp64_c1(ct, orc, c1):
    scene, arrays, osc = c1
    opts, _, _ = scene.frame_opts()
    cam = scene.camera(64, 64)
    a32, _ = osc.render(cam, opts, 64, 64, 0, 4)
    a64, _ = osc.render(cam, opts, 64, 64, 0, 4, precision=64)
    x, r = a32[..., :3] / a32[..., 3:], a64[..., :3] / a64[..., 3:]
    assert float(np.mean((x - r) ** 2 / (r ** 2 + 1e-2))) < 1e-4


# ------------------------------------------------------------- the C ABI --
def test_cabi_exports_every_declared_symbol(ct):
    hdr = (ROOT / "include" / "cudatracer.h").read_text()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    names = set(re.findall(r"\b(ct_[a-z0-9_]+)\s*\(", hdr))
    assert len(names) > 30
    for n in sorted(names):
        assert hasattr(ct.lib, n), f"{n} declared in include/cudatracer.h but not exported"


def test_cabi_struct_layouts_match_the_binding(ct):
    """The ctypes mirrors have the sizes the library is built with (the C side
    checks them with static_asserts in host/capi_host.cpp) and the library
    reports the header's ABI version."""
    import ctypes as C
    assert C.sizeof(ct.Material) == 52 and C.sizeof(ct.Camera) == 40
    assert C.sizeof(ct.FrameOpts) == 52 and C.sizeof(ct.FrameStats) == 88
    ver = int(re.search(r"#define CT_ABI_VERSION (\d+)", (ROOT / "include" / "cudatracer.h").read_text()).group(1))
    assert ct.lib.ct_abi_version() == ver
    assert (ct.TRAV_EXACT, ct.TRAV_FAST, ct.TRAV_COMPACT) == (0, 1, 2)


def test_cabi_host_errors(ct):
    s = ct.HostScene.empty()
    with pytest.raises(ct.CTError):
        s.build()  # no triangles
    m = s.add_material([1, 1, 1])
    with pytest.raises(ct.CTError):
        s.add_triangle(np.array([[0, 0, 0], [1, 0, 0], [np.nan, 1, 0]], np.float32), np.zeros((3, 3), np.float32), m)
    with pytest.raises(ct.CTError):
        s.add_triangle(np.zeros((3, 3), np.float32), np.zeros((3, 3), np.float32), 99)  # unknown material
    with pytest.raises(ct.CTError):
        ct.HostScene.generate(42, 0)
    o = ct.default_opts()
    assert (o.max_depth, o.spp, o.rr_start, o.accel) == (5, 1, -1, ct.ACCEL_BVH4)
    assert abs(o.triangle_epsilon - 1e-3) < 1e-9 and abs(o.normal_epsilon - 1e-5) < 1e-12


def test_ctx_create_rejects_bad_shape_before_device(ct):
    """A launch shape the kernels cannot take is refused before any device
    call, so the error is the same with or without a GPU."""
    for kw in ({"block_size": 100}, {"block_size": 2048}, {"stack_lds": 129}):
        with pytest.raises(ct.CTError) as e:
            ct.Context(0, **kw)
        assert e.value.code == ct.CT_ERR_INVALID_ARG, kw


def test_add_material_forces_roughness_like_reference(ct):
    s = ct.HostScene.empty()
    i = s.add_material([0.5, 0.5, 0.5], roughness=0.05, keep_roughness=False)
    j = s.add_material([0.5, 0.5, 0.5], roughness=0.05, keep_roughness=True)
    s.add_triangle(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), np.array([[0, 0, 1]] * 3, np.float32), i)
    s.build()
    mats = s.arrays()["materials"]
    assert mats[0].roughness == 1.0  # default material 0
    assert mats[i].roughness == 1.0 and abs(mats[j].roughness - 0.05) < 1e-9


# ------------------------------------------------------------ OBJ import --
_CUBE_OBJ = """# unit cube, quads, v/vt/vn and negative indices
v -1 -1 -1
v 1 -1 -1
v 1 1 -1
v -1 1 -1
v -1 -1 1
v 1 -1 1
v 1 1 1
v -1 1 1
vt 0 0
vt 1 0
vt 1 1
vt 0 1
vn 0 0 -1
vn 0 0 1
vn 0 -1 0
vn 0 1 0
vn -1 0 0
vn 1 0 0
g cube
s off
f 1/1/1 4/4/1 3/3/1 2/2/1
f 5/1/2 6/2/2 7/3/2 8/4/2
f 1/1/3 2/2/3 6/3/3 5/4/3
f 4/1/4 8/2/4 7/3/4 3/4/4
f 1/1/5 5/2/5 8/3/5 4/4/5
f -7/1/6 -6/2/6 -2/3/6 -3/4/6
"""


def _host_list(ct, s):
    a = s.arrays()
    return a["vertices"], a["normals"], a["indices"], a["center_normals"], a["material_idx"], a["light_indices"]


def test_obj_import_transform_winding_and_lights(ct, tmp_path):
    """loadModel: quads split as fans with reversed winding (Assimp's
    Triangulate + FlipWindingOrder), corners shared, positions
    mat * (p * scale, 1), normals mat * (n, 0), centre normal = mean of the
    three, triangles of a Light material registered as lights."""
    f = tmp_path / "cube.obj"
    f.write_text(_CUBE_OBJ)
    s = ct.HostScene.empty()
    light = s.add_material((2.0, 2.0, 2.0), mtype=ct.MAT_LIGHT)
    T = np.eye(4, dtype=np.float32)
    T[:3, 3] = (10.0, -5.0, 2.0)
    s.load_model(f, scale=3.0, mat=T, material=light, normalize=False)
    s.build()
    V, N, I, CN, M, L = _host_list(ct, s)
    assert I.shape == (12, 3) and V.shape == (24, 3)  # 6 faces x 4 distinct corners
    assert np.array_equal(L, np.arange(12))
    assert np.all(M == light)
    # first face 1 4 3 2 -> fan (1,4,3),(1,3,2) -> reversed (3,4,1),(2,3,1)
    p = V[I[0]]
    assert np.allclose(p, [np.array(q) * 3 + (10, -5, 2) for q in ([1, 1, -1], [-1, 1, -1], [-1, -1, -1])])
    assert np.allclose(N[I[0]], [[0, 0, -1]] * 3)
    assert np.allclose(CN, N[I].mean(axis=1), atol=1e-7)
    lo, hi = V.min(0), V.max(0)
    assert np.allclose(lo, (7, -8, -1)) and np.allclose(hi, (13, -2, 5))


def test_obj_import_normalize_and_smooth_normals(ct, tmp_path):
    """normalize fits the model into [-1, 1] (centred, divided by half the
    largest extent); a file without normals gets smooth per-position normals
    (an octahedron's vertex normals point along its axes)."""
    v = [(0, 1, 4), (0, 1, -2), (3, 1, 1), (-3, 1, 1), (0, 4, 1), (0, -2, 1)]
    faces = [(1, 3, 5), (1, 5, 4), (1, 4, 6), (1, 6, 3), (2, 5, 3), (2, 4, 5), (2, 6, 4), (2, 3, 6)]
    f = tmp_path / "octa.obj"
    f.write_text("".join(f"v {x} {y} {z}\n" for x, y, z in v) + "".join(f"f {a} {b} {c}\n" for a, b, c in faces))
    s = ct.HostScene.empty()
    s.load_model(f, normalize=True)
    s.build()
    V, N, I, _, M, L = _host_list(ct, s)
    # without an MTL library the faces get the importer's default material (Lambert 0.6)
    assert I.shape == (8, 3) and V.shape == (6, 3) and L.size == 0 and np.all(M == 1)
    m1 = s.arrays()["materials"][1]
    assert m1.type == ct.MAT_LAMBERTIAN and np.allclose(m1.albedo[:], 0.6)
    assert np.allclose(V.min(0), (-1, -1, -1)) and np.allclose(V.max(0), (1, 1, 1))
    c = np.array([0.0, 1.0, 1.0])  # centre of the original box, in original units
    expect = (np.array(v, float) - c) / 3.0
    assert np.allclose(np.sort(V, axis=0), np.sort(expect, axis=0), atol=1e-6)
    d = V / np.linalg.norm(V, axis=1, keepdims=True)
    assert np.allclose(np.linalg.norm(N, axis=1), 1.0, atol=1e-6)
    assert np.all(np.sum(N * d, axis=1) > 0.99)  # outward along each vertex's axis


def test_obj_import_errors(ct, tmp_path):
    s = ct.HostScene.empty()
    with pytest.raises(ct.CTError) as e:
        s.load_model(tmp_path / "missing.obj")
    assert e.value.code == ct.CT_ERR_INVALID_ARG
    bad = tmp_path / "bad.obj"
    bad.write_text("v 0 0 0\nv 1 0 0\nf 1 2 9\n")
    with pytest.raises(ct.CTError) as e:
        s.load_model(bad)
    assert e.value.code == ct.CT_ERR_INVALID_SCENE and "line 3" in str(e.value)
    with pytest.raises(ct.CTError) as e:  # the list is unchanged: still no triangles
        s.build()
    assert e.value.code == ct.CT_ERR_INVALID_SCENE


@pytest.mark.skipif(not Path("/root/reference/Models/sphere.obj").exists(),
                    reason="the reference's model files are only present where the reference is mounted")
def test_obj_import_reference_sphere(ct, orc):
    """The reference's demo model Models/sphere.obj (960 faces, 482 vertices,
    vertex normals): normalised into the unit box it is a closed sphere of
    radius 1 with outward normals, and the oracle's traversal over the
    product's trees equals brute force."""
    s = ct.HostScene.empty()
    s.load_model("/root/reference/Models/sphere.obj", scale=1.0)
    s.build()
    a = s.arrays()
    V, N, I = a["vertices"], a["normals"], a["indices"]
    assert I.shape[0] == 960
    r = np.linalg.norm(V, axis=1)
    assert np.all(np.abs(r - 1.0) < 0.02)
    assert np.all(np.sum(N * V, axis=1) > 0.9)
    osc = orc.Scene(a)
    rng = np.random.default_rng(3)
    o = rng.normal(size=(4096, 3))
    o = (3.0 * o / np.linalg.norm(o, axis=1, keepdims=True)).astype(np.float32)  # on a sphere of radius 3
    tgt = rng.uniform(-0.5, 0.5, (4096, 3))
    d = tgt - o
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    t, p = osc.trace(o, d, 1, tri_eps=1e-6)
    tb, pb = osc.trace(o, d, 2, tri_eps=1e-6)
    assert np.array_equal(t, tb) and np.sum(p != pb) <= 2
    assert np.all(p >= 0)  # every ray aims inside the sphere
