"""Pins of oracle functions that no reference vector covers, against what the
mathematics fixes (CPU only):

- Russian roulette (absent in the reference; Tools.h:91-94 holds an unused
  SurvivalProbability): unbiased in expectation on the box-and-sphere scene,
  and exact in a white furnace (every surviving path carries the furnace
  radiance, the survivors' fraction is the albedo);
- the dielectric branch's refraction (glm::refract,
  include/glm/detail/func_geometric.inl:112-122): Snell's law, the plane of
  incidence, unit length and total internal reflection;
- the microfacet lobes' pdf (mf_eval) against the density the sampler
  (sample_vndf + reflect, Microfacet.cuh:141-230,312-415) actually draws,
  by a chi-square test over the sphere of directions;
- next-event estimation + BSDF sampling + the balance-heuristic weights
  (Kernel.cu:114-136,161-195) on a glossy floor under a quad light: the MIS
  estimate equals the direct-lighting integral computed by quadrature of the
  same BRDF over the light.
A mistake of the kind these catch (a dropped 1/q, a sign of sqrt(k), a pdf
without G1, a MIS weight without the light pdf) fails one of them.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest


# ---------------------------------------------------------- refraction --
@pytest.mark.parametrize("eta", [1 / 1.5, 1.5, 1 / 1.33, 1.33, 1.0])
def test_refraction_snell_plane_and_tir(orc, eta):
    """glm::refract(I, N, eta) for unit I against the surface (N.I < 0):
    valid iff eta^2 sin^2(theta_i) <= 1; then |T| = 1, sin(theta_t) =
    eta sin(theta_i), T in the plane of I and N, on the far side of the
    surface, tangential part eta times I's.  FP32 follows FP64 to 1e-6."""
    rng = np.random.default_rng(int(eta * 1000))
    n_tir = n_ok = 0
    for _ in range(400):
        N = rng.normal(size=3)
        N /= np.linalg.norm(N)
        I = rng.normal(size=3)
        I /= np.linalg.norm(I)
        if np.dot(I, N) > 0:
            I = -I
        if abs(np.dot(I, N)) < 1e-3:
            continue
        If, Nf = I.astype(np.float32), N.astype(np.float32)
        I64, N64 = If.astype(np.float64), Nf.astype(np.float64)  # the values the oracle sees
        T, ok = orc.refract(If, Nf, eta, precision=64)
        cos_i = -float(np.dot(N64, I64))
        sin2_i = max(0.0, 1.0 - cos_i * cos_i)
        k = 1.0 - eta * eta * sin2_i
        assert ok == (k >= 0.0)
        if not ok:
            n_tir += 1
            continue
        n_ok += 1
        # glm::refract is exact only for unit inputs: the float inputs are
        # unit to ~1e-7
        assert abs(np.linalg.norm(T) - 1.0) < 1e-6
        assert np.dot(T, N64) < 0.0
        sin_t = np.linalg.norm(np.cross(N64, T))
        assert abs(sin_t - eta * math.sqrt(sin2_i)) < 1e-6
        assert abs(np.dot(T, np.cross(N64, I64))) < 1e-6  # plane of incidence
        tan_i = I64 - np.dot(I64, N64) * N64
        tan_t = T - np.dot(T, N64) * N64
        assert np.allclose(tan_t, eta * tan_i, atol=1e-6)
        # FP32 rounds k = 1 - eta^2 (1 - cos^2) with an absolute error of
        # ~1e-7, so sqrt(k) is well conditioned only away from k = 0
        T32, ok32 = orc.refract(If, Nf, eta, precision=32)
        assert ok32 or k < 1e-6
        if k > 1e-2:
            assert np.allclose(T32, T, atol=2e-6)
    assert n_ok > 100
    if eta > 1.0:
        assert n_tir > 10  # leaving the denser medium at grazing angles reflects totally
    else:
        assert n_tir == 0


def test_refraction_normal_incidence_and_critical_angle(orc):
    N = np.array([0.0, 0.0, 1.0], np.float32)
    T, ok = orc.refract(np.array([0.0, 0.0, -1.0], np.float32), N, 1 / 1.5)
    assert ok and np.allclose(T, [0, 0, -1], atol=1e-15)
    eta = 1.5
    crit = math.asin(1 / eta)
    for ang, want in ((crit - 1e-3, True), (crit + 1e-3, False)):
        I = np.array([math.sin(ang), 0.0, -math.cos(ang)], np.float32)
        assert orc.refract(I, N, eta)[1] == want


# ---------------------------------------------------- microfacet pdf ----
def _bsdf(orc, mtype, rough, wo, wi=None, r=(0.0, 0.0)):
    wo_ = (C.c_float * 3)(*wo)
    wi_in = (C.c_float * 3)(*wi) if wi is not None else None
    wi_out = (C.c_float * 3)()
    f = (C.c_float * 3)()
    p = C.c_float()
    orc.lib.or_bsdf_local(mtype, rough, wo_, wi_in, float(r[0]), float(r[1]), wi_out, f, C.byref(p))
    return np.array(wi_out[:]), f[0], p.value


@pytest.mark.parametrize("mtype", [4, 5])
@pytest.mark.parametrize("rough,theta_o", [(0.1, 0.4), (0.3, 1.0), (0.6, 1.35)])
def test_microfacet_pdf_matches_sampled_density(orc, mtype, rough, theta_o):
    """Directions wi = reflect(-wo, wm) with wm from the visible-normal
    sampler, binned over (cos theta_i, phi_i) on the upper hemisphere plus
    one bin for the lower one, against the bin masses of mf_eval's pdf
    (integrated by the midpoint rule): chi-square.  The pdf is the density of
    the directions the shading loop draws, so the MIS weights and the
    throughput's deferred 1/pdf use the right value."""
    wo = np.array([math.sin(theta_o), 0.0, math.cos(theta_o)], np.float32)
    rng = np.random.default_rng(17 + mtype)
    n = 30000
    rs = rng.random((n, 2)).astype(np.float32)
    wis = np.array([_bsdf(orc, mtype, rough, wo, r=r)[0] for r in rs])
    nc, nphi = 12, 12
    up = wis[:, 2] > 0
    ci = np.clip(wis[up, 2], 0, 1)
    ph = np.arctan2(wis[up, 1], wis[up, 0])
    hist, _, _ = np.histogram2d(ci, ph, bins=[nc, nphi], range=[[0, 1], [-np.pi, np.pi]])
    sub = 6
    exp = np.zeros((nc, nphi))
    for a in range(nc):
        for b in range(nphi):
            acc = 0.0
            for sa in range(sub):
                c_ = (a + (sa + 0.5) / sub) / nc
                s_ = math.sqrt(max(0.0, 1 - c_ * c_))
                for sb in range(sub):
                    p_ = -np.pi + (b + (sb + 0.5) / sub) * 2 * np.pi / nphi
                    wi = (s_ * math.cos(p_), s_ * math.sin(p_), c_)
                    acc += _bsdf(orc, mtype, rough, wo, wi=wi)[2]
            exp[a, b] = acc * (1.0 / nc) * (2 * np.pi / nphi) / (sub * sub)
    upper = float(exp.sum())
    assert upper <= 1.0 + 2e-3  # a density over the sphere
    e = np.append(exp.ravel() * n, (1.0 - upper) * n)
    o = np.append(hist.ravel(), (~up).sum())
    keep = e > 15
    chi2 = float(np.sum((o[keep] - e[keep]) ** 2 / e[keep]))
    dof = int(keep.sum()) - 1
    # mean dof, sd sqrt(2 dof): 5 sd plus slack for the midpoint rule
    assert chi2 < dof + 5 * math.sqrt(2 * dof) + 10, (chi2, dof)
    assert abs(o[keep].sum() - e[keep].sum()) < 4 * math.sqrt(n) + 0.02 * n


# ------------------------------------------------------ MIS vs integral --
def _glossy_floor_scene(ct, mtype, rough):
    s = ct.HostScene.empty()
    Le = 10.0
    lm = s.add_material([Le, Le, Le], mtype=ct.MAT_LIGHT)
    fm = s.add_material([1.0, 1.0, 1.0], mtype=mtype, roughness=rough, keep_roughness=True)
    h = 4.0
    lv = np.array([[-1.5, h, -1], [1.5, h, -1], [1.5, h, 1], [-1.5, h, 1]], np.float32)
    dn = np.array([[0, -1, 0]] * 3, np.float32)
    s.add_triangle(lv[[0, 1, 2]], dn, lm)
    s.add_triangle(lv[[0, 2, 3]], dn, lm)
    up = np.array([[0, 1, 0]] * 3, np.float32)
    fv = np.array([[-30, 0, -30], [30, 0, -30], [30, 0, 30], [-30, 0, 30]], np.float32)
    s.add_triangle(fv[[0, 1, 2]], up, fm)
    s.add_triangle(fv[[0, 2, 3]], up, fm)
    s.build()
    return s, Le, lv


def _direct_integral(orc, mtype, rough, p, wo, Le, lv, n=240):
    """L(p, wo) = int_light f(wo, wi) Le cos(theta_i) cos(theta_l) / d^2 dA,
    midpoint rule over the light rectangle (f from the oracle's BRDF in the
    local frame z = (0, 1, 0): the floor's normal)."""
    x0, x1 = float(lv[:, 0].min()), float(lv[:, 0].max())
    z0, z1 = float(lv[:, 2].min()), float(lv[:, 2].max())
    h = float(lv[0, 1])
    xs = x0 + (np.arange(n) + 0.5) / n * (x1 - x0)
    zs = z0 + (np.arange(n) + 0.5) / n * (z1 - z0)
    dA = (x1 - x0) * (z1 - z0) / (n * n)
    # local frame: tangent x, bitangent z, normal y (isotropic lobes)
    to_local = lambda v: (float(v[0]), float(v[2]), float(v[1]))  # noqa: E731
    wol = to_local(wo)
    total = 0.0
    for x in xs:
        for z in zs:
            q = np.array([x, h, z])
            d = q - p
            d2 = float(np.dot(d, d))
            wi = d / math.sqrt(d2)
            cos_i = wi[1]
            cos_l = wi[1]  # the light faces down: -wi . (0, -1, 0)
            f = _bsdf(orc, mtype, rough, wol, wi=to_local(wi))[1]
            total += f * Le * cos_i * cos_l / d2 * dA
    return total


@pytest.mark.parametrize("mtype,rough", [(5, 0.2), (5, 0.5), (4, 0.25)])
def test_mis_direct_light_equals_quadrature(ct, orc, mtype, rough):
    """A white glossy floor under a quad light (black sky, depth 2, uniform
    light selection pdf): one view direction (a camera with a 0.01 degree
    field of view, so every sample shades nearly one point).  The FP64
    oracle's estimate (NEE with the balance heuristic plus BSDF samples that
    hit the light) equals the quadrature of the same BRDF over the light."""
    import cudatracer_amd as ctm
    s, Le, lv = _glossy_floor_scene(ct, mtype, rough)
    osc = orc.Scene(s.arrays())
    # the view direction is the mirror image of the direction to the light's
    # centre, so the glossy peak falls on the light (the camera ray meets
    # nothing before the floor)
    p = np.array([3.0, 0.0, 0.3])
    to_light = np.array([0.0, 4.0, 0.0]) - p
    to_light /= np.linalg.norm(to_light)
    cam_pos = p + 7.0 * np.array([-to_light[0], to_light[1], -to_light[2]])
    cam = ctm.look_at(cam_pos, p, 0.01)
    o = ctm.default_opts()
    o.max_depth = 2
    o.bsdf_mode = ct.BSDF_TYPED
    o.light_pdf = ct.LIGHTPDF_UNIFORM
    o.dist_epsilon = 1e-3
    w = h = 4
    spp = 8192
    acc, st = osc.render(cam, o, w, h, 0, spp, precision=64, threads=8)
    assert st["shadow_rays"] > 0 and st["extension_rays"] > 0
    pix = acc[..., 0] / acc[..., 3]
    est, se = float(pix.mean()), float(pix.std(ddof=1) / math.sqrt(pix.size))
    wo = (cam_pos - p) / np.linalg.norm(cam_pos - p)
    ref = _direct_integral(orc, mtype, rough, p, wo, Le, lv)
    print(f"type {mtype} roughness {rough}: estimate {est:.5f} +- {se:.5f}, quadrature {ref:.5f}")
    assert ref > 0.05
    assert abs(est - ref) < 4 * se + 2e-3 * ref, (est, se, ref)


# ------------------------------------------------------ Russian roulette --
def _convex_sphere_scene(ct, rho, E):
    from test_oracle_cpu import _icosphere
    s = ct.HostScene.empty()
    m = s.add_material([rho, rho, rho])
    for tri in _icosphere(3, 10.0):
        n = np.cross(tri[1] - tri[0], tri[2] - tri[0])
        n = n / np.linalg.norm(n)
        if np.dot(n, tri.sum(0)) < 0:
            n = -n
        s.add_triangle(tri, np.array([n, n, n], np.float32), m)
    s.set_sky(True, [E, E, E])
    s.build()
    return s


def test_russian_roulette_white_furnace_exact(ct, orc):
    """Lambert sphere (albedo rho) under a constant sky E with roulette from
    the first bounce: after the sphere the survival probability is
    q = max(throughput / pdf) = rho, and a survivor's throughput is divided
    by q, so every sphere pixel is exactly E (survived) or 0 (terminated) and
    the survivors' fraction is rho: the expectation rho E of the estimator
    without roulette."""
    rho, E = 0.6, 2.0
    s = _convex_sphere_scene(ct, rho, E)
    osc = orc.Scene(s.arrays())
    import cudatracer_amd as ctm
    cam = ctm.look_at([0, 0, 40], [0, 0, 0], 40.0)
    o = ctm.default_opts()
    o.max_depth = 4
    o.rr_start = 0
    survivors = killed = 0
    for frame in range(4):
        acc, _ = osc.render(cam, o, 64, 64, frame, 1, precision=64)
        img = acc[..., 0]
        assert np.all(np.isclose(img, E, rtol=1e-9) | (img == 0.0))
        norr = o.copy()
        norr.rr_start = -1
        ref, _ = osc.render(cam, norr, 64, 64, frame, 1, precision=64)
        sphere = np.isclose(ref[..., 0], rho * E, rtol=1e-6)
        assert 0.1 < sphere.mean() < 0.9
        assert np.all(img[~sphere] == ref[..., 0][~sphere])  # background untouched
        survivors += int(np.sum(img[sphere] > 0))
        killed += int(np.sum(img[sphere] == 0))
    n = survivors + killed
    frac = survivors / n
    assert abs(frac - rho) < 4 * math.sqrt(rho * (1 - rho) / n), (frac, n)


def test_russian_roulette_unbiased_c1(ct, orc, c1):
    """Box-and-sphere scene, depth 4, NEE: the image mean with roulette from
    bounce 0 equals the mean without it within 4 standard errors (16
    independent batches of 16 samples per pixel each), and roulette does
    terminate paths (fewer extension rays)."""
    scene, arrays, osc = c1
    opts, _, _ = scene.frame_opts()
    w = h = 48
    cam = scene.camera(w, h)
    means = {}
    ext = {}
    for rr in (-1, 0):
        o = opts.copy()
        o.rr_start = rr
        batch = []
        n_ext = 0
        for b in range(16):
            acc, st = osc.render(cam, o, w, h, 16 * b, 16, precision=64, threads=8)
            batch.append(float(acc[..., :3].sum() / acc[..., 3].sum()))
            n_ext += st["extension_rays"]
        means[rr] = np.array(batch)
        ext[rr] = n_ext
    m0, m1 = means[-1].mean(), means[0].mean()
    se = math.sqrt(means[-1].var(ddof=1) / 16 + means[0].var(ddof=1) / 16)
    print(f"no RR {m0:.5f}, RR {m1:.5f}, se {se:.5f}; extension rays {ext[-1]} -> {ext[0]}")
    assert abs(m1 - m0) < 4 * se + 1e-9
    assert ext[0] < 0.9 * ext[-1]
