"""GPU tests of textures, normal maps and the sky texture:
the textured atrium (config 4) against the oracle (FP32 per pixel, rel-MSE
against FP64), the invariants of the CPU tests on the GPU (a texture of the
albedo's colour, a flat normal map and a 1x1 sky texture change no bit),
material edits that switch textures on, and the upload's index checks."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from conftest import ArrayView as _View

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def _relmse(x, ref):
    return float(np.mean((x - ref) ** 2 / (ref ** 2 + 1e-2)))


def _mats(ct, arrays):
    """Copies of a scene's materials (the cached arrays stay untouched)."""
    return [ct.Material.from_buffer_copy(m) for m in arrays["materials"]]


def _upload(ctx, view: _View):
    ctx.upload_view(view.view)  # the _View keeps the arrays alive during the call


def _frame(ctx, cam, opts, w, h):
    ctx.resize(w, h)
    ctx.trace_frame(cam, opts)
    return ctx.read_accum()[0]


def test_textured_atrium_vs_oracle(ct, orc, gpu_ctx):
    """Config 4 (tiles + normal map on the floors, stone columns, sky
    texture) at 192x108, 2 spp: every value within 1e-4 of the FP32 oracle,
    rel-MSE vs FP64 within 1e-3."""
    scene = ct.HostScene.generate(4, 7)
    arrays = scene.arrays()
    osc = orc.Scene(arrays)
    gpu_ctx.upload(scene)
    opts, _, _ = scene.frame_opts()
    w, h = 192, 108
    opts.spp = 2
    cam = scene.camera(w, h)
    acc = _frame(gpu_ctx, cam, opts, w, h)
    ref, _ = osc.render(cam, opts, w, h, 0, 2, threads=16)
    diff = np.abs(acc - ref)
    rel = diff / np.maximum(np.abs(ref), 1e-30)
    bad = int(np.sum((rel > 1e-4) & (diff > 1e-7)))
    exact = int(np.sum(_bits(acc) == _bits(ref)))
    print(f"textured atrium: over 1e-4 rel {bad}, bit-identical {exact}/{acc.size}")
    assert bad == 0
    ref64, _ = osc.render(cam, opts, w, h, 0, 2, precision=64, threads=16)
    r = _relmse(acc[..., :3] / acc[..., 3:], ref64[..., :3] / ref64[..., 3:])
    print(f"textured atrium rel-MSE vs FP64: {r:.3e}")
    assert r <= 1e-3
    # the textures are visible: the untextured atrium differs
    gpu_ctx.upload(ct.HostScene.generate(2, 7))
    assert not np.array_equal(_frame(gpu_ctx, cam, opts, w, h), acc)


def test_textured_atrium_roulette_tail(ct, orc, gpu_ctx):
    """Config 4 with Russian roulette from bounce 1 at depth 10 (160x90,
    2 spp), so the roulette-tail launch shades textured, normal-mapped
    surfaces and the textured sky for every bounce from 1 on: every value
    within 1e-4 of the FP32 oracle and the ray counts equal the oracle's."""
    scene = ct.HostScene.generate(4, 7)
    osc = orc.Scene(scene.arrays())
    gpu_ctx.upload(scene)
    opts, _, _ = scene.frame_opts()
    w, h = 160, 90
    opts.spp = 2
    opts.max_depth = 10
    opts.rr_start = 1
    cam = scene.camera(w, h)
    gpu_ctx.resize(w, h)
    st = gpu_ctx.trace_frame(cam, opts)
    acc, _ = gpu_ctx.read_accum()
    assert st.tail_ms > 0.0
    ref, ost = osc.render(cam, opts, w, h, 0, 2, threads=16)
    assert (st.primary_rays, st.extension_rays, st.shadow_rays) == (
        ost["primary_rays"], ost["extension_rays"], ost["shadow_rays"])
    diff = np.abs(acc - ref)
    bad = int(np.sum((diff / np.maximum(np.abs(ref), 1e-30) > 1e-4) & (diff > 1e-7)))
    exact = int(np.sum(_bits(acc) == _bits(ref)))
    print(f"textured atrium, roulette tail: over 1e-4 rel {bad}, bit-identical {exact}/{acc.size}")
    assert bad == 0


def test_gpu_texture_invariants(ct, orc, gpu_ctx, c1, c2):
    scene, arrays, _ = c1
    opts, _, _ = scene.frame_opts()
    w = h = 64
    cam = scene.camera(w, h)
    _upload(gpu_ctx, _View(ct, arrays))
    base = _frame(gpu_ctx, cam, opts, w, h)
    mats = _mats(ct, arrays)
    texs = []
    for m in mats:
        if m.type != ct.MAT_LIGHT:
            t = np.zeros((3, 5, 4), np.float32)
            t[..., :3] = np.array(m.albedo[:], np.float32)
            t[..., 3] = 1.0
            texs.append(t)
            m.diffuse_tex = len(texs) - 1
    flat = np.zeros((2, 2, 4), np.float32)
    flat[..., 2] = 1.0
    texs.append(flat)
    for m in mats:
        if m.type != ct.MAT_LIGHT:
            m.normal_tex = len(texs) - 1
    vw = _View(ct, arrays, texs, mats)
    _upload(gpu_ctx, vw)
    assert np.array_equal(_bits(_frame(gpu_ctx, cam, opts, w, h)), _bits(base))
    # the same textured scene against the oracle with a tilted normal map
    tilted = np.zeros((2, 2, 4), np.float32)
    tilted[..., 0] = 0.6
    tilted[..., 2] = 0.8
    texs2 = texs[:-1] + [tilted]
    _upload(gpu_ctx, _View(ct, arrays, texs2, mats))
    acc = _frame(gpu_ctx, cam, opts, w, h)
    a = dict(arrays)
    a["materials"] = vw.mats
    a["textures"] = texs2
    a["sky_texture"] = -1
    ref, _ = orc.Scene(a).render(cam, opts, w, h, 0, 1, threads=16)
    diff = np.abs(acc - ref)
    assert int(np.sum((diff / np.maximum(np.abs(ref), 1e-30) > 1e-4) & (diff > 1e-7))) == 0
    assert not np.array_equal(acc, base)
    # sky: a 1x1 texture of the constant colour
    scene2, arrays2, _ = c2
    opts2, _, _ = scene2.frame_opts()
    cam2 = scene2.camera(96, 54)
    _upload(gpu_ctx, _View(ct, arrays2))
    sky0 = _frame(gpu_ctx, cam2, opts2, 96, 54)
    one = np.zeros((1, 1, 4), np.float32)
    one[0, 0, :3] = arrays2["sky"]
    _upload(gpu_ctx, _View(ct, arrays2, [one], sky=0))
    assert np.array_equal(_bits(_frame(gpu_ctx, cam2, opts2, 96, 54)), _bits(sky0))


def test_material_update_switches_texture_on(ct, orc, gpu_ctx):
    """ct_material_update with a diffuse texture index renders like an
    upload of the edited scene (oracle of the edited materials)."""
    scene = ct.HostScene.generate(4, 7)
    arrays = scene.arrays()
    gpu_ctx.upload(scene)
    opts, _, _ = scene.frame_opts()
    w, h = 96, 54
    cam = scene.camera(w, h)
    mats = _mats(ct, arrays)
    k = next(i for i, m in enumerate(mats) if m.type == ct.MAT_LAMBERTIAN and m.diffuse_tex < 0 and i > 0)
    m = ct.Material()
    C.memmove(C.byref(m), C.byref(mats[k]), C.sizeof(ct.Material))
    m.diffuse_tex = 2  # the stone texture
    gpu_ctx.material_update(k, m)
    acc = _frame(gpu_ctx, cam, opts, w, h)
    mats[k] = m
    a = dict(arrays)
    a["materials"] = (ct.Material * len(mats))(*mats)
    ref, _ = orc.Scene(a).render(cam, opts, w, h, 0, 1, threads=16)
    diff = np.abs(acc - ref)
    rel = diff / np.maximum(np.abs(ref), 1e-30)
    assert int(np.sum((rel > 1e-4) & (diff > 1e-7))) == 0
    bad = ct.Material()
    C.memmove(C.byref(bad), C.byref(m), C.sizeof(ct.Material))
    bad.normal_tex = 99
    with pytest.raises(ct.CTError):
        gpu_ctx.material_update(k, bad)


def test_texture_index_checks(ct, gpu_ctx, c1):
    scene, arrays, _ = c1
    mats = _mats(ct, arrays)
    mats[1].diffuse_tex = 0  # no textures in the view
    with pytest.raises(ct.CTError):
        _upload(gpu_ctx, _View(ct, arrays, (), mats))
    t = np.ones((2, 2, 4), np.float32)
    with pytest.raises(ct.CTError):
        _upload(gpu_ctx, _View(ct, arrays, [t], sky=1))
    # a zeroed sky_texture without textures keeps the constant sky
    _upload(gpu_ctx, _View(ct, arrays, (), None, sky=0))


def test_textured_atrium_full_frame_schedules(ct, orc, gpu_ctx):
    """Config 4 at its full size (1920x1080, depth 8): eight middle rows
    bit-identical to the FP32 oracle, and the whole frame bit-identical, with
    equal ray counts, under chunks one at a time, the counting frame and a
    context of eight chunks; texture fetches depend on nothing but the hit."""
    scene = ct.HostScene.generate(4, 7)
    osc = orc.Scene(scene.arrays())
    gpu_ctx.upload(scene)
    opts, w, h = scene.frame_opts()
    assert (w, h) == (1920, 1080)
    cam = scene.camera(w, h)
    gpu_ctx.resize(w, h)

    def counts(st):
        return (st.primary_rays, st.extension_rays, st.shadow_rays)

    gpu_ctx.reset()
    st0 = gpu_ctx.trace_frame(cam, opts)
    ref = gpu_ctx.read_accum()[0].copy()
    rows, r0 = 8, (h - 8) // 2
    band = np.zeros((h, w, 4), np.float32)
    osc.render(cam, opts, w, h, 0, 1, rows=(r0, r0 + rows), threads=16, accum=band)
    assert np.array_equal(_bits(ref[r0:r0 + rows]), _bits(band[r0:r0 + rows]))
    print(f"textured atrium full frame: rays {counts(st0)}; rows {r0}..{r0 + rows - 1} bit-identical", flush=True)

    def same(label, ctx, st):
        acc, s = ctx.read_accum()
        diff = int(np.sum(_bits(acc) != _bits(ref)))
        print(f"textured atrium full frame, {label}: words differing {diff}", flush=True)
        assert s == 1 and diff == 0 and counts(st) == counts(st0)

    gpu_ctx.reset()
    same("chunks one at a time", gpu_ctx, gpu_ctx.trace_frame_serial(cam, opts))
    gpu_ctx.reset()
    same("counting frame", gpu_ctx, gpu_ctx.trace_frame_counted(cam, opts)[0])
    eight = ct.Context(0, chunk_paths=1 << 18)
    try:
        eight.upload(scene)
        eight.resize(w, h)
        st = eight.trace_frame(cam, opts)
        assert st.chunks == 8
        same("8 chunks", eight, st)
    finally:
        eight.close()
