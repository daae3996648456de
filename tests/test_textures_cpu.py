"""CPU tests of the texture path: the shared atan2 against
libm, the sky mapping and nearest-texel addressing of SceneData::getTextureColor
against their closed forms (SceneData.cuh:81-97, Kernel.cu:96-100), the
TGA/PPM/PFM loaders' texel order, OBJ + MTL materials, and render invariants in the oracle that any
slip in the shading integration breaks (a texture of the albedo's colour, a
flat normal map and a 1x1 sky texture must leave the image bit-identical;
a non-trivial map must change it)."""
from __future__ import annotations

import math
import struct

import numpy as np
import pytest

from conftest import get_scene


def _ulps(a, ref):
    ref32 = np.asarray(ref).astype(np.float32)
    spacing = np.spacing(np.abs(ref32)).astype(np.float64)
    return np.abs(np.asarray(a, np.float64) - ref) / spacing


def test_atan2_vs_libm(orc):
    rng = np.random.default_rng(3)
    ys = np.concatenate([rng.uniform(-10, 10, 20000), rng.normal(0, 1e-3, 2000), np.zeros(8)]).astype(np.float32)
    xs = np.concatenate([rng.uniform(-10, 10, 20000), rng.normal(0, 1e4, 2000),
                         np.array([1, -1, 0, -0.0, 2, -2, 1e-30, -1e-30])]).astype(np.float32)
    got = np.array([orc.lib.or_atan2f(float(y), float(x)) for y, x in zip(ys, xs)], np.float32)
    ref = np.array([math.atan2(float(y), float(x)) for y, x in zip(ys, xs)])
    ok = np.abs(ref) > 1e-30
    assert _ulps(got[ok], ref[ok]).max() <= 2.0
    # signed zeros and the axes
    assert orc.lib.or_atan2f(0.0, -0.0) == np.float32(math.pi)
    assert math.copysign(1.0, orc.lib.or_atan2f(-0.0, 1.0)) == -1.0
    assert orc.lib.or_atan2f(1.0, 0.0) == np.float32(math.pi / 2)
    assert orc.lib.or_atan2f(-1.0, -1.0) == np.float32(-3 * math.pi / 4)


def test_sky_uv_closed_form(orc):
    rng = np.random.default_rng(4)
    d = rng.normal(size=(3000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d = d.astype(np.float32)
    import ctypes as C
    u, v = C.c_float(), C.c_float()
    for dx, dy, dz in d:
        orc.lib.or_sky_uv(float(dx), float(dy), float(dz), C.byref(u), C.byref(v))
        eu = 1.0 + math.atan2(float(dx), -float(dz)) / math.pi * 0.5
        ev = 1.0 - math.acos(float(dy)) / math.pi
        assert abs(u.value - eu) < 1e-6 and abs(v.value - ev) < 1e-6


def _wrap32(x: np.float32) -> np.float32:
    """fmod(x, 1) then + 1 if negative, in float32 (SceneData.cuh:86-90)."""
    f = np.float32(np.fmod(np.float32(x), np.float32(1.0)))
    if f < 0:
        f = np.float32(f + np.float32(1.0))
    return f


def test_texel_addressing_closed_form(ct, orc):
    """Every texel carries its own (x, y); the lookup must return the texel
    uint(wrap(u) * (w - 1)), uint(wrap(v) * (h - 1)) of the reference."""
    w, h = 37, 23
    tex = np.zeros((h, w, 4), np.float32)
    tex[..., 0] = np.arange(w)[None, :]
    tex[..., 1] = np.arange(h)[:, None]
    tex[..., 3] = 1.0
    arrays = get_scene(ct, orc, 1)[1]
    a = dict(arrays)
    a["textures"] = [tex]
    a["sky_texture"] = -1
    osc = orc.Scene(a, build=False)
    rng = np.random.default_rng(5)
    us = np.concatenate([rng.uniform(-3, 3, 3000), [0.0, -0.0, 1.0, -1.0, 2.0, 0.9999999, -1e-9, 1e-9, 12345.678,
                                                   -7.5, 1e9, -1e9, 0.5]]).astype(np.float32)
    vs = np.concatenate([rng.uniform(-3, 3, 3000), [0.0, 1.0, -1.0, 0.5, 1.5, -0.25, 3.0, -3.0, 8.0, 0.1, 0.0,
                                                   0.999, 0.75]]).astype(np.float32)
    for u, v in zip(us, vs):
        ix = int(np.float32(_wrap32(u) * np.float32(w - 1)))
        iy = int(np.float32(_wrap32(v) * np.float32(h - 1)))
        got = osc.texel(0, float(u), float(v))
        assert (got[0], got[1]) == (ix, iy), (u, v, got, ix, iy)


def test_ppm_and_pfm_texel_order(ct, tmp_path):
    rng = np.random.default_rng(6)
    h, w = 5, 7
    img = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)  # rows top-down, as in the file
    ppm = tmp_path / "t.ppm"
    ppm.write_bytes(b"P6\n# comment\n%d %d\n255\n" % (w, h) + img.tobytes())
    img15 = rng.integers(0, 16, (h, w, 3), dtype=np.uint8)
    ppm15 = tmp_path / "t15.ppm"
    ppm15.write_bytes(b"P6 %d %d 15\n" % (w, h) + img15.tobytes())
    pf = rng.normal(size=(h, w, 3)).astype("<f4")  # rows bottom-up, as in the file
    pfm = tmp_path / "t.pfm"
    pfm.write_bytes(b"PF\n%d %d\n-1.0\n" % (w, h) + pf.tobytes())
    gray = rng.normal(size=(h, w)).astype(">f4")
    pfg = tmp_path / "g.pfm"
    pfg.write_bytes(b"Pf\n%d %d\n1.0\n" % (w, h) + gray.tobytes())
    s = ct.HostScene.empty()
    i0 = s.load_texture(ppm)
    i1 = s.load_texture(pfm)
    i2 = s.load_texture(pfg)
    i3 = s.load_texture(ppm15)
    assert (i0, i1, i2, i3) == (0, 1, 2, 3)
    assert s.load_texture(ppm) == 0  # already loaded
    s.add_triangle(np.eye(3), np.tile([0, 0, 1.0], 3), 0)
    t = s.arrays()["textures"]
    want0 = np.concatenate([img[::-1].astype(np.float32) / np.float32(255.0), np.ones((h, w, 1), np.float32)], 2)
    assert np.array_equal(t[0], want0)
    assert np.array_equal(t[1][..., :3], pf.astype(np.float32)) and np.all(t[1][..., 3] == 1.0)
    g = gray.astype(np.float32)
    assert all(np.array_equal(t[2][..., k], g) for k in range(3))
    want3 = ((img15[::-1].astype(np.uint32) * 255 + 7) // 15).astype(np.float32) / np.float32(255.0)
    assert np.array_equal(t[3][..., :3], want3)


def test_texture_load_errors(ct, tmp_path):
    s = ct.HostScene.empty()
    with pytest.raises(ct.CTError):
        s.load_texture(tmp_path / "missing.ppm")
    png = tmp_path / "x.png"
    png.write_bytes(b"\x89PNG\r\n\x1a\n" + b"\0" * 32)
    with pytest.raises(ct.CTError, match="unsupported"):
        s.load_texture(png)
    short = tmp_path / "short.ppm"
    short.write_bytes(b"P6 4 4 255\n" + b"\0" * 10)
    with pytest.raises(ct.CTError, match="truncated"):
        s.load_texture(short)
    with pytest.raises(ValueError):
        s.add_texture(np.zeros((4, 4, 3), np.float32))
    with pytest.raises(ct.CTError):
        s.set_sky_texture(0)  # no texture yet


# ---------------------------------------------------------- render level --
def _render(orc, arrays, opts, cam, w, h, spp=1):
    osc = orc.Scene(arrays)
    img, _ = osc.render(cam, opts, w, h, 0, spp, threads=8)
    return img


def _with(arrays, textures, mats=None, sky=-1):
    a = dict(arrays)
    a["textures"] = textures
    a["sky_texture"] = sky
    if mats is not None:
        a["materials"] = mats
    return a


def _copy_mats(ct, arrays):
    src = arrays["materials"]
    out = (ct.Material * len(src))()
    for i, m in enumerate(src):
        out[i] = m
    return out


def test_oracle_texture_invariants(ct, orc):
    """C1 at 48x48: a diffuse texture holding each material's albedo, and a
    flat normal map (0, 0, 1), leave the FP32 image bit-identical; a tilted
    normal map and a two-tone texture change it."""
    scene, arrays, _ = get_scene(ct, orc, 1)
    opts, _, _ = scene.frame_opts()
    w = h = 48
    cam = scene.camera(w, h)
    base = _render(orc, _with(arrays, []), opts, cam, w, h)
    mats = _copy_mats(ct, arrays)
    texs = []
    for i, m in enumerate(mats):
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
    same = _render(orc, _with(arrays, texs, mats), opts, cam, w, h)
    assert np.array_equal(same.view(np.uint32), base.view(np.uint32))
    # non-trivial maps change the image
    tilted = np.zeros((2, 2, 4), np.float32)
    tilted[..., 0] = 0.6
    tilted[..., 2] = 0.8
    texs2 = texs[:-1] + [tilted]
    diff = _render(orc, _with(arrays, texs2, mats), opts, cam, w, h)
    assert not np.array_equal(diff, base)
    mats3 = _copy_mats(ct, arrays)
    two = np.zeros((4, 4, 4), np.float32)
    two[:, :2, :3] = 0.9
    two[:, 2:, :3] = 0.1
    for m in mats3:
        if m.type != ct.MAT_LIGHT:
            m.diffuse_tex = 0
    changed = _render(orc, _with(arrays, [two], mats3), opts, cam, w, h)
    assert not np.array_equal(changed, base)


def test_oracle_sky_texture_invariants(ct, orc):
    """C2 at 64x36: a 1x1 sky texture of the constant sky colour gives the
    constant-sky image bit for bit; a two-tone sky changes it."""
    scene, arrays, _ = get_scene(ct, orc, 2)
    opts, _, _ = scene.frame_opts()
    w, h = 64, 36
    cam = scene.camera(w, h)
    base = _render(orc, _with(arrays, []), opts, cam, w, h)
    one = np.zeros((1, 1, 4), np.float32)
    one[0, 0, :3] = arrays["sky"]
    one[0, 0, 3] = 1.0
    same = _render(orc, _with(arrays, [one], sky=0), opts, cam, w, h)
    assert np.array_equal(same.view(np.uint32), base.view(np.uint32))
    two = np.zeros((8, 16, 4), np.float32)
    two[:4, :, :3] = 0.05   # lower half (row 0 is the bottom)
    two[4:, :, :3] = 2.0
    changed = _render(orc, _with(arrays, [two], sky=0), opts, cam, w, h)
    assert not np.array_equal(changed, base)


def test_textured_config_scene(ct, orc):
    """Config 4: the atrium's geometry with textures (same triangle count as
    config 2), a sky texture, textured floors and columns, texture
    coordinates on every vertex."""
    s4 = ct.HostScene.generate(4, 7, build=False)
    s2 = ct.HostScene.generate(2, 7, build=False)
    a4, a2 = s4.arrays(), s2.arrays()
    assert a4["indices"].shape == a2["indices"].shape
    assert np.array_equal(a4["vertices"], a2["vertices"])
    assert len(a4["textures"]) == 4 and a4["sky_texture"] == 3
    mats = a4["materials"]
    assert sum(m.diffuse_tex >= 0 for m in mats) >= 3 and sum(m.normal_tex >= 0 for m in mats) == 2
    assert a4["texcoords"].shape == (a4["vertices"].shape[0], 2)
    assert all(m.diffuse_tex == -1 and m.normal_tex == -1 for m in a2["materials"])


def _tga(path, img, rle=False, top_left=False, alpha=False):
    """Write an 8-bit TGA; img rows top-down (as seen), RGB(A) uint8."""
    h, w = img.shape[:2]
    px = img[..., [2, 1, 0, 3]] if alpha else img[..., [2, 1, 0]]
    rows = px if top_left else px[::-1]  # bottom-up unless the top-left flag is set
    bpp = 32 if alpha else 24
    desc = (0x20 if top_left else 0) | (8 if alpha else 0)
    hdr = struct.pack("<BBBHHBHHHHBB", 0, 0, 10 if rle else 2, 0, 0, 0, 0, 0, w, h, bpp, desc)
    data = rows.reshape(-1, px.shape[-1])
    if not rle:
        body = data.tobytes()
    else:  # runs of equal pixels and raw packets, at most 128 each
        out = bytearray()
        i, n = 0, len(data)
        while i < n:
            j = i + 1
            while j < n and j - i < 128 and np.array_equal(data[j], data[i]):
                j += 1
            if j - i > 1:
                out += bytes([0x80 | (j - i - 1)]) + data[i].tobytes()
                i = j
            else:
                k = i + 1
                while k < n and k - i < 128 and not (k + 1 < n and np.array_equal(data[k], data[k + 1])):
                    k += 1
                out += bytes([k - i - 1]) + data[i:k].tobytes()
                i = k
        body = bytes(out)
    path.write_bytes(hdr + body)


def test_tga_texel_order(ct, tmp_path):
    """TGA as FreeImage loads it: rows bottom-up, B G R(A) bytes, colour / 255,
    alpha 1 (Surface.cpp:77-92 keeps no alpha for 8-bit images)."""
    rng = np.random.default_rng(8)
    img = rng.integers(0, 256, (6, 9, 4), dtype=np.uint8)
    img[2:4, :, :] = img[2, 0]  # runs for the RLE packets
    s = ct.HostScene.empty()
    paths = []
    for i, kw in enumerate([{}, {"rle": True}, {"top_left": True}, {"alpha": True}, {"rle": True, "alpha": True}]):
        p = tmp_path / f"t{i}.tga"
        _tga(p, img if kw.get("alpha") else img[..., :3], **kw)
        paths.append(s.load_texture(p))
    s.add_triangle(np.eye(3), np.tile([0, 0, 1.0], 3), 0)
    t = s.arrays()["textures"]
    want = np.concatenate([img[::-1, :, :3].astype(np.float32) / np.float32(255.0), np.ones((6, 9, 1), np.float32)], 2)
    for k in paths:
        assert np.array_equal(t[k], want), k


@pytest.mark.skipif(not __import__("pathlib").Path("/root/reference/Models/sponza/textures/chain_texture.tga").exists(),
                    reason="the reference's texture files are only present where the reference is mounted")
def test_tga_reference_textures(ct):
    """The reference's own Sponza textures (24- and 32-bit, bottom-up): the
    texels are the file's pixels in file row order."""
    d = "/root/reference/Models/sponza/textures/"
    s = ct.HostScene.empty()
    names = ["chain_texture.tga", "sponza_thorn_ddn.tga"]
    idx = [s.load_texture(d + n) for n in names]
    s.add_triangle(np.eye(3), np.tile([0, 0, 1.0], 3), 0)
    t = s.arrays()["textures"]
    for i, n in zip(idx, names):
        raw = open(d + n, "rb").read()
        w, h = struct.unpack("<HH", raw[12:16])
        bpp = raw[16] // 8
        px = np.frombuffer(raw[18 + raw[0]:18 + raw[0] + w * h * bpp], np.uint8).reshape(h, w, bpp)
        assert t[i].shape == (h, w, 4)
        assert np.array_equal(t[i][..., 0], px[..., 2].astype(np.float32) / np.float32(255.0))
        assert np.array_equal(t[i][..., 2], px[..., 0].astype(np.float32) / np.float32(255.0))


def test_obj_mtl_materials(ct, tmp_path):
    """OBJ + MTL: each usemtl group gets TriangleList::processMesh's material
    (TriangleList.cpp:227-345); faces before any usemtl get the importer's
    default; emissive faces become lights; map_Kd loads the texture."""
    img = np.full((2, 2, 3), 200, np.uint8)
    _tga(tmp_path / "wood.tga", img)
    (tmp_path / "m.mtl").write_text(
        "newmtl diffuse\nKd 0.5 0.4 0.3\nKs 0.1 0.1 0.1\n"
        "newmtl shiny\nKd 0.2 0.2 0.2\nKs 0.9 0.8 0.7\nNs 50\n"
        "newmtl lamp\nKe 4 4 3\nKd 0 0 0\n"
        "newmtl textured\nmap_Kd -bm 1.0 wood.tga\nNs 1.5\n"
        "newmtl mirror\nKd 0 0 0\nKs 0.95 0.95 0.95\n"
        "newmtl glass\nKd 0 0 0\nKs 0 0 0\nTf 0.9 0.95 1\nNi 1.45\n"
        "newmtl broken\nmap_Kd missing.tga\nKd 0.3 0.3 0.3\n")
    faces = ["f 1 2 3"]
    for name in ["diffuse", "shiny", "lamp", "textured", "mirror", "glass", "broken", "unknown"]:
        faces += [f"usemtl {name}", "f 1 2 3"]
    (tmp_path / "m.obj").write_text("mtllib m.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0 0\n" + "\n".join(faces) + "\n")
    s = ct.HostScene.empty()
    s.load_model(tmp_path / "m.obj", normalize=False)
    a = s.arrays()
    mats, mi = a["materials"], a["material_idx"]
    m = [mats[int(k)] for k in mi]
    assert len(mi) == 9
    assert m[0].type == ct.MAT_LAMBERTIAN and np.allclose(m[0].albedo[:], 0.6)       # default
    assert m[1].type == ct.MAT_LAMBERTIAN and np.allclose(m[1].albedo[:], (0.5, 0.4, 0.3))
    assert m[2].type == ct.MAT_GGX and np.allclose(m[2].albedo[:], (0.9, 0.8, 0.7))  # Ks brighter than Kd
    assert m[2].roughness == 1.0  # addMaterial forces roughness 1
    assert m[3].type == ct.MAT_LIGHT and np.allclose(m[3].albedo[:], (4, 4, 3))
    assert list(a["light_indices"]) == [3]
    assert m[4].type == ct.MAT_LAMBERTIAN and m[4].diffuse_tex == 0 and np.allclose(m[4].albedo[:], 1.0)
    assert np.allclose(a["textures"][0][..., :3], 200 / 255.0)
    assert m[5].type == ct.MAT_SPECULAR and np.allclose(m[5].albedo[:], 0.95)
    assert m[6].type == ct.MAT_FRESNEL and abs(m[6].refract_idx - 1.45) < 1e-6
    assert m[7].diffuse_tex == -1 and np.allclose(m[7].albedo[:], 0.3)              # texture failed to load
    assert int(mi[8]) == int(mi[0])                                                   # unknown name: default
