"""ctypes binding of liboracle.so — the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only to compute expected values / the CPU baseline.  It never
imports the product package; scene arrays come in as numpy arrays.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "liboracle.so"


class Material(C.Structure):
    _fields_ = [("albedo", C.c_float * 3), ("refract_idx", C.c_float), ("diffuse_tex", C.c_int32),
                ("normal_tex", C.c_int32), ("mask_tex", C.c_int32), ("displace_tex", C.c_int32),
                ("type", C.c_int32), ("roughness", C.c_float), ("absorption", C.c_float * 3)]


class Camera(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("forward", C.c_float * 3), ("up", C.c_float * 3),
                ("plane_distance", C.c_float)]


class FrameOpts(C.Structure):
    _fields_ = [("max_depth", C.c_int32), ("spp", C.c_int32), ("rr_start", C.c_int32), ("accel", C.c_int32),
                ("bsdf_mode", C.c_int32), ("traversal", C.c_int32), ("shadow", C.c_int32),
                ("indirect", C.c_int32), ("normal_epsilon", C.c_float), ("dist_epsilon", C.c_float),
                ("triangle_epsilon", C.c_float), ("sort_by_material", C.c_int32),
                ("light_pdf", C.c_int32)]


class RenderStats(C.Structure):
    _fields_ = [("primary_rays", C.c_uint64), ("extension_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("samples", C.c_uint64), ("node_tests_closest", C.c_uint64), ("tri_tests_closest", C.c_uint64),
                ("node_tests_shadow", C.c_uint64), ("tri_tests_shadow", C.c_uint64)]

    def as_dict(self) -> dict:
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class Texture(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("rgba", C.POINTER(C.c_float))]


if not LIB_PATH.exists():
    raise ImportError(f"{LIB_PATH} missing: run make -C oracle")
lib = C.CDLL(str(LIB_PATH))
_P = C.c_void_p
_fptr = C.POINTER(C.c_float)
_u32p = C.POINTER(C.c_uint32)
_sig = {
    "or_scene_create": (_P, [C.c_uint32, _fptr, _fptr, C.c_uint32, _u32p, _fptr, _u32p, C.c_uint32, _u32p,
                             C.c_uint32, C.POINTER(Material), C.c_int32, _fptr]),
    "or_scene_destroy": (None, [_P]),
    "or_scene_set_textures": (C.c_int, [_P, _fptr, C.c_uint32, C.POINTER(Texture), C.c_int32]),
    "or_texel": (C.c_int, [_P, C.c_int32, C.c_float, C.c_float, _fptr]),
    "or_atan2f": (C.c_float, [C.c_float, C.c_float]),
    "or_sky_uv": (None, [C.c_float, C.c_float, C.c_float, _fptr, _fptr]),
    "or_scene_build": (C.c_int, [_P]),
    "or_scene_tree_sizes": (C.c_int, [_P, _u32p, _u32p, _u32p]),
    "or_scene_tree_export": (C.c_int, [_P, _P, _P, _u32p]),
    "or_scene_set_trees": (C.c_int, [_P, C.c_uint32, _P, C.c_uint32, _P, _u32p]),
    "or_trace": (C.c_int, [_P, C.c_int, C.c_int, C.c_uint32, _fptr, _fptr, C.c_float, C.POINTER(C.c_double),
                           C.POINTER(C.c_int32), _u32p, C.c_int]),
    "or_trace_shadow": (C.c_int, [_P, C.c_int, C.c_int, C.c_uint32, _fptr, _fptr, _fptr, C.c_float,
                                  C.POINTER(C.c_uint8), _u32p, C.c_int]),
    "or_render": (C.c_int, [_P, C.c_int, C.POINTER(Camera), C.POINTER(FrameOpts), C.c_int32, C.c_int32,
                            C.c_uint32, C.c_int32, C.c_int32, C.c_int32, _fptr, C.POINTER(C.c_double), C.c_int,
                            C.POINTER(RenderStats)]),
    "or_ct_sinf": (C.c_float, [C.c_float]),
    "or_ct_cosf": (C.c_float, [C.c_float]),
    "or_ct_expf": (C.c_float, [C.c_float]),
    "or_ct_logf": (C.c_float, [C.c_float]),
    "or_ct_acosf": (C.c_float, [C.c_float]),
    "or_ct_powf": (C.c_float, [C.c_float, C.c_float]),
    "or_ct_erff": (C.c_float, [C.c_float]),
    "or_ct_erfinvf": (C.c_float, [C.c_float]),
    "or_wang_hash": (C.c_uint32, [C.c_uint32]),
    "or_xorshift": (C.c_uint32, [C.c_uint32]),
    "or_mf_D": (C.c_double, [C.c_int, C.c_double, C.POINTER(C.c_double)]),
    "or_mf_lambda": (C.c_double, [C.c_int, C.c_double, C.POINTER(C.c_double)]),
    "or_mf_sample_vndf32": (None, [C.c_int, C.c_float, _fptr, C.c_float, C.c_float, _fptr]),
    "or_fresnel_dielectric": (C.c_double, [C.c_double, C.c_double]),
    "or_refract": (C.c_int, [C.c_int, _fptr, _fptr, C.c_float, C.POINTER(C.c_double)]),
    "or_bsdf_local": (C.c_int, [C.c_int, C.c_float, _fptr, _fptr, C.c_float, C.c_float, _fptr, _fptr, _fptr]),
}
for _n, (_r, _a) in _sig.items():
    _f = getattr(lib, _n)
    _f.restype = _r
    _f.argtypes = _a


def _fp(a):
    return a.ctypes.data_as(_fptr)


def _up(a):
    return a.ctypes.data_as(_u32p)


def as_camera(cam) -> Camera:
    """Copy any ctypes struct with the ct_camera layout."""
    c = Camera()
    C.memmove(C.byref(c), C.byref(cam), C.sizeof(Camera))
    return c


def as_opts(o) -> FrameOpts:
    r = FrameOpts()
    C.memmove(C.byref(r), C.byref(o), C.sizeof(FrameOpts))
    return r


class Scene:
    """Oracle scene from TriangleList-style arrays; builds its own trees."""

    def __init__(self, arrays: dict, build: bool = True):
        self._keep = {k: np.ascontiguousarray(v) for k, v in arrays.items() if isinstance(v, np.ndarray)}
        a = self._keep
        v = a["vertices"].astype(np.float32, copy=False)
        mats_src = arrays["materials"]
        self._mats = (Material * len(mats_src))()
        C.memmove(self._mats, C.byref(mats_src[0]) if len(mats_src) else None, C.sizeof(Material) * len(mats_src))
        li = a["light_indices"].astype(np.uint32, copy=False)
        self._li = li if li.size else np.zeros(1, np.uint32)
        sky = np.asarray(arrays.get("sky", np.zeros(3, np.float32)), np.float32)
        self._sky = sky
        self.n_tris = int(a["indices"].shape[0])
        self._h = lib.or_scene_create(v.shape[0], _fp(v), _fp(a["normals"]), self.n_tris, _up(a["indices"]),
                                      _fp(a["center_normals"]), _up(a["material_idx"]), int(li.size),
                                      _up(self._li), len(mats_src), self._mats, int(arrays.get("sky_enabled", 0)),
                                      _fp(self._sky))
        texs = arrays.get("textures") or []
        if texs or "texcoords" in arrays:
            self._tex = [np.ascontiguousarray(t, np.float32) for t in texs]
            self._tex_views = (Texture * max(len(texs), 1))()
            for i, t in enumerate(self._tex):
                self._tex_views[i] = Texture(t.shape[1], t.shape[0], _fp(t))
            tc = arrays.get("texcoords")
            self._uv = np.ascontiguousarray(tc, np.float32) if tc is not None and len(tc) else None
            if lib.or_scene_set_textures(self._h, _fp(self._uv) if self._uv is not None else None, len(texs),
                                         self._tex_views, int(arrays.get("sky_texture", -1))) != 0:
                raise ValueError("texture index out of range")
        if build:
            if lib.or_scene_build(self._h) != 0:
                raise RuntimeError("oracle build failed")

    def texel(self, texture: int, u: float, v: float) -> np.ndarray:
        out = np.zeros(4, np.float32)
        if lib.or_texel(self._h, texture, u, v, _fp(out)) != 0:
            raise ValueError("bad texture index")
        return out

    def set_trees(self, bvh, mbvh, prim_indices) -> None:
        """Traverse the given trees (an upload's, possibly edited) instead of
        the built ones; bvh / mbvh as the bytes of the scene arrays."""
        b = np.ascontiguousarray(bvh, np.uint8)
        m = np.ascontiguousarray(mbvh, np.uint8)
        pr = np.ascontiguousarray(prim_indices, np.uint32)
        if lib.or_scene_set_trees(self._h, b.size // 32, b.ctypes.data_as(_P), m.size // 128, m.ctypes.data_as(_P),
                                  _up(pr)) != 0:
            raise ValueError("tree arrays do not fit the scene")

    def trees(self):
        nb, nm, npr = C.c_uint32(), C.c_uint32(), C.c_uint32()
        lib.or_scene_tree_sizes(self._h, C.byref(nb), C.byref(nm), C.byref(npr))
        bvh = np.zeros(nb.value * 32, np.uint8)
        mbvh = np.zeros(nm.value * 128, np.uint8)
        prims = np.zeros(npr.value, np.uint32)
        lib.or_scene_tree_export(self._h, bvh.ctypes.data_as(_P), mbvh.ctypes.data_as(_P), _up(prims))
        return bvh, mbvh, prims

    def trace(self, origins, dirs, accel: int, precision: int = 32, tri_eps: float = 1e-3, counters=False,
              threads: int = 8):
        o = np.ascontiguousarray(origins, np.float32)
        d = np.ascontiguousarray(dirs, np.float32)
        n = o.shape[0]
        t = np.empty(n, np.float64)
        p = np.empty(n, np.int32)
        cnt = np.zeros((n, 2), np.uint32) if counters else None
        r = lib.or_trace(self._h, accel, precision, n, _fp(o), _fp(d), tri_eps, t.ctypes.data_as(C.POINTER(C.c_double)),
                         p.ctypes.data_as(C.POINTER(C.c_int32)), _up(cnt) if counters else None, threads)
        if r != 0:
            raise RuntimeError("or_trace failed")
        return (t, p, cnt) if counters else (t, p)

    def trace_shadow(self, origins, dirs, tmax, accel: int, precision: int = 32, tri_eps: float = 1e-3,
                     threads: int = 8, counters=False):
        o = np.ascontiguousarray(origins, np.float32)
        d = np.ascontiguousarray(dirs, np.float32)
        tm = np.ascontiguousarray(tmax, np.float32)
        n = o.shape[0]
        occ = np.empty(n, np.uint8)
        cnt = np.zeros((n, 2), np.uint32) if counters else None
        r = lib.or_trace_shadow(self._h, accel, precision, n, _fp(o), _fp(d), _fp(tm), tri_eps,
                                occ.ctypes.data_as(C.POINTER(C.c_uint8)), _up(cnt) if counters else None, threads)
        if r != 0:
            raise RuntimeError("or_trace_shadow failed")
        return (occ, cnt) if counters else occ

    def render(self, cam, opts, width: int, height: int, frame_base: int = 0, spp: int = 1, rows=None,
               precision: int = 32, threads: int = 8, accum=None):
        r0, r1 = rows if rows is not None else (0, height)
        st = RenderStats()
        c = as_camera(cam)
        o = as_opts(opts)
        if precision == 64:
            acc = accum if accum is not None else np.zeros((height, width, 4), np.float64)
            r = lib.or_render(self._h, 64, C.byref(c), C.byref(o), width, height, frame_base, spp, r0, r1, None,
                              acc.ctypes.data_as(C.POINTER(C.c_double)), threads, C.byref(st))
        else:
            acc = accum if accum is not None else np.zeros((height, width, 4), np.float32)
            r = lib.or_render(self._h, 32, C.byref(c), C.byref(o), width, height, frame_base, spp, r0, r1,
                              _fp(acc), None, threads, C.byref(st))
        if r != 0:
            raise RuntimeError("or_render failed")
        return acc, st.as_dict()

    def __del__(self):
        if getattr(self, "_h", None):
            lib.or_scene_destroy(self._h)
            self._h = None


def refract(I, N, eta: float, precision: int = 64):
    """glm::refract as the oracle's dielectric branch uses it: (direction,
    valid); valid is False under total internal reflection."""
    i = np.ascontiguousarray(I, np.float32)
    n = np.ascontiguousarray(N, np.float32)
    out = np.zeros(3, np.float64)
    ok = lib.or_refract(precision, _fp(i), _fp(n), eta, out.ctypes.data_as(C.POINTER(C.c_double)))
    return out, bool(ok)
