"""ctypes binding of libcudatracer.so — the MI355X (gfx950) wavefront path tracer.

The native library (HIP kernels + C++ host classes behind the C ABI in
include/cudatracer.h) is the product; this module only marshals arguments.
It never falls back to another implementation: if the library is missing the
import fails, and if no GPU is present creating a Context raises CTError.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "libcudatracer.so"

# ----------------------------------------------------------------- status --
CT_OK = 0
CT_ERR_INVALID_ARG = 1
CT_ERR_OUT_OF_MEMORY = 2
CT_ERR_DEVICE = 3
CT_ERR_NO_SCENE = 4
CT_ERR_ADMISSION = 5
CT_ERR_NOT_FOUND = 6
CT_ERR_STACK_OVERFLOW = 7
CT_ERR_CANCELLED = 8
CT_ERR_INVALID_SCENE = 9

MAT_LIGHT, MAT_SPECULAR, MAT_FRESNEL, MAT_LAMBERTIAN, MAT_BECKMANN, MAT_GGX = 0, 1, 2, 3, 4, 5
ACCEL_BVH2, ACCEL_BVH4 = 0, 1
BSDF_REFERENCE, BSDF_TYPED = 0, 1
TRAV_EXACT, TRAV_FAST, TRAV_COMPACT = 0, 1, 2
LIGHTPDF_REFERENCE, LIGHTPDF_UNIFORM = 0, 1
SCENE_C1, SCENE_C2, SCENE_C3, SCENE_C2_TEXTURED = 1, 2, 3, 4
CT_BUILD_WIDE = 1


class CTError(RuntimeError):
    def __init__(self, code: int, msg: str = ""):
        super().__init__(f"cudatracer error {code}: {msg}")
        self.code = code


# ------------------------------------------------------------- structures --
class Material(C.Structure):
    _fields_ = [("albedo", C.c_float * 3), ("refract_idx", C.c_float), ("diffuse_tex", C.c_int32),
                ("normal_tex", C.c_int32), ("mask_tex", C.c_int32), ("displace_tex", C.c_int32),
                ("type", C.c_int32), ("roughness", C.c_float), ("absorption", C.c_float * 3)]


class BVHNode(C.Structure):
    _fields_ = [("min", C.c_float * 3), ("left_first", C.c_int32), ("max", C.c_float * 3), ("count", C.c_int32)]


class MBVHNode(C.Structure):
    _fields_ = [("bminx", C.c_float * 4), ("bmaxx", C.c_float * 4), ("bminy", C.c_float * 4),
                ("bmaxy", C.c_float * 4), ("bminz", C.c_float * 4), ("bmaxz", C.c_float * 4),
                ("childs", C.c_int32 * 4), ("counts", C.c_int32 * 4)]


class Texture(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("rgba", C.POINTER(C.c_float))]


class SceneView(C.Structure):
    _fields_ = [("n_vertices", C.c_uint32), ("vertices", C.POINTER(C.c_float)), ("normals", C.POINTER(C.c_float)),
                ("texcoords", C.POINTER(C.c_float)), ("n_triangles", C.c_uint32),
                ("indices", C.POINTER(C.c_uint32)), ("center_normals", C.POINTER(C.c_float)),
                ("material_idx", C.POINTER(C.c_uint32)), ("n_lights", C.c_uint32),
                ("light_indices", C.POINTER(C.c_uint32)), ("n_materials", C.c_uint32),
                ("materials", C.POINTER(Material)), ("prim_indices", C.POINTER(C.c_uint32)),
                ("n_bvh_nodes", C.c_uint32), ("bvh_nodes", C.POINTER(BVHNode)),
                ("n_mbvh_nodes", C.c_uint32), ("mbvh_nodes", C.POINTER(MBVHNode)),
                ("sky_enabled", C.c_int32), ("sky", C.c_float * 3),
                ("n_textures", C.c_uint32), ("textures", C.POINTER(Texture)), ("sky_texture", C.c_int32)]


class Camera(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("forward", C.c_float * 3), ("up", C.c_float * 3),
                ("plane_distance", C.c_float)]


class FrameOpts(C.Structure):
    _fields_ = [("max_depth", C.c_int32), ("spp", C.c_int32), ("rr_start", C.c_int32), ("accel", C.c_int32),
                ("bsdf_mode", C.c_int32), ("traversal", C.c_int32), ("shadow", C.c_int32),
                ("indirect", C.c_int32), ("normal_epsilon", C.c_float), ("dist_epsilon", C.c_float),
                ("triangle_epsilon", C.c_float), ("sort_by_material", C.c_int32),
                ("light_pdf", C.c_int32)]

    def copy(self) -> "FrameOpts":
        o = FrameOpts()
        C.pointer(o)[0] = self
        return o


class FrameStats(C.Structure):
    _fields_ = [("primary_rays", C.c_uint64), ("extension_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("samples", C.c_uint64), ("gpu_ms", C.c_double), ("extend_ms", C.c_double),
                ("shade_ms", C.c_double), ("connect_ms", C.c_double), ("other_ms", C.c_double),
                ("chunks", C.c_uint32), ("extend_launches", C.c_uint32), ("tail_ms", C.c_double)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class BuildStats(C.Structure):
    _fields_ = [("n_bvh_nodes", C.c_uint32), ("n_mbvh_nodes", C.c_uint32), ("levels", C.c_uint32),
                ("max_depth", C.c_uint32), ("long_chains", C.c_uint32), ("jump_rounds", C.c_uint32),
                ("device_ms", C.c_float), ("total_ms", C.c_float)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class DeviceOpts(C.Structure):
    _fields_ = [("device", C.c_int32), ("mem_budget", C.c_uint64), ("block_size", C.c_int32),
                ("stack_lds", C.c_int32), ("chunk_paths", C.c_int32)]


class View(C.Structure):
    _fields_ = [("camera", Camera), ("width", C.c_int32), ("height", C.c_int32), ("spp", C.c_int32),
                ("frame_base", C.c_uint32), ("accum_out", C.POINTER(C.c_float))]


class BatchStats(C.Structure):
    _fields_ = [("primary_rays", C.c_uint64), ("extension_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("samples", C.c_uint64), ("gpu_ms", C.c_double), ("chunks", C.c_int32)]


class TuneResult(C.Structure):
    _fields_ = [("block_size", C.c_int32), ("stack_lds", C.c_int32), ("chunk_paths", C.c_int32),
                ("best_ms", C.c_double), ("trials", C.c_int32)]


class ServerOpts(C.Structure):
    _fields_ = [("mem_budget", C.c_uint64), ("chunk_paths", C.c_int32), ("max_active", C.c_int32)]


class Request(C.Structure):
    _fields_ = [("camera", Camera), ("width", C.c_int32), ("height", C.c_int32), ("spp", C.c_int32),
                ("frame_base", C.c_uint32), ("opts", FrameOpts)]


class Result(C.Structure):
    _fields_ = [("status", C.c_int32), ("width", C.c_int32), ("height", C.c_int32), ("spp", C.c_int32),
                ("submit_ms", C.c_double), ("admit_ms", C.c_double), ("done_ms", C.c_double),
                ("rays", C.c_uint64)]


# ------------------------------------------------------------------ loading --
if not LIB_PATH.exists():
    raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() (make -C cudatracer_amd/csrc)")
lib = C.CDLL(str(LIB_PATH))

_P = C.c_void_p
_sig = {
    "ct_abi_version": (C.c_int, []),
    "ct_frame_opts_default": (None, [C.POINTER(FrameOpts)]),
    "ct_ctx_create": (C.c_int, [C.POINTER(DeviceOpts), C.POINTER(_P)]),
    "ct_ctx_destroy": (None, [_P]),
    "ct_last_error": (C.c_char_p, [_P]),
    "ct_device_name": (C.c_int, [_P, C.c_char_p, C.c_size_t]),
    "ct_scene_upload": (C.c_int, [_P, C.POINTER(SceneView)]),
    "ct_scene_upload_build": (C.c_int, [_P, C.POINTER(SceneView), C.c_uint32, C.POINTER(BuildStats)]),
    "ct_scene_tree_sizes": (C.c_int, [_P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "ct_scene_read_trees": (C.c_int, [_P, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p,
                                      C.c_uint32]),
    "ct_material_update": (C.c_int, [_P, C.c_uint32, C.POINTER(Material)]),
    "ct_resize": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "ct_reset": (C.c_int, [_P]),
    "ct_trace_frame": (C.c_int, [_P, C.POINTER(Camera), C.POINTER(FrameOpts), C.POINTER(FrameStats)]),
    "ct_trace_frame_counted": (C.c_int, [_P, C.POINTER(Camera), C.POINTER(FrameOpts), C.POINTER(FrameStats),
                                         C.POINTER(C.c_uint64)]),
    "ct_trace_frame_serial": (C.c_int, [_P, C.POINTER(Camera), C.POINTER(FrameOpts), C.POINTER(FrameStats)]),
    "ct_read_accum": (C.c_int, [_P, C.POINTER(C.c_float), C.c_size_t, C.POINTER(C.c_int32)]),
    "ct_write_accum": (C.c_int, [_P, C.POINTER(C.c_float), C.c_size_t, C.c_int32]),
    "ct_read_image": (C.c_int, [_P, C.POINTER(C.c_float), C.c_size_t]),
    "ct_read_image_rgba8": (C.c_int, [_P, C.POINTER(C.c_uint8), C.c_size_t]),
    "ct_trace_rays": (C.c_int, [_P, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_uint32, C.POINTER(FrameOpts),
                                C.POINTER(C.c_float), C.POINTER(C.c_int32)]),
    "ct_trace_shadow_rays": (C.c_int, [_P, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float),
                                       C.c_uint32, C.POINTER(FrameOpts), C.POINTER(C.c_uint8)]),
    "ct_trace_rays_device": (C.c_int, [_P, _P, _P, C.c_uint32, C.POINTER(FrameOpts), _P, C.POINTER(C.c_float)]),
    "ct_trace_rays_device_noalloc": (C.c_int, [_P, _P, _P, C.c_uint32, C.POINTER(FrameOpts), _P,
                                               C.POINTER(C.c_float)]),
    "ct_render_batch": (C.c_int, [_P, C.POINTER(View), C.c_int32, C.POINTER(FrameOpts), C.POINTER(BatchStats)]),
    "ct_autotune": (C.c_int, [_P, C.POINTER(Camera), C.POINTER(FrameOpts), C.c_int32, C.c_int32,
                              C.POINTER(TuneResult)]),
    "ct_server_create": (C.c_int, [_P, C.POINTER(ServerOpts), C.POINTER(_P)]),
    "ct_server_submit": (C.c_int, [_P, C.POINTER(Request), C.POINTER(C.c_uint64)]),
    "ct_server_poll": (C.c_int, [_P, C.c_uint64, C.POINTER(C.c_int32), C.POINTER(Result), C.POINTER(C.c_float),
                                 C.c_size_t]),
    "ct_server_wait": (C.c_int, [_P, C.c_uint64, C.POINTER(Result), C.POINTER(C.c_float), C.c_size_t]),
    "ct_server_cancel": (C.c_int, [_P, C.c_uint64]),
    "ct_server_clock_ms": (C.c_double, [_P]),
    "ct_server_last_error": (C.c_char_p, [_P]),
    "ct_server_destroy": (None, [_P]),
    "ct_host_scene_create": (C.c_int, [C.POINTER(_P)]),
    "ct_host_scene_generate": (C.c_int, [C.c_int32, C.c_uint64, C.POINTER(_P)]),
    "ct_host_scene_destroy": (None, [_P]),
    "ct_host_scene_add_material": (C.c_int, [_P, C.POINTER(Material), C.c_int32]),
    "ct_host_scene_add_triangle": (C.c_int, [_P, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_uint32,
                                             C.POINTER(C.c_float)]),
    "ct_host_scene_build": (C.c_int, [_P]),
    "ct_host_scene_load_model": (C.c_int, [_P, C.c_char_p, C.c_float, C.POINTER(C.c_float), C.c_int32, C.c_int32]),
    "ct_host_scene_last_error": (C.c_char_p, [_P]),
    "ct_host_scene_view": (C.c_int, [_P, C.POINTER(SceneView)]),
    "ct_host_scene_set_sky": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_float)]),
    "ct_host_scene_add_texture": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.POINTER(C.c_float)]),
    "ct_host_scene_load_texture": (C.c_int, [_P, C.c_char_p]),
    "ct_host_scene_set_sky_texture": (C.c_int, [_P, C.c_int32]),
    "ct_host_scene_camera": (C.c_int, [_P, C.c_int32, C.c_int32, C.POINTER(Camera)]),
    "ct_host_scene_frame_opts": (C.c_int, [_P, C.POINTER(FrameOpts), C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "ct_host_scene_stats": (C.c_int, [_P, C.POINTER(C.c_uint64)]),
    "ct_camera_look_at": (None, [C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_float, C.POINTER(Camera)]),
    "ct_build_wide_nodes": (C.c_int, [C.POINTER(SceneView), C.POINTER(C.c_uint32), C.c_uint32, C.POINTER(C.c_uint32),
                                      C.POINTER(C.c_uint32)]),
}
for _name, (_res, _args) in _sig.items():
    _f = getattr(lib, _name)
    _f.restype = _res
    _f.argtypes = _args

EXPORTED = tuple(_sig)


def build_wide_nodes(scene: "HostScene"):
    """The fast traversal's wide tree for a host scene: (uint32[n, 32] node
    words, worst-case stack entries)."""
    v = scene.view()
    n, bound = C.c_uint32(), C.c_uint32()
    _check(lib.ct_build_wide_nodes(C.byref(v), None, 0, C.byref(n), C.byref(bound)), None)
    out = np.zeros((n.value, 32), np.uint32)
    _check(lib.ct_build_wide_nodes(C.byref(v), out.ctypes.data_as(C.POINTER(C.c_uint32)), n.value, C.byref(n),
                                   C.byref(bound)), None)
    return out, bound.value


def _fp(a: np.ndarray, ct=C.c_float):
    return a.ctypes.data_as(C.POINTER(ct))


def default_opts() -> FrameOpts:
    o = FrameOpts()
    lib.ct_frame_opts_default(C.byref(o))
    return o


def look_at(pos, target, fov_deg: float) -> Camera:
    cam = Camera()
    p = np.asarray(pos, np.float32)
    t = np.asarray(target, np.float32)
    lib.ct_camera_look_at(_fp(p), _fp(t), C.c_float(fov_deg), C.byref(cam))
    return cam


# --------------------------------------------------------------- host scene --
class HostScene:
    """TriangleList + BVHTree + MBVHTree (C++), built on the host."""

    def __init__(self, handle):
        self._h = handle
        self._view = None

    @classmethod
    def empty(cls) -> "HostScene":
        h = _P()
        _check(lib.ct_host_scene_create(C.byref(h)), None)
        return cls(h)

    @classmethod
    def generate(cls, config: int, seed: int = 0, build: bool = True) -> "HostScene":
        """A configuration scene; build=False skips the host tree build (for
        Context.upload_build, which builds the trees on the device)."""
        h = _P()
        _check(lib.ct_host_scene_generate(config, seed, C.byref(h)), None)
        s = cls(h)
        if build:
            s.build()
        return s

    def add_material(self, albedo, mtype=MAT_LAMBERTIAN, roughness=0.0, ior=1.0, keep_roughness=False,
                     diffuse_tex: int = -1, normal_tex: int = -1) -> int:
        """TriangleList::addMaterial: roughness is forced to 1 unless keep_roughness."""
        m = Material()
        m.albedo[:] = list(map(float, albedo))
        m.refract_idx = ior
        m.diffuse_tex, m.normal_tex = diffuse_tex, normal_tex
        m.mask_tex = m.displace_tex = -1
        m.type = mtype
        m.roughness = roughness
        r = lib.ct_host_scene_add_material(self._h, C.byref(m), 1 if keep_roughness else 0)
        if r < 0:
            raise CTError(-r, "add_material")
        return r

    def add_triangle(self, p, n, material: int, uv=None) -> None:
        p = np.ascontiguousarray(p, np.float32).reshape(9)
        n = np.ascontiguousarray(n, np.float32).reshape(9)
        uvp = _fp(np.ascontiguousarray(uv, np.float32).reshape(6)) if uv is not None else None
        _check(lib.ct_host_scene_add_triangle(self._h, _fp(p), _fp(n), material, uvp), None)

    def load_model(self, path, scale: float = 1.0, mat=None, material: int = -1, normalize: bool = True) -> None:
        """TriangleList::loadModel for an OBJ file.  mat: 4x4 transform
        (row-major numpy-style matrix, applied to column vectors); None =
        identity."""
        mp = None
        if mat is not None:
            m = np.asarray(mat, np.float32).reshape(4, 4)
            mp = _fp(np.ascontiguousarray(m.T))  # column-major, like glm
        r = lib.ct_host_scene_load_model(self._h, str(path).encode(), scale, mp, material, 1 if normalize else 0)
        if r != CT_OK:
            raise CTError(r, lib.ct_host_scene_last_error(self._h).decode())

    def add_texture(self, rgba) -> int:
        """A texture from an (height, width, 4) float array, row 0 at the
        bottom (TriangleList::loadTexture's texel order).  Returns its index."""
        a = np.ascontiguousarray(rgba, np.float32)
        if a.ndim != 3 or a.shape[2] != 4:
            raise ValueError("texture must be (height, width, 4)")
        r = lib.ct_host_scene_add_texture(self._h, a.shape[1], a.shape[0], _fp(a))
        if r < 0:
            raise CTError(-r, "add_texture")
        return r

    def load_texture(self, path) -> int:
        """TriangleList::loadTexture for binary PPM (P6) / PFM files."""
        r = lib.ct_host_scene_load_texture(self._h, str(path).encode())
        if r < 0:
            raise CTError(-r, lib.ct_host_scene_last_error(self._h).decode())
        return r

    def set_sky_texture(self, texture: int) -> None:
        _check(lib.ct_host_scene_set_sky_texture(self._h, texture), None)

    def set_sky(self, enabled: bool, rgb=(0, 0, 0)) -> None:
        a = np.asarray(rgb, np.float32)
        _check(lib.ct_host_scene_set_sky(self._h, 1 if enabled else 0, _fp(a)), None)

    def build(self) -> None:
        _check(lib.ct_host_scene_build(self._h), None)
        self._view = None

    def view(self) -> SceneView:
        v = SceneView()
        _check(lib.ct_host_scene_view(self._h, C.byref(v)), None)
        self._view = v
        return v

    def camera(self, width: int, height: int) -> Camera:
        cam = Camera()
        _check(lib.ct_host_scene_camera(self._h, width, height, C.byref(cam)), None)
        return cam

    def frame_opts(self):
        o = FrameOpts()
        w, h = C.c_int32(), C.c_int32()
        _check(lib.ct_host_scene_frame_opts(self._h, C.byref(o), C.byref(w), C.byref(h)), None)
        return o, w.value, h.value

    def stats(self) -> dict:
        a = (C.c_uint64 * 8)()
        _check(lib.ct_host_scene_stats(self._h, a), None)
        keys = ["triangles", "bvh_nodes", "mbvh_nodes", "max_depth", "leaves", "lights", "bvh_build_us",
                "mbvh_build_us"]
        return dict(zip(keys, list(a)))

    def arrays(self) -> dict:
        """numpy copies of the scene view (vertices, trees, ...)."""
        v = self.view()
        nv, nt = v.n_vertices, v.n_triangles

        def arr(ptr, n, dt):
            if n == 0 or not ptr:
                return np.zeros(0, dt)
            # a typed view of the memory (np.ctypeslib.as_array on a pointer
            # copies at ~130 MB/s, 17 s for the 7.24M-triangle scene)
            addr = C.cast(ptr, C.c_void_p).value
            return np.frombuffer((C.c_uint8 * (n * np.dtype(dt).itemsize)).from_address(addr), dt).copy()

        out = {
            "vertices": arr(v.vertices, 3 * nv, np.float32).reshape(-1, 3),
            "normals": arr(v.normals, 3 * nv, np.float32).reshape(-1, 3),
            "indices": arr(v.indices, 3 * nt, np.uint32).reshape(-1, 3),
            "center_normals": arr(v.center_normals, 3 * nt, np.float32).reshape(-1, 3),
            "material_idx": arr(v.material_idx, nt, np.uint32),
            "light_indices": arr(v.light_indices, v.n_lights, np.uint32),
            "prim_indices": arr(v.prim_indices, nt, np.uint32),
            "texcoords": arr(v.texcoords, 2 * nv, np.float32).reshape(-1, 2),
            "sky_enabled": int(v.sky_enabled),
            "sky": np.array(list(v.sky), np.float32),
            "textures": [arr(v.textures[t].rgba, 4 * v.textures[t].width * v.textures[t].height,
                             np.float32).reshape(v.textures[t].height, v.textures[t].width, 4)
                         for t in range(v.n_textures)],
            "sky_texture": int(v.sky_texture),
        }
        out["materials"] = (Material * v.n_materials).from_buffer_copy(
            C.string_at(v.materials, C.sizeof(Material) * v.n_materials))
        out["bvh"] = arr(C.cast(v.bvh_nodes, C.POINTER(C.c_uint8)), 32 * v.n_bvh_nodes, np.uint8)
        out["mbvh"] = arr(C.cast(v.mbvh_nodes, C.POINTER(C.c_uint8)), 128 * v.n_mbvh_nodes, np.uint8)
        return out

    def __del__(self):
        if getattr(self, "_h", None):
            lib.ct_host_scene_destroy(self._h)
            self._h = None


def _check(r: int, ctx) -> None:
    if r != CT_OK:
        msg = lib.ct_last_error(ctx).decode() if ctx else ""
        raise CTError(r, msg)


# ------------------------------------------------------------------- context --
class Context:
    """One device context (ct_ctx).  Raises CTError(CT_ERR_DEVICE) without a GPU."""

    def __init__(self, device: int = 0, mem_budget: int = 0, block_size: int = 0, stack_lds: int = 0,
                 chunk_paths: int = 0):
        d = DeviceOpts(device, mem_budget, block_size, stack_lds, chunk_paths)
        h = _P()
        r = lib.ct_ctx_create(C.byref(d), C.byref(h))
        if r != CT_OK:
            raise CTError(r, "ct_ctx_create (no usable HIP device?)")
        self._h = h
        self.width = self.height = 0

    @property
    def handle(self):
        return self._h

    def name(self) -> str:
        b = C.create_string_buffer(256)
        lib.ct_device_name(self._h, b, 256)
        return b.value.decode()

    def upload(self, scene: HostScene) -> None:
        v = scene.view()
        _check(lib.ct_scene_upload(self._h, C.byref(v)), self._h)

    def upload_build(self, scene, wide: bool = False) -> dict:
        """Upload a scene and build its trees on the device (byte-identical to
        the host build).  scene: HostScene or SceneView."""
        v = scene.view() if isinstance(scene, HostScene) else scene
        st = BuildStats()
        _check(lib.ct_scene_upload_build(self._h, C.byref(v), CT_BUILD_WIDE if wide else 0, C.byref(st)), self._h)
        return st.as_dict()

    def read_trees(self) -> dict:
        """The current scene's binary tree, quad tree (raw bytes, host export
        layouts) and primitive order."""
        n2, n4, nt = C.c_uint32(), C.c_uint32(), C.c_uint32()
        _check(lib.ct_scene_tree_sizes(self._h, C.byref(n2), C.byref(n4), C.byref(nt)), self._h)
        bvh = np.zeros(32 * n2.value, np.uint8)
        mbvh = np.zeros(128 * n4.value, np.uint8)
        prims = np.zeros(nt.value, np.uint32)
        _check(lib.ct_scene_read_trees(self._h, bvh.ctypes.data, n2.value, mbvh.ctypes.data, n4.value,
                                       prims.ctypes.data, nt.value), self._h)
        return {"bvh": bvh, "mbvh": mbvh, "prim_indices": prims}

    def upload_view(self, v: SceneView) -> None:
        _check(lib.ct_scene_upload(self._h, C.byref(v)), self._h)

    def resize(self, w: int, h: int) -> None:
        _check(lib.ct_resize(self._h, w, h), self._h)
        self.width, self.height = w, h

    def reset(self) -> None:
        _check(lib.ct_reset(self._h), self._h)

    def material_update(self, index: int, m: Material) -> None:
        _check(lib.ct_material_update(self._h, index, C.byref(m)), self._h)

    def trace_frame(self, cam: Camera, opts: FrameOpts | None = None, stats: bool = True):
        st = FrameStats()
        _check(lib.ct_trace_frame(self._h, C.byref(cam), C.byref(opts) if opts is not None else None,
                                  C.byref(st) if stats else None), self._h)
        return st

    def trace_frame_counted(self, cam: Camera, opts: FrameOpts):
        st = FrameStats()
        w = (C.c_uint64 * 4)()
        _check(lib.ct_trace_frame_counted(self._h, C.byref(cam), C.byref(opts), C.byref(st), w), self._h)
        return st, list(w)

    def trace_frame_serial(self, cam: Camera, opts: FrameOpts):
        """A frame with its chunks one at a time and every kernel timed
        (each kernel's own device time, for measurement)."""
        st = FrameStats()
        _check(lib.ct_trace_frame_serial(self._h, C.byref(cam), C.byref(opts), C.byref(st)), self._h)
        return st

    def read_accum(self):
        out = np.empty((self.height, self.width, 4), np.float32)
        s = C.c_int32()
        _check(lib.ct_read_accum(self._h, _fp(out), out.size, C.byref(s)), self._h)
        return out, s.value

    def write_accum(self, acc: np.ndarray, samples: int) -> None:
        a = np.ascontiguousarray(acc, np.float32)
        _check(lib.ct_write_accum(self._h, _fp(a), a.size, samples), self._h)

    def read_image(self):
        out = np.empty((self.height, self.width, 4), np.float32)
        _check(lib.ct_read_image(self._h, _fp(out), out.size), self._h)
        return out

    def read_image_rgba8(self):
        out = np.empty((self.height, self.width, 4), np.uint8)
        _check(lib.ct_read_image_rgba8(self._h, out.ctypes.data_as(C.POINTER(C.c_uint8)), out.size), self._h)
        return out

    def trace_rays(self, origins, dirs, opts: FrameOpts):
        o = np.ascontiguousarray(origins, np.float32)
        d = np.ascontiguousarray(dirs, np.float32)
        n = o.shape[0]
        t = np.empty(n, np.float32)
        p = np.empty(n, np.int32)
        _check(lib.ct_trace_rays(self._h, _fp(o), _fp(d), n, C.byref(opts), _fp(t), _fp(p, C.c_int32)), self._h)
        return t, p

    def trace_shadow_rays(self, origins, dirs, tmax, opts: FrameOpts):
        o = np.ascontiguousarray(origins, np.float32)
        d = np.ascontiguousarray(dirs, np.float32)
        tm = np.ascontiguousarray(tmax, np.float32)
        n = o.shape[0]
        occ = np.empty(n, np.uint8)
        _check(lib.ct_trace_shadow_rays(self._h, _fp(o), _fp(d), _fp(tm), n, C.byref(opts), _fp(occ, C.c_uint8)),
               self._h)
        return occ

    def trace_rays_device(self, d_o: int, d_d: int, n: int, d_hits: int, opts: FrameOpts) -> float:
        ms = C.c_float()
        _check(lib.ct_trace_rays_device(self._h, C.c_void_p(d_o), C.c_void_p(d_d), n, C.byref(opts),
                                        C.c_void_p(d_hits), C.byref(ms)), self._h)
        return ms.value

    def trace_rays_device_noalloc(self, d_o: int, d_d: int, n: int, d_hits: int, opts: FrameOpts) -> float:
        """ct_trace_rays_device without device allocations (the library's
        per-GPU scratch array; 48-entry traversal stack)."""
        ms = C.c_float()
        _check(lib.ct_trace_rays_device_noalloc(self._h, C.c_void_p(d_o), C.c_void_p(d_d), n, C.byref(opts),
                                                C.c_void_p(d_hits), C.byref(ms)), self._h)
        return ms.value

    def render_batch(self, views: list, opts: FrameOpts):
        """views: list of (camera, width, height, spp, frame_base); returns list of accumulators."""
        outs = [np.zeros((h, w, 4), np.float32) for (_, w, h, _, _) in views]
        arr = (View * len(views))()
        for i, (cam, w, h, spp, fb) in enumerate(views):
            arr[i].camera = cam
            arr[i].width, arr[i].height, arr[i].spp, arr[i].frame_base = w, h, spp, fb
            arr[i].accum_out = _fp(outs[i])
        st = BatchStats()
        _check(lib.ct_render_batch(self._h, arr, len(views), C.byref(opts), C.byref(st)), self._h)
        return outs, st

    def autotune(self, cam: Camera, opts: FrameOpts, w: int = 256, h: int = 256) -> TuneResult:
        r = TuneResult()
        _check(lib.ct_autotune(self._h, C.byref(cam), C.byref(opts), w, h, C.byref(r)), self._h)
        return r

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib.ct_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


class Server:
    """Continuous-batching render server (owns the context's stream while alive)."""

    def __init__(self, ctx: Context, mem_budget: int = 0, chunk_paths: int = 0, max_active: int = 0):
        o = ServerOpts(mem_budget, chunk_paths, max_active)
        h = _P()
        _check(lib.ct_server_create(ctx.handle, C.byref(o), C.byref(h)), ctx.handle)
        self._h = h
        self._ctx = ctx
        self._sizes = {}

    def submit(self, cam: Camera, w: int, h: int, spp: int, opts: FrameOpts, frame_base: int = 0) -> int:
        r = Request()
        r.camera = cam
        r.width, r.height, r.spp, r.frame_base = w, h, spp, frame_base
        r.opts = opts
        rid = C.c_uint64()
        rc = lib.ct_server_submit(self._h, C.byref(r), C.byref(rid))
        if rc != CT_OK:
            raise CTError(rc, lib.ct_server_last_error(self._h).decode())
        self._sizes[rid.value] = (w, h)
        return rid.value

    def poll(self, rid: int):
        done = C.c_int32()
        res = Result()
        w, h = self._sizes[rid]
        out = np.empty((h, w, 4), np.float32)
        _check(lib.ct_server_poll(self._h, rid, C.byref(done), C.byref(res), _fp(out), out.size), self._ctx.handle)
        return (res, out) if done.value else None

    def wait(self, rid: int):
        res = Result()
        w, h = self._sizes[rid]
        out = np.empty((h, w, 4), np.float32)
        _check(lib.ct_server_wait(self._h, rid, C.byref(res), _fp(out), out.size), self._ctx.handle)
        return res, out

    def cancel(self, rid: int) -> None:
        _check(lib.ct_server_cancel(self._h, rid), self._ctx.handle)

    def clock_ms(self) -> float:
        return lib.ct_server_clock_ms(self._h)

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib.ct_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()
