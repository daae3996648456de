"""Shared fixtures: config scenes, their oracle twins, seeded ray sets and a
GPU context (session scoped, created only by tests marked gpu)."""
from __future__ import annotations

import ctypes as C
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long-running")


@pytest.fixture(scope="session")
def ct():
    import cudatracer_amd
    return cudatracer_amd


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    return oracle


_SCENES = {}


def get_scene(ct, orc, config: int, seed: int = 7):
    key = (config, seed)
    if key not in _SCENES:
        s = ct.HostScene.generate(config, seed)
        arrays = s.arrays()
        _SCENES[key] = (s, arrays, orc.Scene(arrays))
    return _SCENES[key]


@pytest.fixture(scope="session")
def c1(ct, orc):
    return get_scene(ct, orc, 1)


@pytest.fixture(scope="session")
def c2(ct, orc):
    return get_scene(ct, orc, 2)


@pytest.fixture(scope="session")
def gpu_ctx(ct):
    ctx = ct.Context(0)
    yield ctx
    ctx.close()


@pytest.fixture(scope="session")
def c3(ct, orc, gpu_ctx):
    """The 7.24M-triangle scene with the trees of its device build (byte-
    identical to the host build, which the builder tests pin against the
    oracle's reference-algorithm build on smaller scenes); the oracle
    traverses those trees instead of rebuilding them (a 30-plane rescan of
    7.24M triangles takes minutes).  Tests upload it again with
    gpu_ctx.upload_build(scene)."""
    scene = ct.HostScene.generate(ct.SCENE_C3, 7, build=False)
    gpu_ctx.upload_build(scene)
    trees = gpu_ctx.read_trees()
    arrays = scene.arrays()
    arrays.update(trees)
    osc = orc.Scene(arrays, build=False)
    osc.set_trees(trees["bvh"], trees["mbvh"], trees["prim_indices"])
    return scene, arrays, osc


class DeviceArray:
    """Device memory for the device-pointer entry points, through the HIP
    runtime the library itself links (torch ships its own ROCm runtime,
    which sees no GPU once the library's runtime owns it in the process)."""

    _hip = None

    def __init__(self, host: np.ndarray):
        if DeviceArray._hip is None:
            import cudatracer_amd  # noqa: F401  (loads the library and its HIP runtime)
            path = next((ln.split()[-1] for ln in open("/proc/self/maps") if "libamdhip64" in ln), None)
            if path is None:
                raise RuntimeError("the library's HIP runtime is not loaded")
            lib = C.CDLL(path)
            lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            lib.hipFree.argtypes = [C.c_void_p]
            DeviceArray._hip = lib
        self.host = np.ascontiguousarray(host)
        self.ptr = C.c_void_p()
        if DeviceArray._hip.hipMalloc(C.byref(self.ptr), self.host.nbytes) != 0:
            raise MemoryError("hipMalloc")
        self.upload()

    def upload(self):
        assert DeviceArray._hip.hipMemcpy(self.ptr, self.host.ctypes.data, self.host.nbytes, 1) == 0  # H2D

    def download(self) -> np.ndarray:
        out = np.empty_like(self.host)
        assert DeviceArray._hip.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) == 0  # D2H
        return out

    @property
    def addr(self) -> int:
        return self.ptr.value

    def __del__(self):
        if getattr(self, "ptr", None) and self.ptr.value and DeviceArray._hip is not None:
            DeviceArray._hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()


def camera_rays(cam, width: int, height: int, n: int, seed: int):
    """Jittered pinhole rays of the camera (same basis as the kernels)."""
    rng = np.random.default_rng(seed)
    pos = np.array(cam.position, np.float32)
    w = np.array(cam.forward, np.float32)
    up = np.array(cam.up, np.float32)
    u = np.cross(w, up)
    u = u / np.linalg.norm(u)
    v = np.cross(u, w)
    v = v / np.linalg.norm(v)
    hor = u * cam.plane_distance * (width / height if width > height else 1.0)
    ver = v * cam.plane_distance * (1.0 if width > height else width / height)
    px = rng.uniform(0, width, n).astype(np.float32)
    py = rng.uniform(0, height, n).astype(np.float32)
    sx = 2.0 * px / width - 1.0
    sy = 1.0 - 2.0 * py / height
    d = w[None] + hor[None] * sx[:, None] + ver[None] * sy[:, None]
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    o = np.repeat(pos[None], n, axis=0).astype(np.float32)
    return o, d


def surface_rays(arrays: dict, n: int, seed: int, offset: float = 1e-3):
    """Cosine-distributed rays leaving seeded points on random triangles."""
    rng = np.random.default_rng(seed)
    idx = arrays["indices"]
    V = arrays["vertices"].astype(np.float64)
    tri = rng.integers(0, idx.shape[0], n)
    r1 = rng.random(n)
    r2 = rng.random(n)
    flip = r1 + r2 > 1
    r1[flip], r2[flip] = 1 - r1[flip], 1 - r2[flip]
    p0, p1, p2 = V[idx[tri, 0]], V[idx[tri, 1]], V[idx[tri, 2]]
    p = (1 - r1 - r2)[:, None] * p0 + r1[:, None] * p1 + r2[:, None] * p2
    nrm = np.cross(p1 - p0, p2 - p0)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    side = np.where(rng.random(n) < 0.5, -1.0, 1.0)
    nrm *= side[:, None]
    a = np.where(np.abs(nrm[:, 0:1]) > 0.9, np.array([[0.0, 1.0, 0.0]]), np.array([[1.0, 0.0, 0.0]]))
    t = np.cross(nrm, a)
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    b = np.cross(nrm, t)
    u1, u2 = rng.random(n), rng.random(n)
    r = np.sqrt(u1)
    phi = 2 * np.pi * u2
    d = (r * np.cos(phi))[:, None] * t + (r * np.sin(phi))[:, None] * b + np.sqrt(1 - u1)[:, None] * nrm
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    o = p + offset * nrm
    return o.astype(np.float32), d.astype(np.float32)


def ray_set(scene, arrays, n: int, seed: int):
    """Half camera rays, half surface rays (the parity ray set)."""
    cam = scene.camera(640, 480)
    o1, d1 = camera_rays(cam, 640, 480, n // 2, seed)
    o2, d2 = surface_rays(arrays, n - n // 2, seed + 1)
    return np.concatenate([o1, o2]), np.concatenate([d1, d2])


class ArrayView:
    """A ct_scene_view over numpy copies of a scene's arrays (textures and
    materials optional), kept alive with the object."""

    def __init__(self, ct, arrays, textures=(), mats=None, sky=-1):
        self.keep = []

        def f(a, dt=np.float32):
            a = np.ascontiguousarray(a, dt)
            self.keep.append(a)
            return a

        v = ct.SceneView()
        vert = f(arrays["vertices"])
        v.n_vertices = vert.shape[0]
        v.vertices = vert.ctypes.data_as(C.POINTER(C.c_float))
        v.normals = f(arrays["normals"]).ctypes.data_as(C.POINTER(C.c_float))
        v.texcoords = f(arrays["texcoords"]).ctypes.data_as(C.POINTER(C.c_float))
        idx = f(arrays["indices"], np.uint32)
        v.n_triangles = idx.shape[0]
        v.indices = idx.ctypes.data_as(C.POINTER(C.c_uint32))
        v.center_normals = f(arrays["center_normals"]).ctypes.data_as(C.POINTER(C.c_float))
        v.material_idx = f(arrays["material_idx"], np.uint32).ctypes.data_as(C.POINTER(C.c_uint32))
        li = f(arrays["light_indices"], np.uint32)
        v.n_lights = li.shape[0]
        v.light_indices = li.ctypes.data_as(C.POINTER(C.c_uint32))
        src = mats if mats is not None else arrays["materials"]
        self.mats = (ct.Material * len(src))(*src)
        v.n_materials = len(src)
        v.materials = self.mats
        v.prim_indices = f(arrays["prim_indices"], np.uint32).ctypes.data_as(C.POINTER(C.c_uint32))
        bvh = f(arrays["bvh"], np.uint8)
        v.n_bvh_nodes = bvh.size // 32
        v.bvh_nodes = C.cast(bvh.ctypes.data, C.POINTER(ct.BVHNode))
        mbvh = f(arrays["mbvh"], np.uint8)
        v.n_mbvh_nodes = mbvh.size // 128
        v.mbvh_nodes = C.cast(mbvh.ctypes.data, C.POINTER(ct.MBVHNode))
        v.sky_enabled = arrays["sky_enabled"] or sky >= 0
        v.sky[:] = list(arrays["sky"])
        self.tex = [f(t) for t in textures]
        self.tv = (ct.Texture * max(len(self.tex), 1))()
        for i, t in enumerate(self.tex):
            self.tv[i] = ct.Texture(t.shape[1], t.shape[0], t.ctypes.data_as(C.POINTER(C.c_float)))
        v.n_textures = len(self.tex)
        v.textures = self.tv
        v.sky_texture = sky
        self.view = v
