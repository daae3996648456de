"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only).

The scene loaders (OBJ + MTL, TGA, PPM, PFM), the scene generators, both
tree builders and the wide-tree export are built with g++ and the two
sanitizers (tests/sanitize/host_driver.cpp) and run on:
  - seeded mutations of small valid files (flipped bytes, truncation,
    inserted numbers, separators, NaN/inf, huge values), which the loaders
    must load or refuse with an error code without any sanitizer report;
  - the C1 scene: generate, binary + quad build, view, wide-node export.
"""
from __future__ import annotations

import os
import random
import shutil
import struct
import subprocess
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HOST = ROOT / "cudatracer_amd" / "csrc" / "host"
SOURCES = ["bvh.cpp", "capi_host.cpp", "cnode.cpp", "objload.cpp", "scene.cpp", "scenegen.cpp", "texload.cpp"]
FLAGS = ["-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fopenmp",
         f"-I{ROOT / 'include'}"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1", UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1",
           OMP_NUM_THREADS="2")

_OBJ = b"""mtllib m.mtl
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
usemtl a
f 1/1/1 2/2/1 3/3/1 4/4/1
f -1/-1/-2 -2/-2/-2 -3/-3/-2
usemtl b
f 5//2 6//2 7//2 8//2
f 1 5 8
f 2/3 6/4 7/1
"""
_MTL = b"""newmtl a
Kd 0.5 0.5 0.5
map_Kd t.ppm
newmtl b
Kd 0.8 0.2 0.2
Ke 2 2 2
map_bump t.tga
"""
_INSERTS = [b"-", b"9999999999", b"/", b" ", b"\n", b"0", b"-0", b"nan", b"inf", b"1e39", b"f 1 2", b"//",
            b"\xff\xff", b"65535"]


def _tga(w: int, h: int, rle: bool) -> bytes:
    hdr = bytes([0, 0, 10 if rle else 2]) + bytes(5) + struct.pack("<HHHH", 0, 0, w, h) + bytes([24, 0])
    if rle:
        return hdr + b"".join(bytes([0, i % 256, 7, 9]) for i in range(w * h))
    return hdr + bytes((i * 7) % 256 for i in range(w * h * 3))


def _seeds() -> dict[str, bytes]:
    return {"a.obj": _OBJ, "m.mtl": _MTL, "t.tga": _tga(4, 3, False), "r.tga": _tga(4, 3, True),
            "t.ppm": b"P6\n# c\n4 3\n255\n" + bytes(range(36)),
            "t.pfm": b"PF\n4 3\n-1.0\n" + struct.pack("<36f", *[i / 36 for i in range(36)])}


def _mutate(rng: random.Random, b: bytes) -> bytes:
    b = bytearray(b)
    for _ in range(rng.randint(1, 6)):
        k = rng.random()
        if k < 0.3 and b:
            b[rng.randrange(len(b))] = rng.randrange(256)
        elif k < 0.45:
            b = b[:rng.randrange(len(b) + 1)]
        elif k < 0.65:
            i = rng.randrange(len(b) + 1)
            b[i:i] = rng.choice(_INSERTS)
        elif k < 0.8 and len(b) > 2:
            i = rng.randrange(len(b))
            del b[i:min(len(b), i + rng.randint(1, 8))]
        else:
            i = rng.randrange(len(b) + 1)
            b[i:i] = bytes([rng.randrange(256)]) * rng.randint(1, 4)
    return bytes(b)


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("sanitize")
    srcs = [HOST / s for s in SOURCES] + [ROOT / "tests" / "sanitize" / "host_driver.cpp"]

    def compile_one(src: Path) -> Path:
        obj = d / (src.stem + ".o")
        subprocess.run(["g++", *FLAGS, "-c", str(src), "-o", str(obj)], check=True, capture_output=True)
        return obj

    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(compile_one, srcs))
    exe = d / "host_driver"
    subprocess.run(["g++", *FLAGS, *map(str, objs), "-o", str(exe)], check=True, capture_output=True)
    return exe


def _run(args: list[str], cwd: Path) -> str:
    r = subprocess.run(args, capture_output=True, text=True, env=ENV, cwd=cwd, timeout=600)
    report = "AddressSanitizer" in r.stderr or "runtime error" in r.stderr or "LeakSanitizer" in r.stderr
    assert r.returncode == 0 and not report, r.stderr[-4000:]
    return r.stdout


def test_loaders_on_mutated_files(driver, tmp_path):
    rng = random.Random(20261202)
    files = []
    for n in range(150):
        d = tmp_path / str(n)
        d.mkdir()
        for name, data in _seeds().items():
            (d / name).write_bytes(_mutate(rng, data) if rng.random() < 0.6 else data)
        files += [str(d / f) for f in ("a.obj", "t.tga", "r.tga", "t.ppm", "t.pfm")]
    out = _run([str(driver), "load", *files], tmp_path)
    loaded, refused = (int(x) for x in out.split()[1::2])
    print(out.strip())
    assert loaded + refused == len(files) and loaded > 0 and refused > 0


def test_generator_and_builders_c1(driver, tmp_path):
    out = _run([str(driver), "gen", "1"], tmp_path)
    print(out.strip())
    assert "rc 0" in out
