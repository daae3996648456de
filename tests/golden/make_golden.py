"""Write the golden fixtures under tests/golden/.

rng.json   WangHash / xorshift32 / RandomFloat known-answer vectors, computed
           here with plain Python integers from the closed forms of
           src/Core/Random.cuh:10-30 (independent of the C++ code).
trees.json SHA-256 of the binary tree, quad tree and primitive order of the
           seeded config scenes as built by the product's BVHTree/MBVHTree.
           The oracle's own 30-plane builder must reproduce the same bytes.

Run from the repository root: python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
import struct
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
OUT = Path(__file__).resolve().parent


def wang_hash(s: int) -> int:
    s &= 0xFFFFFFFF
    s = (s ^ 61) ^ (s >> 16)
    s = (s * 9) & 0xFFFFFFFF
    s = s ^ (s >> 4)
    s = (s * 0x27D4EB2D) & 0xFFFFFFFF
    s = s ^ (s >> 15)
    return s


def xorshift(s: int) -> int:
    # Random.cuh:19-25 shifts left three times (13, 17, 5), unlike
    # Marsaglia's xorshift32 (13 left, 17 right, 5 left).
    s ^= (s << 13) & 0xFFFFFFFF
    s ^= (s << 17) & 0xFFFFFFFF
    s ^= (s << 5) & 0xFFFFFFFF
    return s & 0xFFFFFFFF


def f32(x: float) -> float:
    return struct.unpack("f", struct.pack("f", x))[0]


def rand_float(u: int) -> float:
    # float(u) rounded to float32, times float32(2.3283064365387e-10), rounded
    return f32(f32(float(u)) * f32(2.3283064365387e-10))


def rng_vectors() -> dict:
    seeds = [0, 1, 61, 12345, 0xDEADBEEF, 0xFFFFFFFF, 720898027, 147565741]
    out = {"wang_hash": [[s, wang_hash(s)] for s in seeds], "xorshift": [], "rand_float": []}
    for s in seeds[1:]:
        x = s
        seq = []
        for _ in range(4):
            x = xorshift(x)
            seq.append(x)
        out["xorshift"].append([s, seq])
        out["rand_float"].append([s, [rand_float(v) for v in seq]])
    # shade seeds of Kernel.cu:144 for a few (pixel, frame, pathLength)
    out["shade_seed"] = [[p, f, k, wang_hash((p * 16789 + f * 1791 + k * 720898027) & 0xFFFFFFFF)]
                         for p, f, k in [(0, 0, 0), (1, 0, 0), (1000, 3, 2), (2073599, 15, 5), (65535, 1, 7)]]
    out["primary_seed"] = [[p, f, ((p + f * 147565741) * 720898027) & 0xFFFFFFFF]
                           for p, f in [(0, 0), (1, 0), (77, 5), (2073599, 15)]]
    return out


def tree_hashes() -> dict:
    sys.path.insert(0, str(ROOT))
    import cudatracer_amd as ct
    out = {}
    for cfg in (1, 2):
        s = ct.HostScene.generate(cfg, 7)
        a = s.arrays()
        out[f"c{cfg}_seed7"] = {
            "bvh_nodes": int(a["bvh"].size // 32),
            "mbvh_nodes": int(a["mbvh"].size // 128),
            "bvh_sha256": hashlib.sha256(a["bvh"].tobytes()).hexdigest(),
            "mbvh_sha256": hashlib.sha256(a["mbvh"].tobytes()).hexdigest(),
            "prims_sha256": hashlib.sha256(a["prim_indices"].tobytes()).hexdigest(),
        }
    return out


if __name__ == "__main__":
    (OUT / "rng.json").write_text(json.dumps(rng_vectors(), indent=1))
    (OUT / "trees.json").write_text(json.dumps(tree_hashes(), indent=1))
    print("wrote", OUT / "rng.json", OUT / "trees.json")
