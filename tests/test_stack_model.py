"""Host model of the traversal stack's address arithmetic (LaneStack in
cudatracer_amd/csrc/device/traverse.h), checked before the GPU test that
overflows it: every LDS access stays inside the block's LDS stack
(depth x blockDim x 4 bytes), every spilled access inside the block's spill
region ((kMaxStack - depth) x blockDim x 4 bytes), pops return what was
pushed (LIFO), and the overflow flag is raised exactly when a push needs a
position at or beyond kMaxStack (the lane then stops: no access follows).
The model mirrors the kernel's integer expressions one for one (uint32
arithmetic, byte offsets)."""
from __future__ import annotations

import numpy as np
import pytest

K_MAX_STACK = 128
U32 = 0xFFFFFFFF


class LaneStackModel:
    def __init__(self, block: int, depth: int, tid: int, max_stack: int = K_MAX_STACK):
        depth = min(depth, max_stack)  # make_stack: no spill when the LDS part holds the whole bound
        self.stride4 = 4 * block
        self.lds_end = depth * self.stride4
        self.max_end = max_stack * self.stride4
        self.region = (max_stack - depth) * self.stride4
        self.tid = tid
        self.lds = {}
        self.spill = {}
        self.accesses = []  # (kind, byte offset)
        self.top = 4 * tid
        self.overflow = False

    def empty(self) -> bool:
        return self.top < self.stride4

    def _lds_at(self, a):
        assert 0 <= a < self.lds_end, ("lds", a)
        assert a % self.stride4 == 4 * self.tid  # this thread's column
        return a

    def put(self, a, v):
        if a < self.lds_end:
            self.lds[self._lds_at(a)] = v
        elif a < self.max_end:
            off = a - self.lds_end
            assert 0 <= off < self.region, ("spill", off)
            self.spill[off] = v
        # else: dropped (push4 flags the overflow)

    def top_pop(self):
        self.top = (self.top - self.stride4) & U32
        # the kernel reads LDS at a clamped offset, the spill only if spilled
        clamp = self.top if self.top < self.lds_end else self.lds_end - 4
        assert 0 <= clamp < self.lds_end
        v = self.lds.get(clamp)
        if self.top >= self.lds_end:
            off = self.top - self.lds_end
            assert 0 <= off < self.region, ("spill read", off)
            v = self.spill.get(off)
        return v

    def push4(self, vals, keep):
        a0 = self.top
        a1 = a0 + (self.stride4 if keep[0] else 0)
        a2 = a1 + (self.stride4 if keep[1] else 0)
        a3 = a2 + (self.stride4 if keep[2] else 0)
        self.top = a3 + (self.stride4 if keep[3] else 0)
        if a3 < self.lds_end:
            for a, v in zip((a0, a1, a2, a3), vals):
                self.lds[self._lds_at(a)] = v
        else:
            for a, v in zip((a0, a1, a2, a3), vals):
                self.put(a, v)
            if self.top >= self.max_end + self.stride4:
                self.overflow = True
                self.top = self.max_end + 4 * self.tid


@pytest.mark.parametrize("max_stack", [K_MAX_STACK, 48])
@pytest.mark.parametrize("block,depth", [(64, 8), (128, 16), (256, 16), (256, 24), (512, 32), (1024, 8),
                                         (256, 127)])
def test_lane_stack_bounds_and_lifo(block, depth, max_stack):
    rng = np.random.default_rng(block * 131 + depth + max_stack)
    for trial in range(60):
        tid = int(rng.integers(0, block))
        st = LaneStackModel(block, depth, tid, max_stack)
        ref = []
        st.push4([0, 0, 0, 0], [True, False, False, False])  # the root, as trace_persistent pushes it
        ref.append(0)
        bias = rng.uniform(0.3, 0.9)  # how often a step pushes rather than pops
        for step in range(2000):
            if not ref or rng.random() < bias:
                keep = [bool(k) for k in rng.random(4) < 0.6]
                vals = [int(x) for x in rng.integers(1, 1 << 30, 4)]
                need = len(ref) + sum(keep)
                st.push4(vals, keep)
                if need > max_stack:
                    assert st.overflow, (need, trial)
                    break  # the lane stops
                assert not st.overflow
                ref.extend(v for v, k in zip(vals, keep) if k)
            else:
                v = st.top_pop()
                assert v == ref.pop(), (trial, step)
            assert st.empty() == (len(ref) == 0)


def test_overflow_exactly_past_max_stack():
    """Filling one position at a time: the push that needs position
    kMaxStack (the 129th entry) is the first to flag."""
    for block, depth in ((256, 16), (64, 32)):
        st = LaneStackModel(block, depth, 5)
        for n in range(1, K_MAX_STACK + 2):
            st.push4([n, 0, 0, 0], [True, False, False, False])
            assert st.overflow == (n > K_MAX_STACK), n
            if st.overflow:
                break
        # a four-entry push from 126 entries needs positions 126..129
        st = LaneStackModel(block, depth, 0)
        for n in range(126):
            st.push4([n, 0, 0, 0], [True, False, False, False])
        st.push4([1, 2, 3, 4], [True, True, False, False])
        assert not st.overflow  # 128 entries: positions 0..127
        st.push4([5, 6, 7, 8], [False, False, False, True])
        assert st.overflow


def test_noalloc_grid_fits_scratch():
    """The allocation-free launch (runtime/noalloc.hip) cuts the persistent
    grid to the blocks whose spill regions fit the 32 MiB array: block b's
    region starts at b * (48 - depth) * blockDim * 4, so every spilled
    access of the model above lies inside the array for any launch shape the
    context accepts; shapes that fit fewer than 32 blocks (one per fetch
    range) are refused.  The control part holds the counters and flags."""
    spill_bytes, ctl_bytes, stack = 32 << 20, 64 << 10, 48
    counter_words = 192 + 2 * 32 * 96 + 64  # kCounterWords
    assert counter_words * 4 <= ctl_bytes - 256
    for block in (64, 128, 256, 512, 1024):
        for depth in (1, 8, 16, 24, 32, 47, 48, 64, 128):
            d = min(depth, stack)
            per_block = (stack - d) * block * 4
            fit = spill_bytes // per_block if per_block else 1 << 30
            if fit < 32:
                continue  # refused with CT_ERR_INVALID_ARG
            grid = min(2048, fit)
            assert grid * per_block <= spill_bytes
            top = LaneStackModel(block, depth, block - 1, stack)
            assert top.region == per_block


def _strip_stride(n: int) -> int:
    """Host rule of runtime.hip's strip_stride: the integer nearest 0.618 n
    that is coprime to n (1 for n <= 2)."""
    import math
    if n <= 2:
        return 1
    t = int(math.floor(0.6180339887498949 * n + 0.5))
    for d in range(n):
        if t + d < n and math.gcd(t + d, n) == 1:
            return t + d
        if t >= d + 1 and math.gcd(t - d - 1, n) == 1:
            return t - d - 1
    return 1


def _pixel_of(o, w: int, h: int, mul: int, rows_per_strip: int = 2):
    """pixel_of of device/dev_common.h over an array of ordinals."""
    o = np.asarray(o, dtype=np.int64)
    k, r = np.divmod(o, rows_per_strip * w)
    nfull = h // rows_per_strip
    strip = np.where(k < nfull, (k * mul) % max(nfull, 1), k)
    y0 = strip * rows_per_strip
    rows = np.minimum(rows_per_strip, h - y0)
    x, dy = np.divmod(r, rows)
    return (y0 + dy) * w + x


@pytest.mark.parametrize("rows", [2, 4])
@pytest.mark.parametrize("w,h", [(1, 1), (1, 7), (5, 3), (33, 17), (64, 6), (37, 23), (256, 256), (480, 270),
                                 (1920, 1080), (3840, 135), (3840, 2160), (7, 4 * 97 + 3)])
def test_pixel_order_is_a_bijection_and_spreads_chunks(w, h, rows):
    """pixel_of (device/dev_common.h, strips of 2 rows; 4 checked too) maps
    the ordinals [0, w h) onto the pixels one to one for every size; on
    frames of 256 rows and more each half of the ordinals (a frame's two
    wavefront chunks) covers rows spread over the whole image, so the chunks
    carry similar work."""
    mul = _strip_stride(h // rows)
    pix = _pixel_of(np.arange(w * h), w, h, mul, rows)
    assert np.array_equal(np.sort(pix), np.arange(w * h))
    if h >= 256:
        for half in (pix[: (w * h) // 2], pix[(w * h) // 2:]):
            hist = np.histogram(half // w, bins=8, range=(0, h))[0]
            assert hist.min() > 0.6 * hist.mean(), hist
