// Ray traversal on gfx950.
//
// Exact mode keeps the reference's arithmetic, node order and culling so the
// closest-hit primitive ids are bit-exact, ties included:
//   quad tree   MBVHNode::intersect / traverseMBVH[Shadow], src/BVH/MBVHNode.cuh:69-205
//   binary tree BVHNode::Intersect / traverseBVH[Shadow],   src/BVH/BVHNode.cuh:50-204
//   triangle    triangle::intersect,                        src/Core/Triangle.cuh:13-44
// Empty quad slots (count 0, always "hit" in the reference) are not pushed:
// popping them is a no-op, so the visit order of real nodes is unchanged.
//
// Execution shape (what changes is only how a wave64 interleaves its lanes;
// every lane performs the reference's sequence of pops, node tests, pushes
// and triangle tests for its own ray):
//   - persistent waves claim rays for their idle lanes with one atomic once
//     three quarters of the wave is idle, from queue ranges of their XCD
//     (see Fetcher);
//   - one step per loop iteration and lane: either pop + node test, or one
//     triangle of the current leaf, so long leaves do not stall node lanes;
//   - the per-lane stack lives in LDS, laid out [depth][thread]; pushes are
//     branch-free stores at the stack top (a store for a culled child is
//     overwritten by the next one); entries beyond the LDS depth spill to
//     the block's global region, laid out the same way.  The builder's depth limit (64) bounds the
//     stack at 97 entries; the stack bound (Queues::max_stack, 128 by
//     default) guards user-supplied trees.
#pragma once

#include "dev_common.h"

extern __shared__ uint32_t ct_lds_stack[];

namespace ctd {

CT_DEV uint32_t fbits(float f) { return __float_as_uint(f); }
CT_DEV uint32_t lane_id() { return __lane_id(); }

// Optional per-lane work counters: node and triangle tests of closest-hit
// rays, and of any-hit rays.
struct TravCount {
    uint32_t nodes, tris, snodes, stris;
};

// Moller-Trumbore with precomputed edges (edge1 = p1 - p0 and edge2 = p2 - p0
// are computed on the host with the same float subtraction).
CT_DEV bool tri_test(f3 o, f3 d, float& t, f3 p0, f3 edge1, f3 edge2, float eps) {
    const f3 h = cross(d, edge2);
    const float a = dot(edge1, h);
    if (a > -eps && a < eps) return false;
    const float f = 1.0f / a;
    const f3 s = o - p0;
    const float u = f * dot(s, h);
    if (u < 0.0f || u > 1.0f) return false;
    const f3 q = cross(s, edge1);
    const float v = f * dot(d, q);
    if (v < 0.0f || u + v > 1.0f) return false;
    const float tt = f * dot(edge2, q);
    if (tt > eps && t > tt) {
        t = tt;
        return true;
    }
    return false;
}

CT_DEV float tag(float v, uint32_t slot) { return __uint_as_float((fbits(v) & 0xFFFFFFFCu) | slot); }

// Per-lane stack: LDS positions [0, depth), global spill [depth, max_stack).
// Both are laid out [position][thread]: position p of thread i is at byte
// 4 * (p * blockDim + i) of the LDS stack or, spilled, at byte
// 4 * ((p - depth) * blockDim + i) of the block's spill region.  The stack is
// held as `top`, the LDS byte offset of the next free position, so a push or
// a pop is an add and the depth checks compare with uniform limits.  The
// spill region is addressed through a raw buffer resource (vector buffer
// loads/stores, bounds-checked by the hardware), which also keeps the
// compiler from folding LDS and spill reads into one flat access.
struct LaneStack {
    uint32_t top;      // p * stride4 + 4 * threadIdx.x for the next free position p
    uint32_t stride4;  // 4 * blockDim.x
    uint32_t lds_end;  // depth * stride4: positions from here on are spilled
    uint32_t max_end;  // max_stack * stride4
    __amdgpu_buffer_rsrc_t spill;  // this block's spill region
    bool overflow;

    CT_DEV static uint32_t& lds(uint32_t byte) {
        return *reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(ct_lds_stack) + byte);
    }
    CT_DEV void reset() {
        top = 4u * threadIdx.x;
        overflow = false;
    }
    CT_DEV bool empty() const { return top < stride4; }
    // Store at byte a: LDS, spill region, or dropped beyond max_stack (push4
    // flags the overflow).
    CT_DEV void put(uint32_t a, uint32_t v) {
        if (a < lds_end) lds(a) = v;
        else if (a < max_end) __builtin_amdgcn_raw_buffer_store_b32(v, spill, a - lds_end, 0, 0);
    }
    CT_DEV uint32_t top_pop() {
        top -= stride4;
        // LDS read at a clamped offset (always in bounds), buffer read only
        // for spilled entries
        uint32_t v = lds(top < lds_end ? top : lds_end - 4u);
        if (top >= lds_end) v = __builtin_amdgcn_raw_buffer_load_b32(spill, top - lds_end, 0, 0);
        return v;
    }
    // Push v0..v3 where k0..k3 (in that order).  Every value is stored at the
    // next free position; a culled one is overwritten by the next store or
    // left above the new top.  Branch-free when all four positions are in LDS.
    CT_DEV void push4(uint32_t v0, bool k0, uint32_t v1, bool k1, uint32_t v2, bool k2, uint32_t v3, bool k3) {
        const uint32_t a0 = top, a1 = a0 + (k0 ? stride4 : 0u), a2 = a1 + (k1 ? stride4 : 0u),
                       a3 = a2 + (k2 ? stride4 : 0u);
        top = a3 + (k3 ? stride4 : 0u);
        if (a3 < lds_end) {
            lds(a0) = v0;
            lds(a1) = v1;
            lds(a2) = v2;
            lds(a3) = v3;
        } else {
            put(a0, v0);
            put(a1, v1);
            put(a2, v2);
            put(a3, v3);
            if (top >= max_end + stride4) {  // a position at or beyond max_stack was needed
                overflow = true;
                top = max_end + 4u * threadIdx.x;
            }
        }
    }
};

// Push the hit children of a quad node, sorted by the tagged tmin (nearest
// popped first) with the reference's 5-comparator network.
CT_DEV void push_children(uint32_t hm, float m0, float m1, float m2, float m3, uint32_t e0, uint32_t e1,
                          uint32_t e2, uint32_t e3, LaneStack& st) {
    if (hm == 0u) return;
    float s0 = m0, s1 = m1, s2 = m2, s3 = m3, tmp;
    if (s0 > s1) { tmp = s0; s0 = s1; s1 = tmp; }
    if (s2 > s3) { tmp = s2; s2 = s3; s3 = tmp; }
    if (s0 > s2) { tmp = s0; s0 = s2; s2 = tmp; }
    if (s1 > s3) { tmp = s1; s1 = s3; s3 = tmp; }
    if (s2 > s3) { tmp = s2; s2 = s3; s3 = tmp; }
    auto entry = [&](uint32_t slot) { return slot == 0 ? e0 : (slot == 1 ? e1 : (slot == 2 ? e2 : e3)); };
    const uint32_t q3 = fbits(s3) & 3u, q2 = fbits(s2) & 3u, q1 = fbits(s1) & 3u, q0 = fbits(s0) & 3u;
    st.push4(entry(q3), (hm >> q3) & 1u, entry(q2), (hm >> q2) & 1u, entry(q1), (hm >> q1) & 1u, entry(q0),
             (hm >> q0) & 1u);
}

// Quad node test (MBVHNode::intersect): four slabs in the reference's float
// order, tmin tagged with the slot, the 5-comparator network, then pushes in
// the order of traverseMBVH and traverseMBVHShadow (sorted position 3 first).
CT_DEV void node4_step(const float4* __restrict__ nodes4, uint32_t e, f3 o, f3 inv, float t, LaneStack& st) {
    const float4* n = nodes4 + 8 * (size_t)e;
    const float4 bminx = n[0], bmaxx = n[1], bminy = n[2], bmaxy = n[3], bminz = n[4], bmaxz = n[5];
    const float4 ch = n[6], cn = n[7];
    float m0, m1, m2, m3, x0, x1, x2, x3;
#define CT_SLAB(MN, MX, lo, hi)                                                 \
    {                                                                           \
        float t1 = (bminx.lo - o.x) * inv.x, t2 = (bmaxx.lo - o.x) * inv.x;     \
        float mn = gmin(t1, t2), mx = gmax(t1, t2);                             \
        t1 = (bminy.lo - o.y) * inv.y;                                          \
        t2 = (bmaxy.lo - o.y) * inv.y;                                          \
        mn = gmax(mn, gmin(t1, t2));                                            \
        mx = gmin(mx, gmax(t1, t2));                                            \
        t1 = (bminz.lo - o.z) * inv.z;                                          \
        t2 = (bmaxz.lo - o.z) * inv.z;                                          \
        mn = gmax(mn, gmin(t1, t2));                                            \
        mx = gmin(mx, gmax(t1, t2));                                            \
        MN = tag(mn, hi);                                                       \
        MX = mx;                                                                \
    }
    CT_SLAB(m0, x0, x, 0u)
    CT_SLAB(m1, x1, y, 1u)
    CT_SLAB(m2, x2, z, 2u)
    CT_SLAB(m3, x3, w, 3u)
#undef CT_SLAB
    const int c0 = (int)fbits(cn.x), c1 = (int)fbits(cn.y), c2 = (int)fbits(cn.z), c3 = (int)fbits(cn.w);
    // hit mask on the tagged values; empty slots (count 0) never pushed
    const uint32_t hm = ((x0 > 0.0f && m0 <= x0 && m0 < t && c0 != 0) ? 1u : 0u) |
                        ((x1 > 0.0f && m1 <= x1 && m1 < t && c1 != 0) ? 2u : 0u) |
                        ((x2 > 0.0f && m2 <= x2 && m2 < t && c2 != 0) ? 4u : 0u) |
                        ((x3 > 0.0f && m3 <= x3 && m3 < t && c3 != 0) ? 8u : 0u);
    const uint32_t e0 = c0 == -1 ? fbits(ch.x) : (kLeafBit | fbits(ch.x));
    const uint32_t e1 = c1 == -1 ? fbits(ch.y) : (kLeafBit | fbits(ch.y));
    const uint32_t e2 = c2 == -1 ? fbits(ch.z) : (kLeafBit | fbits(ch.z));
    const uint32_t e3 = c3 == -1 ? fbits(ch.w) : (kLeafBit | fbits(ch.w));
    push_children(hm, m0, m1, m2, m3, e0, e1, e2, e3, st);
}

// Fast quad node test: the same tagged entry distances, hit flags and push
// order as node4_step, from the hardware min/max instructions (one v_min or
// v_max per slab pair, v_min3/v_max3 across axes) instead of the
// compare-and-select that (y < x) ? y : x costs.  That is exact under
// conditions that fast_ray() checks once per ray and this function once per
// node:
//  - inv is finite with |inv| > 0.5 and o is finite, so no slab value
//    (plane - o) * inv is NaN (planes are finite, see plane_flags), and a
//    slab value is zero only when plane - o is: a non-zero difference times
//    |inv| > 0.5 stays above half the smallest denormal;
//  - plane - o is never -0 (o has no +0 component, or no plane is -0), so
//    every zero slab value is +0 * inv, of the sign of inv on its axis;
//  - inside a slab pair (one axis) equal values therefore have equal bits,
//    and v_min / v_max return what (y < x) ? y : x and (x < y) ? y : x
//    return;
//  - across axes only zeros of opposite sign can tie.  For the exit distance
//    that does not matter (a zero fails x > 0 in either sign); where an
//    entry distance is zero (the origin on a face of the box) the four entry
//    distances are recomputed with the reference's fold, which keeps the
//    zero of the first axis (exact slab pairs, then (x < y) ? y : x);
//  - no key is NaN, so the reference network's last comparator (2,3) never
//    swaps: after (0,2) and (1,3), s2 = max(s0, s2) <= max(s1, s3) = s3.  The
//    entries ride through the first four comparators instead of being looked
//    up from the sorted keys' slot tags.
// A culled child in the sort; also the entry of an empty slot, whose child
// is 0xFFFFFFFF in the builders' trees (kLeafBit | 0xFFFFFFFF), which the
// upload checks (kEmptyEntry) because the sorted path relies on it.
constexpr uint32_t kNoEntry = 0xFFFFFFFFu;

CT_DEV bool fast_ray(f3 o, f3 inv, uint32_t plane_flags) {
    constexpr float kMax = 3.40282347e38f;
    const bool inv_ok = fabsf(inv.x) > 0.5f && fabsf(inv.x) <= kMax && fabsf(inv.y) > 0.5f &&
                        fabsf(inv.y) <= kMax && fabsf(inv.z) > 0.5f && fabsf(inv.z) <= kMax;
    const bool o_ok = fabsf(o.x) <= kMax && fabsf(o.y) <= kMax && fabsf(o.z) <= kMax;
    const bool z_ok = !(plane_flags & kPlaneNegZero) || (fbits(o.x) != 0u && fbits(o.y) != 0u && fbits(o.z) != 0u);
    return !(plane_flags & (kPlaneNaN | kEmptyEntry)) && inv_ok && o_ok && z_ok;
}

// nodes4f: the tree in the fast layout (k_prepare_quad): six plane rows and
// the four stack entries, seven loads per node.
CT_DEV void node4_fast(const float4* __restrict__ nodes4f, uint32_t e, f3 o, f3 inv, float t, LaneStack& st) {
    const float4* n = nodes4f + 8 * (size_t)e;
    const float4 bminx = n[0], bmaxx = n[1], bminy = n[2], bmaxy = n[3], bminz = n[4], bmaxz = n[5];
    const float4 ent = n[6];
    float m0, m1, m2, m3, x0, x1, x2, x3;
#define CT_SLABF(MN, MX, lo)                                                      \
    {                                                                             \
        const float ax = (bminx.lo - o.x) * inv.x, bx = (bmaxx.lo - o.x) * inv.x; \
        const float ay = (bminy.lo - o.y) * inv.y, by = (bmaxy.lo - o.y) * inv.y; \
        const float az = (bminz.lo - o.z) * inv.z, bz = (bmaxz.lo - o.z) * inv.z; \
        MN = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fminf(az, bz));           \
        MX = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fmaxf(az, bz));           \
    }
    CT_SLABF(m0, x0, x)
    CT_SLABF(m1, x1, y)
    CT_SLABF(m2, x2, z)
    CT_SLABF(m3, x3, w)
#undef CT_SLABF
    // keep the exit distances in this block: sunk past the branch below they
    // would hold all 24 slab values live (and be re-canonicalised there)
    __asm__ volatile("" ::"v"(x0), "v"(x1), "v"(x2), "v"(x3));
    if (fminf(fminf(fabsf(m0), fabsf(m1)), fminf(fabsf(m2), fabsf(m3))) == 0.0f) {
        // rare: the planes are reloaded (an opaque index) one child at a time
        // rather than held live beside everything else
        uint32_t r = e;
        __asm__ volatile("" : "+v"(r));
        const float* p = reinterpret_cast<const float*>(nodes4f + 8 * (size_t)r);
#pragma unroll 1
        for (uint32_t i = 0; i < 4; ++i) {
            const float v = gmax(gmax(fminf((p[i] - o.x) * inv.x, (p[4 + i] - o.x) * inv.x),
                                      fminf((p[8 + i] - o.y) * inv.y, (p[12 + i] - o.y) * inv.y)),
                                 fminf((p[16 + i] - o.z) * inv.z, (p[20 + i] - o.z) * inv.z));
            m0 = i == 0 ? v : m0;
            m1 = i == 1 ? v : m1;
            m2 = i == 2 ? v : m2;
            m3 = i == 3 ? v : m3;
        }
    }
    m0 = tag(m0, 0u);
    m1 = tag(m1, 1u);
    m2 = tag(m2, 2u);
    m3 = tag(m3, 3u);
    // An empty slot's entry is kNoEntry (its child is 0xFFFFFFFF: no
    // kEmptyEntry), so the count test is the entry test: the sorted path
    // carries kNoEntry through the network and never pushes it.
    const uint32_t e0 = fbits(ent.x), e1 = fbits(ent.y), e2 = fbits(ent.z), e3 = fbits(ent.w);
    const bool h0 = x0 > 0.0f && m0 <= x0 && m0 < t;
    const bool h1 = x1 > 0.0f && m1 <= x1 && m1 < t;
    const bool h2 = x2 > 0.0f && m2 <= x2 && m2 < t;
    const bool h3 = x3 > 0.0f && m3 <= x3 && m3 < t;
    if (!(h0 || h1 || h2 || h3)) return;
    // comparators (0,1) and (2,3) on keys and entries, then (0,2) and (1,3)
    // on entries (their keys are not compared again)
    uint32_t q0 = h0 ? e0 : kNoEntry, q1 = h1 ? e1 : kNoEntry, q2 = h2 ? e2 : kNoEntry, q3 = h3 ? e3 : kNoEntry;
    const bool s01 = m0 > m1, s23 = m2 > m3;
    const float a = s01 ? m1 : m0, b = s01 ? m0 : m1, c = s23 ? m3 : m2, d = s23 ? m2 : m3;
    const uint32_t ea = s01 ? q1 : q0, eb = s01 ? q0 : q1, ec = s23 ? q3 : q2, ed = s23 ? q2 : q3;
    const bool s02 = a > c, s13 = b > d;
    q0 = s02 ? ec : ea;
    q2 = s02 ? ea : ec;
    q1 = s13 ? ed : eb;
    q3 = s13 ? eb : ed;
    st.push4(q3, q3 != kNoEntry, q2, q2 != kNoEntry, q1, q1 != kNoEntry, q0, q0 != kNoEntry);
}

// The compact quad tree (ACCEL kTreeCompact): node i of the quad tree in 16
// words (64 bytes, half a cache line, against the 112 bytes node4_fast
// reads), written by k_prepare_quad:
//   0..2    p.x, p.y, p.z: the node's grid corner (float bits)
//   3       biased exponents of the x, y, z grid steps 2^k (bytes 0..2)
//   4, 5    low and high x offsets of slots 0..3 (byte i = slot i)
//   6, 7    low and high y offsets;  8, 9  low and high z offsets
//   10, 11  zero
//   12..15  the four stack entries (kNoEntry for an empty slot)
// A child's box is [p + lo 2^k, p + hi 2^k] per axis, quantised outwards, so
// it contains the quad node's box (why no further margin is needed:
// compact_axis in device/kernels.hip).  The ray moves into the node's grid once per node,
// t = q (2^k / d) + (p - o) / d: one conversion and one fma per plane, with
// the hardware min/max (no slab value is NaN for a compact_ray).  Entry
// distances carry the slot tag and go through node4_fast's network, so the
// children are visited nearest first, in the exact traversal's order except
// where the quantised distances order two children differently.
constexpr int kTreeCompact = 5;

// Rays the compact node test takes: finite origin, |1/d| <= 2^64 (so
// 2^k / d stays finite for the steps k <= kCompactMaxExp the encoder allows).
// The others use node4_step on the same tree.
constexpr int kCompactMaxExp = 60;
CT_DEV bool compact_ray(f3 o, f3 inv) {
    constexpr float kMax = 3.40282347e38f, kInvMax = 18446744073709551616.0f;  // 2^64
    return fabsf(inv.x) <= kInvMax && fabsf(inv.y) <= kInvMax && fabsf(inv.z) <= kInvMax && fabsf(o.x) <= kMax &&
           fabsf(o.y) <= kMax && fabsf(o.z) <= kMax;
}

CT_DEV void cnode4_step(const uint4* __restrict__ cnodes, uint32_t e, f3 o, f3 inv, float t, LaneStack& st) {
    const uint4* n = cnodes + 4 * (size_t)e;
    const uint4 h = n[0], qa = n[1], qb = n[2], ent = n[3];
    const float ax = __uint_as_float((h.w & 0xFFu) << 23) * inv.x;
    const float ay = __uint_as_float(((h.w >> 8) & 0xFFu) << 23) * inv.y;
    const float az = __uint_as_float(((h.w >> 16) & 0xFFu) << 23) * inv.z;
    const float bx = (__uint_as_float(h.x) - o.x) * inv.x, by = (__uint_as_float(h.y) - o.y) * inv.y,
                bz = (__uint_as_float(h.z) - o.z) * inv.z;
    float m0, m1, m2, m3, x0, x1, x2, x3;
#define CT_CSLAB(MN, MX, SH)                                                                              \
    {                                                                                                     \
        const float lx = fmaf((float)((qa.x >> SH) & 0xFFu), ax, bx), hx = fmaf((float)((qa.y >> SH) & 0xFFu), ax, bx); \
        const float ly = fmaf((float)((qa.z >> SH) & 0xFFu), ay, by), hy = fmaf((float)((qa.w >> SH) & 0xFFu), ay, by); \
        const float lz = fmaf((float)((qb.x >> SH) & 0xFFu), az, bz), hz = fmaf((float)((qb.y >> SH) & 0xFFu), az, bz); \
        MN = fmaxf(fmaxf(fminf(lx, hx), fminf(ly, hy)), fminf(lz, hz));                                  \
        MX = fminf(fminf(fmaxf(lx, hx), fmaxf(ly, hy)), fmaxf(lz, hz));                                  \
    }
    CT_CSLAB(m0, x0, 0u)
    CT_CSLAB(m1, x1, 8u)
    CT_CSLAB(m2, x2, 16u)
    CT_CSLAB(m3, x3, 24u)
#undef CT_CSLAB
    m0 = tag(m0, 0u);
    m1 = tag(m1, 1u);
    m2 = tag(m2, 2u);
    m3 = tag(m3, 3u);
    const bool h0 = x0 > 0.0f && m0 <= x0 && m0 < t;
    const bool h1 = x1 > 0.0f && m1 <= x1 && m1 < t;
    const bool h2 = x2 > 0.0f && m2 <= x2 && m2 < t;
    const bool h3 = x3 > 0.0f && m3 <= x3 && m3 < t;
    if (!(h0 || h1 || h2 || h3)) return;
    uint32_t q0 = h0 ? ent.x : kNoEntry, q1 = h1 ? ent.y : kNoEntry, q2 = h2 ? ent.z : kNoEntry,
             q3 = h3 ? ent.w : kNoEntry;
    const bool s01 = m0 > m1, s23 = m2 > m3;
    const float a = s01 ? m1 : m0, b = s01 ? m0 : m1, c = s23 ? m3 : m2, d = s23 ? m2 : m3;
    const uint32_t ea = s01 ? q1 : q0, eb = s01 ? q0 : q1, ec = s23 ? q3 : q2, ed = s23 ? q2 : q3;
    const bool s02 = a > c, s13 = b > d;
    q0 = s02 ? ec : ea;
    q2 = s02 ? ea : ec;
    q1 = s13 ? ed : eb;
    q3 = s13 ? eb : ed;
    st.push4(q3, q3 != kNoEntry, q2, q2 != kNoEntry, q1, q1 != kNoEntry, q0, q0 != kNoEntry);
}

// Wide compressed node (128 B, up to 8 children, layout in host/cnode.h).
// The ray is moved into the node's grid once per node, t = q * (s / d) +
// (p - o) / d, so each child plane costs one conversion and one fma (the
// decoded boxes contain the exact ones by a grid step, which covers the
// rounding of this form).  The nearest hit child is pushed last (popped
// first), the others in slot order: a full sort of 8 measured slower.
// sort = false (any-hit) pushes in slot order.
CT_DEV void fnode8_step(bool sort, const uint4* __restrict__ fnodes, uint32_t e, f3 o, f3 inv, float t,
                        LaneStack& st) {
    const uint4* n = fnodes + 8 * (size_t)e;
    const uint4 h = n[0], qx = n[1], qy = n[2], qz = n[3], ca = n[4], cb = n[5];
    const float sx = __uint_as_float((h.w & 0xFFu) << 23), sy = __uint_as_float(((h.w >> 8) & 0xFFu) << 23),
                sz = __uint_as_float(((h.w >> 16) & 0xFFu) << 23);
    // the ray in the node's grid: t = q * (s / d) + (p - o) / d
    const float ax = sx * inv.x, ay = sy * inv.y, az = sz * inv.z;
    const float bx = (__uint_as_float(h.x) - o.x) * inv.x, by = (__uint_as_float(h.y) - o.y) * inv.y,
                bz = (__uint_as_float(h.z) - o.z) * inv.z;
    const uint32_t valid = h.w >> 24;
    uint32_t hm = 0;
    float best = 3.0e38f;
    uint32_t bslot = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const uint32_t sh = 8u * (uint32_t)(i & 3);
        const uint32_t lxw = i < 4 ? qx.x : qx.y, hxw = i < 4 ? qx.z : qx.w;
        const uint32_t lyw = i < 4 ? qy.x : qy.y, hyw = i < 4 ? qy.z : qy.w;
        const uint32_t lzw = i < 4 ? qz.x : qz.y, hzw = i < 4 ? qz.z : qz.w;
        float t1 = fmaf((float)((lxw >> sh) & 0xFFu), ax, bx), t2 = fmaf((float)((hxw >> sh) & 0xFFu), ax, bx);
        float mn = gmin(t1, t2), mx = gmax(t1, t2);
        t1 = fmaf((float)((lyw >> sh) & 0xFFu), ay, by);
        t2 = fmaf((float)((hyw >> sh) & 0xFFu), ay, by);
        mn = gmax(mn, gmin(t1, t2));
        mx = gmin(mx, gmax(t1, t2));
        t1 = fmaf((float)((lzw >> sh) & 0xFFu), az, bz);
        t2 = fmaf((float)((hzw >> sh) & 0xFFu), az, bz);
        mn = gmax(mn, gmin(t1, t2));
        mx = gmin(mx, gmax(t1, t2));
        const bool hit = ((valid >> i) & 1u) && mx > 0.0f && mn <= mx && mn < t;
        hm |= hit ? (1u << i) : 0u;
        if (hit && mn < best) { best = mn; bslot = (uint32_t)i; }
    }
    if (hm == 0u) return;
    const uint32_t ent[8] = {ca.x, ca.y, ca.z, ca.w, cb.x, cb.y, cb.z, cb.w};
    if (sort) {
        // nearest hit child on top; slot 7 takes the nearest child's place
        const uint32_t a0 = (bslot & 1u) ? ent[1] : ent[0], a1 = (bslot & 1u) ? ent[3] : ent[2];
        const uint32_t a2 = (bslot & 1u) ? ent[5] : ent[4], a3 = (bslot & 1u) ? ent[7] : ent[6];
        const uint32_t b0 = (bslot & 2u) ? a1 : a0, b1 = (bslot & 2u) ? a3 : a2;
        const uint32_t near = (bslot & 4u) ? b1 : b0;
        const uint32_t h7 = (hm >> 7) & 1u;
        uint32_t x[7], k[7];
#pragma unroll
        for (int i = 0; i < 7; ++i) {
            const bool at = (uint32_t)i == bslot;
            x[i] = at ? ent[7] : ent[i];
            k[i] = at ? h7 : ((hm >> i) & 1u);
        }
        st.push4(x[0], k[0], x[1], k[1], x[2], k[2], x[3], k[3]);
        st.push4(x[4], k[4], x[5], k[5], x[6], k[6], near, true);
        return;
    }
    st.push4(ent[0], hm & 1u, ent[1], (hm >> 1) & 1u, ent[2], (hm >> 2) & 1u, ent[3], (hm >> 3) & 1u);
    st.push4(ent[4], (hm >> 4) & 1u, ent[5], (hm >> 5) & 1u, ent[6], (hm >> 6) & 1u, ent[7], (hm >> 7) & 1u);
}

// BVHNode::Intersect (BVHNode.cuh:50-62): no culling against the current t.
CT_DEV bool box2(float4 lo, float4 hi, f3 o, f3 inv, float& tmin) {
    const f3 t1 = (xyz(lo) - o) * inv;
    const f3 t2 = (xyz(hi) - o) * inv;
    const f3 mn = mk3(gmin(t1.x, t2.x), gmin(t1.y, t2.y), gmin(t1.z, t2.z));
    const f3 mx = mk3(gmax(t1.x, t2.x), gmax(t1.y, t2.y), gmax(t1.z, t2.z));
    tmin = gmax(mn.x, gmax(mn.y, mn.z));
    const float tmax = gmin(mx.x, gmin(mx.y, mx.z));
    return tmax >= 0.0f && tmin < tmax;
}

// Binary interior step (traverseBVH, BVHNode.cuh:107-139): both children
// tested; if both hit, the strictly nearer is pushed first.
CT_DEV void node2_step(const float4* __restrict__ nodes2, uint32_t lf, f3 o, f3 inv, LaneStack& st) {
    const float4 l0 = nodes2[2 * (size_t)lf], l1 = nodes2[2 * (size_t)lf + 1];
    const float4 r0 = nodes2[2 * (size_t)(lf + 1)], r1 = nodes2[2 * (size_t)(lf + 1) + 1];
    float n1, n2;
    const bool hl = box2(l0, l1, o, inv, n1);
    const bool hr = box2(r0, r1, o, inv, n2);
    const bool left_first = !(hl && hr) || (n1 < n2);
    const uint32_t a = left_first ? lf : lf + 1, b = left_first ? lf + 1 : lf;
    const bool ka = left_first ? hl : hr, kb = left_first ? hr : hl;
    st.push4(a, ka, b, kb, 0u, false, 0u, false);
}

// ----------------------------------------------------------------------
// Persistent traversal driver.  Source: how ray i is loaded; Sink: what is
// done with the result.  ACCEL 4 = quad tree, kTreeCompact = the quad tree
// in the compact layout, 8 = wide compressed tree, 2 = binary tree.
struct RayIn {
    f3 o, d;
    float tmax;
};

constexpr uint32_t kNoLeaf = 0xFFFFFFFFu;

// What a launch traces: every ray closest-hit (extend, ray-level closest
// hit), every ray any-hit (ray-level shadow rays), or two queues at once, a
// closest-hit queue and an any-hit queue, each lane in the mode of the ray
// it holds (the frame loop's bounce b: the extension rays of bounce b and
// the shadow rays of bounce b - 1 in one launch, so neither launch's tail
// idles the machine).
enum TraceMode { kClosest = 0, kAny = 1, kDual = 2 };

// The queues [0, n) and [0, n2) are cut into kFetchRanges contiguous ranges
// each, four per XCD: workgroups are dispatched to the XCDs round-robin by
// block id, so the blocks b = x (mod 8) share XCD x's L2 and trace
// neighbouring rays of the ranges x, x + 8, x + 16 and x + 24 (range = block
// id mod 32).  Range r's claim counter counts through its rays of the first
// queue, then its rays of the second; each counter is on its own cache line:
// with all counters in one line the atomics of all XCDs serialised and
// throughput halved.  A range is drained only by its own blocks (moving
// waves to other ranges at the end measured slower); every range has blocks
// as long as the grid has at least kFetchRanges of them, smaller grids move
// on to the next ranges.
//
// A wave claims rays for its idle lanes once kRefillIdle of its 64 lanes are
// idle (or all, near the end): 4% faster than waiting for the whole wave.
constexpr uint32_t kRefillIdle = 48;

struct Fetcher {
    uint32_t* base;  // counter of range 0 (range r at base + r * kFetchSlot)
    uint32_t n, n2;  // rays of the first and the second queue
    int range;
    bool steal;

    // Claim up to `want` indices of a range's claim space: [first, first +
    // count).  Index v < nfirst is ray lo + v of the first queue, otherwise
    // ray lo2 + v - nfirst of the second.  Wave-uniform.
    CT_DEV bool get(uint32_t want, uint32_t& first, uint32_t& count, uint32_t& lo, uint32_t& nfirst,
                    uint32_t& lo2) {
        for (int tries = 0; tries < kFetchRanges; ++tries) {
            lo = (uint32_t)(((uint64_t)n * (uint32_t)range) / kFetchRanges);
            const uint32_t hi = (uint32_t)(((uint64_t)n * (uint32_t)(range + 1)) / kFetchRanges);
            lo2 = (uint32_t)(((uint64_t)n2 * (uint32_t)range) / kFetchRanges);
            const uint32_t hi2 = (uint32_t)(((uint64_t)n2 * (uint32_t)(range + 1)) / kFetchRanges);
            nfirst = hi - lo;
            const uint32_t total = nfirst + (hi2 - lo2);
            uint32_t got = 0;
            if (lane_id() == 0) got = atomicAdd(base + range * kFetchSlot, want);
            got = __shfl(got, 0);
            if (got < total) {
                first = got;
                count = min(want, total - got);
                return true;
            }
            if (!steal) return false;
            range = (range + 1) % kFetchRanges;
        }
        return false;
    }
};

// n rays of the first queue and, with MODE kDual, n2 any-hit rays of the
// second.  src(i, any) loads ray i of its queue, sink(i, any, t, prim) takes
// the result: prim = the hit primitive, -1 none, -2 stack overflow (the
// error flag is raised and the ray stops).
template <int ACCEL, int MODE, bool COUNT, class Source, class Sink>
CT_DEV void trace_persistent(const DevScene& sc, float eps, uint32_t n, uint32_t n2, uint32_t* fetch, LaneStack& st,
                             TravCount& cnt, uint32_t* err, Source src, Sink sink) {
    bool alive = false, exhausted = false, fast = false;
    bool any = MODE == kAny;  // the lane's ray is an any-hit query
    uint32_t ray = 0, k = kNoLeaf;
    f3 o = mk3(0, 0, 0), d = mk3(0, 0, 0), inv = mk3(0, 0, 0);
    float t = 0.0f;
    int prim = -1;
    Fetcher fe{fetch, n, MODE == kDual ? n2 : 0u, (int)(blockIdx.x % kFetchRanges), gridDim.x < (uint32_t)kFetchRanges};
    while (true) {
        // ---- refill the idle lanes once kRefillIdle of them are idle (or
        // all of them), one claim per wave (wave-uniform)
        const uint64_t idle = __ballot(!alive);
        const uint32_t nidle = (uint32_t)__popcll(idle);
        if (!exhausted && (nidle >= kRefillIdle || idle == __ballot(true))) {
            uint32_t first = 0, count = 0, lo = 0, nfirst = 0, lo2 = 0;
            if (!fe.get(nidle, first, count, lo, nfirst, lo2)) exhausted = true;
            if (!alive && !exhausted) {
                const uint32_t rank = (uint32_t)__popcll(idle & ((1ull << lane_id()) - 1ull));
                const uint32_t v = first + rank;
                if (rank < count) {
                    if (MODE == kDual) any = v >= nfirst;
                    const uint32_t idx = (MODE == kDual && any) ? lo2 + (v - nfirst) : lo + v;
                    const RayIn r = src(idx, any);
                    ray = idx;
                    o = r.o;
                    d = r.d;
                    t = r.tmax;
                    prim = -1;
                    inv = mk3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
                    fast = (ACCEL == 4 && fast_ray(o, inv, sc.plane_flags)) ||
                           (ACCEL == kTreeCompact && compact_ray(o, inv));
                    st.reset();
                    st.push4(0u, true, 0u, false, 0u, false, 0u, false);
                    k = kNoLeaf;
                    alive = true;
                }
            }
        }
        if (__ballot(alive) == 0) {
            if (exhausted) break;
            continue;
        }
        if (!alive) continue;
        // ---- one step: pop (+ node test) and/or one triangle
        if (k == kNoLeaf) {
            const uint32_t e = st.top_pop();
            if (ACCEL == 4 || ACCEL == kTreeCompact || ACCEL == 8) {
                if (e & kLeafBit) {
                    k = e & ~kLeafBit;
                } else {
                    if (COUNT) {
                        if (any) ++cnt.snodes;
                        else ++cnt.nodes;
                    }
                    if (ACCEL == 4) {
                        // any-hit rays visit children nearest first too, like
                        // traverseMBVHShadow (MBVHNode.cuh:165-205): in a
                        // two-queue launch every lane then runs one code path
                        // (slot order for any-hit lanes measured 2% slower)
                        if (fast) node4_fast(sc.nodes4f, e, o, inv, t, st);
                        else node4_step(sc.nodes4, e, o, inv, t, st);
                    } else if (ACCEL == kTreeCompact) {
                        if (fast) cnode4_step(sc.cnodes, e, o, inv, t, st);
                        else node4_step(sc.nodes4, e, o, inv, t, st);
                    } else {
                        // any-hit lanes of a two-queue launch sort too (one
                        // code path per wave)
                        fnode8_step(MODE == kAny ? false : (MODE == kDual ? true : !any), sc.fnodes, e, o, inv, t,
                                    st);
                    }
                }
            } else {
                const float4 lo = sc.nodes2[2 * (size_t)e];
                const float4 hi = sc.nodes2[2 * (size_t)e + 1];
                const int count = (int)fbits(hi.w);
                if (count > -1) {
                    // leaf: triangles until the end-of-leaf flag (an empty
                    // leaf of a user-supplied tree has none)
                    if (count > 0) k = fbits(lo.w);
                } else {
                    if (COUNT) {
                        if (any) cnt.snodes += 2;
                        else cnt.nodes += 2;
                    }
                    node2_step(sc.nodes2, fbits(lo.w), o, inv, st);
                }
            }
        }
        bool done = false;
        if (k != kNoLeaf) {
            if (COUNT) {
                if (any) ++cnt.stris;
                else ++cnt.tris;
            }
            const float4 a = sc.tris[3 * (size_t)k];
            const float4 b = sc.tris[3 * (size_t)k + 1];
            const float4 c = sc.tris[3 * (size_t)k + 2];
            // p0 is needed only after the determinant test: left to the
            // compiler its load follows that test, a second memory round trip
            // on most triangle tests.  Held here, all three rows are in flight
            // together.
            __asm__ volatile("" ::"v"(a.x), "v"(a.y), "v"(a.z));
            const uint32_t w = fbits(a.w);
            if (tri_test(o, d, t, xyz(a), xyz(b), xyz(c), eps)) {
                prim = (int)(w & ~kLastInLeaf);
                if (any) done = true;
            }
            k = (w & kLastInLeaf) ? kNoLeaf : k + 1;
        }
        if (st.overflow) {
            atomicOr(err, 1u);
            done = true;
            prim = -2;
        }
        if (done || (k == kNoLeaf && st.empty())) {
            sink(ray, any, t, prim);
            alive = false;
        }
    }
}

// One ray per lane, traced to its end: the step loop of trace_persistent
// (the same node and triangle tests, pushes and visit order, so the same
// result for the ray) for kernels whose lanes each follow a ray of their own
// (k_tail).  t: in, the ray's extent; out, the closest hit's distance.
// Returns the hit primitive, -1 for none, -2 when the stack bound was
// exceeded (the error flag is raised).  any: stop at the first hit (a
// per-lane flag, so closest-hit and any-hit lanes step together).
template <int ACCEL>
CT_DEV int trace_single(const DevScene& sc, float eps, f3 o, f3 d, float& t, bool any, LaneStack& st,
                        uint32_t* err) {
    const f3 inv = mk3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    const bool fast = (ACCEL == 4 && fast_ray(o, inv, sc.plane_flags)) ||
                      (ACCEL == kTreeCompact && compact_ray(o, inv));
    int prim = -1;
    uint32_t k = kNoLeaf;
    st.reset();
    st.push4(0u, true, 0u, false, 0u, false, 0u, false);
    while (true) {
        if (k == kNoLeaf) {
            const uint32_t e = st.top_pop();
            if (ACCEL == 4 || ACCEL == kTreeCompact || ACCEL == 8) {
                if (e & kLeafBit) {
                    k = e & ~kLeafBit;
                } else if (ACCEL == 4) {
                    if (fast) node4_fast(sc.nodes4f, e, o, inv, t, st);
                    else node4_step(sc.nodes4, e, o, inv, t, st);
                } else if (ACCEL == kTreeCompact) {
                    if (fast) cnode4_step(sc.cnodes, e, o, inv, t, st);
                    else node4_step(sc.nodes4, e, o, inv, t, st);
                } else {
                    fnode8_step(true, sc.fnodes, e, o, inv, t, st);
                }
            } else {
                const float4 lo = sc.nodes2[2 * (size_t)e];
                const float4 hi = sc.nodes2[2 * (size_t)e + 1];
                const int count = (int)fbits(hi.w);
                if (count > -1) {
                    if (count > 0) k = fbits(lo.w);
                } else {
                    node2_step(sc.nodes2, fbits(lo.w), o, inv, st);
                }
            }
        }
        if (k != kNoLeaf) {
            const float4 a = sc.tris[3 * (size_t)k];
            const float4 b = sc.tris[3 * (size_t)k + 1];
            const float4 c = sc.tris[3 * (size_t)k + 2];
            const uint32_t w = fbits(a.w);
            if (tri_test(o, d, t, xyz(a), xyz(b), xyz(c), eps)) {
                prim = (int)(w & ~kLastInLeaf);
                if (any) break;
            }
            k = (w & kLastInLeaf) ? kNoLeaf : k + 1;
        }
        if (st.overflow) {
            atomicOr(err, 1u);
            return -2;
        }
        if (k == kNoLeaf && st.empty()) break;
    }
    return prim;
}

}  // namespace ctd
