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
// The traversal stack is per lane: the first `lds_depth` entries live in LDS
// laid out [depth][thread] (consecutive lanes hit consecutive banks), deeper
// entries spill to a per-thread global area.  The builder's depth limit (64)
// bounds the stack at 97 entries; kMaxStack guards user-supplied trees.
#pragma once

#include "dev_common.h"

namespace ctd {

struct LaneStack {
    uint32_t* lds;    // &lds_base[threadIdx.x]; entry k at lds[k * stride]
    uint32_t* spill;  // this thread's overflow area (kMaxStack - lds_depth entries)
    int stride;
    int depth;
    int sp;
    bool overflow;

    CT_DEV void push(uint32_t v) {
        if (sp < depth) lds[sp * stride] = v;
        else if (sp < kMaxStack) spill[sp - depth] = v;
        else { overflow = true; return; }
        ++sp;
    }
    CT_DEV uint32_t pop() {
        --sp;
        return sp < depth ? lds[sp * stride] : spill[sp - depth];
    }
};

CT_DEV uint32_t fbits(float f) { return __float_as_uint(f); }

// Optional per-ray work counters (node tests, triangle tests).
struct TravCount {
    uint32_t nodes, tris;
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

// Walk one leaf run; returns true on the first hit for any-hit queries.
template <bool AnyHit>
CT_DEV bool leaf_run(const float4* __restrict__ tris, uint32_t k, f3 o, f3 d, float& t, int& prim, float eps,
                     TravCount* cnt) {
    while (true) {
        if (cnt) ++cnt->tris;
        const float4 a = tris[3 * (size_t)k];
        const float4 b = tris[3 * (size_t)k + 1];
        const float4 c = tris[3 * (size_t)k + 2];
        const uint32_t w = fbits(a.w);
        if (tri_test(o, d, t, xyz(a), xyz(b), xyz(c), eps)) {
            if (AnyHit) return true;
            prim = (int)(w & ~kLastInLeaf);
        }
        if (w & kLastInLeaf) break;
        ++k;
    }
    return false;
}

CT_DEV float tag(float v, uint32_t slot) { return __uint_as_float((fbits(v) & 0xFFFFFFFCu) | slot); }

// Exact quad-tree traversal (MBVHNode.cuh:126-205).
template <bool AnyHit>
CT_DEV bool trace_bvh4_exact(const DevScene& sc, f3 o, f3 d, float& t, int& prim, float eps, LaneStack& st,
                             TravCount* cnt) {
    const f3 inv = mk3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    st.sp = 0;
    st.push(0u);
    while (st.sp > 0) {
        const uint32_t e = st.pop();
        if (e & kLeafBit) {
            if (leaf_run<AnyHit>(sc.tris, e & ~kLeafBit, o, d, t, prim, eps, cnt)) return true;
            continue;
        }
        const float4* n = sc.nodes4 + 8 * (size_t)e;
        const float4 bminx = n[0], bmaxx = n[1], bminy = n[2], bmaxy = n[3], bminz = n[4], bmaxz = n[5];
        const float4 ch = n[6], cn = n[7];
        if (cnt) ++cnt->nodes;
        float tmin[4], tmax[4];
        const float mnx[4] = {bminx.x, bminx.y, bminx.z, bminx.w}, mxx[4] = {bmaxx.x, bmaxx.y, bmaxx.z, bmaxx.w};
        const float mny[4] = {bminy.x, bminy.y, bminy.z, bminy.w}, mxy[4] = {bmaxy.x, bmaxy.y, bmaxy.z, bmaxy.w};
        const float mnz[4] = {bminz.x, bminz.y, bminz.z, bminz.w}, mxz[4] = {bmaxz.x, bmaxz.y, bmaxz.z, bmaxz.w};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            float t1 = (mnx[i] - o.x) * inv.x, t2 = (mxx[i] - o.x) * inv.x;
            float mn = gmin(t1, t2), mx = gmax(t1, t2);
            t1 = (mny[i] - o.y) * inv.y;
            t2 = (mxy[i] - o.y) * inv.y;
            mn = gmax(mn, gmin(t1, t2));
            mx = gmin(mx, gmax(t1, t2));
            t1 = (mnz[i] - o.z) * inv.z;
            t2 = (mxz[i] - o.z) * inv.z;
            mn = gmax(mn, gmin(t1, t2));
            mx = gmin(mx, gmax(t1, t2));
            tmin[i] = tag(mn, (uint32_t)i);
            tmax[i] = mx;
        }
        bool res[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) res[i] = tmax[i] > 0.0f && tmin[i] <= tmax[i] && tmin[i] < t;
        float s0 = tmin[0], s1 = tmin[1], s2 = tmin[2], s3 = tmin[3], tmp;
        if (s0 > s1) { tmp = s0; s0 = s1; s1 = tmp; }
        if (s2 > s3) { tmp = s2; s2 = s3; s3 = tmp; }
        if (s0 > s2) { tmp = s0; s0 = s2; s2 = tmp; }
        if (s1 > s3) { tmp = s1; s1 = s3; s3 = tmp; }
        if (s2 > s3) { tmp = s2; s2 = s3; s3 = tmp; }
        const float sorted[4] = {s0, s1, s2, s3};
        const uint32_t chi[4] = {fbits(ch.x), fbits(ch.y), fbits(ch.z), fbits(ch.w)};
        const int cntv[4] = {(int)fbits(cn.x), (int)fbits(cn.y), (int)fbits(cn.z), (int)fbits(cn.w)};
#pragma unroll
        for (int i = 3; i >= 0; --i) {
            const uint32_t idx = fbits(sorted[i]) & 3u;
            const bool r = idx == 0 ? res[0] : (idx == 1 ? res[1] : (idx == 2 ? res[2] : res[3]));
            const int c = idx == 0 ? cntv[0] : (idx == 1 ? cntv[1] : (idx == 2 ? cntv[2] : cntv[3]));
            const uint32_t child = idx == 0 ? chi[0] : (idx == 1 ? chi[1] : (idx == 2 ? chi[2] : chi[3]));
            if (r && c != 0) st.push(c == -1 ? child : (kLeafBit | child));
        }
        if (st.overflow) return false;
    }
    return false;
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

// Exact binary traversal (BVHNode.cuh:82-204): both children tested, the
// strictly nearer pushed first (so popped last), leaves over
// prim_indices[leftFirst, +count).
template <bool AnyHit>
CT_DEV bool trace_bvh2_exact(const DevScene& sc, f3 o, f3 d, float& t, int& prim, float eps, LaneStack& st,
                             TravCount* cnt) {
    const f3 inv = mk3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    st.sp = 0;
    st.push(0u);
    while (st.sp > 0) {
        const uint32_t ni = st.pop();
        const float4 lo = sc.nodes2[2 * (size_t)ni];
        const float4 hi = sc.nodes2[2 * (size_t)ni + 1];
        const int count = (int)fbits(hi.w);
        const uint32_t lf = fbits(lo.w);
        if (count > -1) {
            for (int i = 0; i < count; ++i) {
                const uint32_t k = lf + (uint32_t)i;
                if (cnt) ++cnt->tris;
                const float4 a = sc.tris[3 * (size_t)k];
                const float4 b = sc.tris[3 * (size_t)k + 1];
                const float4 c = sc.tris[3 * (size_t)k + 2];
                if (tri_test(o, d, t, xyz(a), xyz(b), xyz(c), eps)) {
                    if (AnyHit) return true;
                    prim = (int)(fbits(a.w) & ~kLastInLeaf);
                }
            }
            continue;
        }
        const float4 l0 = sc.nodes2[2 * (size_t)lf], l1 = sc.nodes2[2 * (size_t)lf + 1];
        const float4 r0 = sc.nodes2[2 * (size_t)(lf + 1)], r1 = sc.nodes2[2 * (size_t)(lf + 1) + 1];
        if (cnt) cnt->nodes += 2;
        float n1, n2;
        const bool hl = box2(l0, l1, o, inv, n1);
        const bool hr = box2(r0, r1, o, inv, n2);
        if (hl && hr) {
            if (n1 < n2) { st.push(lf); st.push(lf + 1); }
            else { st.push(lf + 1); st.push(lf); }
        } else if (hl) {
            st.push(lf);
        } else if (hr) {
            st.push(lf + 1);
        }
        if (st.overflow) return false;
    }
    return false;
}

template <bool AnyHit>
CT_DEV bool trace_ray(const DevScene& sc, const TraceOpts& to, f3 o, f3 d, float& t, int& prim, LaneStack& st,
                      TravCount* cnt) {
    if (to.accel == 0) return trace_bvh2_exact<AnyHit>(sc, o, d, t, prim, to.teps, st, cnt);
    return trace_bvh4_exact<AnyHit>(sc, o, d, t, prim, to.teps, st, cnt);
}

}  // namespace ctd
