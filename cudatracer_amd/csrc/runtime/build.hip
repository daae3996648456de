// Device build of the binary SAH tree and its quad collapse.
//
// The result is byte-identical to BVHTree::constructBVH followed by
// MBVHTree::constructBVH on the host (host/bvh.cpp, after the reference's
// src/BVH/BVHNode.cu:25-191, BVHTree.cpp:22-58, MBVHNode.cu:126-160,226-277,
// MBVHTree.cpp:15-30): the same split planes, float cost expressions and
// tie-breaking, the same partition order of the primitive list and the same
// node numbering.
//
// Level-synchronous: all nodes of one depth are split together.
//   - Every primitive slot carries its box and primitive id in two ping-pong
//     buffers; level L reads buffer L & 1 and writes the other.
//   - A node of at most kChunk primitives is handled by one workgroup from
//     binning to partition.  A larger node is cut into kChunk-slot chunks
//     whose bins are reduced per node by a second kernel.
//   - The reference's 10 planes per axis cut the centroids into 11 bins (the
//     first plane with centroid <= plane).  Candidate i sees bins 1..i on the
//     left.  Box unions are exact and order-free (ordered-integer min/max
//     atomics in LDS), so the 30 costs, the chosen plane and the child boxes
//     equal the host's.
//   - Partition (BVHNode.cu:168-178) is an in-place swap scan: the left
//     primitives keep their order, the right ones are permuted by the swaps.
//     In parallel: the left primitive of rank k goes to slot k; slot p >=
//     nleft receives the primitive found by following p -> rank(p) while the
//     slot holds a left primitive (the chain of swaps that carried it there).
//   - Leaves write their primitives straight into the leaf-ordered triangle
//     records (last-in-leaf bit included).
//   - After the last level, subtree counts (bottom-up) and preorder ranks
//     (top-down) give every node the index the single-threaded recursion
//     allocates: the children of the split node of preorder rank r sit at
//     2 + 2r and 3 + 2r; a quad node's index is its preorder rank among the
//     split nodes at even depth (MergeNodes allocates depth-first).
#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "runtime.h"

namespace {

constexpr int kBins = 11;    // BVHNode.cu:8: planes 1..10, bins 1..11
constexpr int kPlanes = kBins - 1;
constexpr int kCands = 3 * kPlanes;
constexpr int kBlock = 256;
constexpr uint32_t kChunk = 2048;  // slots per workgroup
constexpr uint32_t kTiny = 256;    // nodes up to this many slots take a one-wave workgroup
constexpr uint32_t kGroup = 64;    // chunk bins reduced per block in the first stage
constexpr uint32_t kChaseCap = 256;  // longer swap chains of large nodes are resolved by pointer jumping
constexpr int kBinWords = 3 * kBins * 7;  // per axis and bin: ordered min xyz, max xyz, count
constexpr uint32_t kMaxPrims = 4;  // BVHNode.cu:6
constexpr int kMaxDepth = 64;      // BVHNode.cu:7; the root has depth 2
constexpr uint32_t kRank = 0x7FFFFFFFu;

// error bits
constexpr uint32_t kErrNoPlane = 1;    // a split without a finite candidate (extent overflow)
constexpr uint32_t kErrCount = 2;      // partition count differs from the bin count

struct NodeRec {  // BVHNode layout: bounds, first slot, count
    float mn[3];
    uint32_t first;
    float mx[3];
    uint32_t count;
};
static_assert(sizeof(NodeRec) == 32, "NodeRec");

struct Decision {
    uint32_t split, axis;
    float coord;
    uint32_t nleft;
};

struct Slots {  // ping-pong slot buffers
    float4* box[2];     // 2 float4 per slot: min (xyz), max (xyz)
    uint32_t* ref[2];   // primitive id per slot
};

struct Geo {
    const float* vert;
    const uint32_t* idx;
    float4* tris;  // leaf-ordered records, 3 float4 per slot
};

__device__ __forceinline__ uint32_t ord(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float unord(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}
// _mm_min_ps / _mm_max_ps lane semantics (AABB.h:311-314)
__device__ __forceinline__ float smin(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float smax(float a, float b) { return a > b ? a : b; }
// AABB::Area (AABB.h:113-121)
__device__ __forceinline__ float area(const float* mn, const float* mx) {
    const float e0 = mx[0] - mn[0], e1 = mx[1] - mn[1], e2 = mx[2] - mn[2];
    const float a = e0 * e1 + e0 * e2 + e1 * e2;
    return a > 0.0f ? a : 0.0f;
}
__device__ __forceinline__ bool active(uint32_t count, int level) {
    return count >= kMaxPrims && level + 2 < kMaxDepth;
}
__device__ __forceinline__ float centroid(const float4& lo, const float4& hi, uint32_t axis) {
    return axis == 0 ? (lo.x + hi.x) * 0.5f : axis == 1 ? (lo.y + hi.y) * 0.5f : (lo.z + hi.z) * 0.5f;
}

// Leaf-ordered triangle record of slot `pos` (the same values the host
// staging of ct_scene_upload produces).
__device__ void write_record(const Geo& g, uint32_t pos, uint32_t prim, bool last) {
    const uint32_t i0 = g.idx[3 * (size_t)prim], i1 = g.idx[3 * (size_t)prim + 1], i2 = g.idx[3 * (size_t)prim + 2];
    const float* p0 = g.vert + 3 * (size_t)i0;
    const float* p1 = g.vert + 3 * (size_t)i1;
    const float* p2 = g.vert + 3 * (size_t)i2;
    const float x0 = p0[0], y0 = p0[1], z0 = p0[2];
    g.tris[3 * (size_t)pos] = make_float4(x0, y0, z0, __uint_as_float(prim | (last ? 0x80000000u : 0u)));
    g.tris[3 * (size_t)pos + 1] = make_float4(p1[0] - x0, p1[1] - y0, p1[2] - z0, 0.0f);
    g.tris[3 * (size_t)pos + 2] = make_float4(p2[0] - x0, p2[1] - y0, p2[2] - z0, 0.0f);
}

// Exclusive scan over a block of BS threads (all threads call it).
template <int BS = kBlock>
__device__ uint32_t block_scan(uint32_t v, uint32_t* s_w, uint32_t& total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t t = __shfl_up(x, d, 64);
        if (lane >= d) x += t;
    }
    if (lane == 63) s_w[w] = x;
    __syncthreads();
    uint32_t base = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < BS / 64; ++k) {
        const uint32_t s = s_w[k];
        base += k < w ? s : 0u;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return base + x - v;
}

// ------------------------------------------------------------- shared state
struct BinLds {
    uint32_t bins[kBinWords];
    float planes[3 * kBins];  // planes[axis * kBins + i], i = 1..10
    float cost[kCands];
    uint32_t lcount[kCands];
    float cbox[kCands * 12];  // left min, left max, right min, right max
    Decision dec;
    uint32_t w[kBlock / 64];
};

__device__ __forceinline__ void init_bins(BinLds& s, const NodeRec& nd) {
    const int t = threadIdx.x;
    for (int w = t; w < kBinWords; w += blockDim.x) {
        const int k = w % 7;
        s.bins[w] = k < 3 ? ord(1e34f) : k < 6 ? ord(-1e34f) : 0u;
    }
    if (t < kPlanes) {
        // plane i of every axis (constant field indices keep the record in registers)
        const int i = t + 1;
        const float off = (float)i / (float)kBins;
#pragma unroll
        for (int axis = 0; axis < 3; ++axis) s.planes[axis * kBins + i] = nd.mn[axis] + (nd.mx[axis] - nd.mn[axis]) * off;
    }
}

__device__ __forceinline__ void bin_slot(BinLds& s, const float4& lo, const float4& hi) {
    const float c[3] = {(lo.x + hi.x) * 0.5f, (lo.y + hi.y) * 0.5f, (lo.z + hi.z) * 0.5f};
    const uint32_t o[6] = {ord(lo.x), ord(lo.y), ord(lo.z), ord(hi.x), ord(hi.y), ord(hi.z)};
#pragma unroll
    for (int axis = 0; axis < 3; ++axis) {
        int b = 1;
        while (b < kBins && !(c[axis] <= s.planes[axis * kBins + b])) ++b;
        uint32_t* w = s.bins + (axis * kBins + (b - 1)) * 7;
        atomicMin(w + 0, o[0]);
        atomicMin(w + 1, o[1]);
        atomicMin(w + 2, o[2]);
        atomicMax(w + 3, o[3]);
        atomicMax(w + 4, o[4]);
        atomicMax(w + 5, o[5]);
        atomicAdd(w + 6, 1u);
    }
}

// Bin a run of consecutive slots: slots that fall in the same bin as the
// previous one are merged in registers first (neighbouring slots are
// spatially close), which cuts the LDS atomics on the few hot bin words.
__device__ void bin_run(BinLds& s, const float4* box, size_t first, uint32_t n) {
    int cur[3] = {-1, -1, -1};
    float mn[3][3], mx[3][3];
    uint32_t cnt[3] = {0, 0, 0};
    auto flush = [&](int axis) {
        if (cur[axis] < 0) return;
        uint32_t* w = s.bins + (axis * kBins + (cur[axis] - 1)) * 7;
        atomicMin(w + 0, ord(mn[axis][0]));
        atomicMin(w + 1, ord(mn[axis][1]));
        atomicMin(w + 2, ord(mn[axis][2]));
        atomicMax(w + 3, ord(mx[axis][0]));
        atomicMax(w + 4, ord(mx[axis][1]));
        atomicMax(w + 5, ord(mx[axis][2]));
        atomicAdd(w + 6, cnt[axis]);
    };
    for (uint32_t q = 0; q < n; ++q) {
        const float4 lo = box[2 * (first + q)], hi = box[2 * (first + q) + 1];
        const float c[3] = {(lo.x + hi.x) * 0.5f, (lo.y + hi.y) * 0.5f, (lo.z + hi.z) * 0.5f};
#pragma unroll
        for (int axis = 0; axis < 3; ++axis) {
            int b = 1;
            while (b < kBins && !(c[axis] <= s.planes[axis * kBins + b])) ++b;
            if (b != cur[axis]) {
                flush(axis);
                cur[axis] = b;
                mn[axis][0] = lo.x;
                mn[axis][1] = lo.y;
                mn[axis][2] = lo.z;
                mx[axis][0] = hi.x;
                mx[axis][1] = hi.y;
                mx[axis][2] = hi.z;
                cnt[axis] = 1;
            } else {
                mn[axis][0] = smin(mn[axis][0], lo.x);
                mn[axis][1] = smin(mn[axis][1], lo.y);
                mn[axis][2] = smin(mn[axis][2], lo.z);
                mx[axis][0] = smax(mx[axis][0], hi.x);
                mx[axis][1] = smax(mx[axis][1], hi.y);
                mx[axis][2] = smax(mx[axis][2], hi.z);
                ++cnt[axis];
            }
        }
    }
#pragma unroll
    for (int axis = 0; axis < 3; ++axis) flush(axis);
}

// Costs of the 30 candidates (one thread each), then the first minimum in
// axis-outer, plane-inner order (BVHNode.cu:133-165).  Ends with a barrier;
// s.dec holds the decision.
__device__ __forceinline__ void decide(BinLds& s, const NodeRec& nd, NodeRec* kid, uint32_t* split_flag, Decision* dec_out,
                       uint32_t* err) {
    const int t = threadIdx.x;
    if (t < kCands) {
        const int axis = t / kPlanes, i = t % kPlanes + 1;
        float lmn[3] = {1e34f, 1e34f, 1e34f}, lmx[3] = {-1e34f, -1e34f, -1e34f};
        float rmn[3] = {1e34f, 1e34f, 1e34f}, rmx[3] = {-1e34f, -1e34f, -1e34f};
        uint32_t lc = 0, rc = 0;
        for (int b = 1; b <= i; ++b) {
            const uint32_t* w = s.bins + (axis * kBins + (b - 1)) * 7;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                lmn[k] = smin(lmn[k], unord(w[k]));
                lmx[k] = smax(lmx[k], unord(w[3 + k]));
            }
            lc += w[6];
        }
        for (int b = i + 1; b <= kBins; ++b) {
            const uint32_t* w = s.bins + (axis * kBins + (b - 1)) * 7;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                rmn[k] = smin(rmn[k], unord(w[k]));
                rmx[k] = smax(rmx[k], unord(w[3 + k]));
            }
            rc += w[6];
        }
        s.cost[t] = area(lmn, lmx) * (float)(int)lc + area(rmn, rmx) * (float)(int)rc;
        s.lcount[t] = lc;
        float* cb = s.cbox + t * 12;
        for (int k = 0; k < 3; ++k) {
            cb[k] = lmn[k];
            cb[3 + k] = lmx[k];
            cb[6 + k] = rmn[k];
            cb[9 + k] = rmx[k];
        }
    }
    __syncthreads();
    if (t < 32) {
        // first minimum below 1e34 in candidate order (strict <, as the
        // sequential scan keeps the first of equal costs): the smallest cost,
        // ties to the smaller index, NaN never chosen
        float key = INFINITY;
        int best = 255;
        if (t < kCands && s.cost[t] < 1e34f) {
            key = s.cost[t];
            best = t;
        }
#pragma unroll
        for (int d = 16; d >= 1; d >>= 1) {
            const float ok = __shfl_xor(key, d, 32);
            const int ob = __shfl_xor(best, d, 32);
            if (ok < key || (ok == key && ob < best)) {
                key = ok;
                best = ob;
            }
        }
        if (best == 255) best = -1;
        const float lowest = best >= 0 ? key : 1e34f;
    if (t == 0) {
        const float parent = area(nd.mn, nd.mx) * (float)(int)nd.count;
        Decision d{0u, 0u, 0.0f, 0u};
        if (!(parent < lowest)) {
            if (best < 0) {
                atomicOr(err, kErrNoPlane);
            } else {
                d.split = 1;
                d.axis = (uint32_t)(best / kPlanes);
                d.coord = s.planes[d.axis * kBins + best % kPlanes + 1];
                d.nleft = s.lcount[best];
                const float* cb = s.cbox + best * 12;
                NodeRec l{}, r{};
                const uint32_t nr = nd.count - d.nleft;
                for (int k = 0; k < 3; ++k) {
                    // a child without primitives keeps zero bounds (no CalculateBounds call)
                    l.mn[k] = d.nleft ? cb[k] : 0.0f;
                    l.mx[k] = d.nleft ? cb[3 + k] : 0.0f;
                    r.mn[k] = nr ? cb[6 + k] : 0.0f;
                    r.mx[k] = nr ? cb[9 + k] : 0.0f;
                }
                l.first = nd.first;
                l.count = d.nleft;
                r.first = nd.first + d.nleft;
                r.count = nr;
                kid[0] = l;
                kid[1] = r;
            }
        }
        *split_flag = d.split;
        if (dec_out) *dec_out = d;
        s.dec = d;
    }
    }
    __syncthreads();
}

// Move slot src -> dst of the next level, or into the final records if the
// destination child will not be split further.
__device__ __forceinline__ void move_slot(const Slots& b, int in, const Geo& g, uint32_t src, uint32_t dst, bool fin,
                                          bool last) {
    const uint32_t prim = b.ref[in][src];
    if (fin) {
        write_record(g, dst, prim, last);
    } else {
        b.ref[in ^ 1][dst] = prim;
        b.box[in ^ 1][2 * (size_t)dst] = b.box[in][2 * (size_t)src];
        b.box[in ^ 1][2 * (size_t)dst + 1] = b.box[in][2 * (size_t)src + 1];
    }
}

// ------------------------------------------------------------------ kernels
// Slot boxes (triangle::getBounds, Triangle.cuh:59-65 via TriangleList) and
// per-block partial bounds of the root.
__global__ __launch_bounds__(kBlock) void k_init(Geo g, Slots b, uint32_t n, float* part) {
    __shared__ float s_b[kBlock / 64][6];
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    float mn[3] = {1e34f, 1e34f, 1e34f}, mx[3] = {-1e34f, -1e34f, -1e34f};
    if (i < n) {
        const uint32_t i0 = g.idx[3 * (size_t)i], i1 = g.idx[3 * (size_t)i + 1], i2 = g.idx[3 * (size_t)i + 2];
        const float* p0 = g.vert + 3 * (size_t)i0;
        const float* p1 = g.vert + 3 * (size_t)i1;
        const float* p2 = g.vert + 3 * (size_t)i2;
        float lo[3], hi[3];
        for (int k = 0; k < 3; ++k) {
            lo[k] = fminf(p0[k], fminf(p1[k], p2[k])) - 0.001f;
            hi[k] = fmaxf(p0[k], fmaxf(p1[k], p2[k])) + 0.01f;
        }
        b.box[0][2 * (size_t)i] = make_float4(lo[0], lo[1], lo[2], 0.0f);
        b.box[0][2 * (size_t)i + 1] = make_float4(hi[0], hi[1], hi[2], 0.0f);
        b.ref[0][i] = i;
        for (int k = 0; k < 3; ++k) {
            mn[k] = lo[k];
            mx[k] = hi[k];
        }
    }
    for (int k = 0; k < 3; ++k)
        for (int d = 32; d >= 1; d >>= 1) {
            mn[k] = smin(mn[k], __shfl_xor(mn[k], d, 64));
            mx[k] = smax(mx[k], __shfl_xor(mx[k], d, 64));
        }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0)
        for (int k = 0; k < 3; ++k) {
            s_b[w][k] = mn[k];
            s_b[w][3 + k] = mx[k];
        }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int k = threadIdx.x;
        float v = s_b[0][k];
        for (int q = 1; q < kBlock / 64; ++q) v = k < 3 ? smin(v, s_b[q][k]) : smax(v, s_b[q][k]);
        part[(size_t)blockIdx.x * 6 + k] = v;
    }
}

// Root record: CalculateBounds over all primitives (BVHNode.cu:25-38).
__global__ __launch_bounds__(kBlock) void k_root(const float* part, uint32_t nparts, uint32_t n, NodeRec* root) {
    __shared__ float s_v[kBlock][6];
    float v[6] = {1e34f, 1e34f, 1e34f, -1e34f, -1e34f, -1e34f};
    for (uint32_t p = threadIdx.x; p < nparts; p += kBlock)
        for (int k = 0; k < 6; ++k) v[k] = k < 3 ? smin(v[k], part[(size_t)p * 6 + k]) : smax(v[k], part[(size_t)p * 6 + k]);
    for (int k = 0; k < 6; ++k) s_v[threadIdx.x][k] = v[k];
    __syncthreads();
    if (threadIdx.x < 6) {
        const int k = threadIdx.x;
        float r = k < 3 ? 1e34f : -1e34f;
        for (int t = 0; t < kBlock; ++t) r = k < 3 ? smin(r, s_v[t][k]) : smax(r, s_v[t][k]);
        if (k < 3) root->mn[k] = r;
        else root->mx[k - 3] = r;
        if (k == 0) {
            root->first = 0;
            root->count = n;
        }
    }
}

// Per level node: chunk counts of the nodes that try to split (count >= 4,
// depth < 64); clears the split flag.
// Per level node: which of the three node kernels splits it (tiny: one wave,
// single: one workgroup, multi: chunks), for the nodes that try to split
// (count >= 4, depth < 64); clears the split flag.
__global__ void k_plan(const NodeRec* nodes, uint32_t n, int level, uint32_t* nt, uint32_t* ns, uint32_t* nm,
                       uint32_t* split) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint32_t count = nodes[j].count;
    const bool act = active(count, level);
    const uint32_t nch = (count + kChunk - 1) / kChunk;
    nt[j] = act && count <= kTiny ? 1u : 0u;
    ns[j] = act && count > kTiny && nch == 1 ? 1u : 0u;
    nm[j] = act && nch > 1 ? nch : 0u;
    split[j] = 0;
}

__global__ void k_lists(uint32_t n, const uint32_t* nt, const uint32_t* off_t, const uint32_t* ns,
                        const uint32_t* off_s, uint32_t* tch, uint32_t* sch) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    if (nt[j]) tch[off_t[j]] = j;
    if (ns[j]) sch[off_s[j]] = j;
}

// Chunk c of the large nodes: its node is the last with chunk offset <= c.
__global__ void k_mchunks(uint32_t tm, uint32_t n, const uint32_t* off_m, uint32_t* mch_node, uint32_t* mch_k) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= tm) return;
    uint32_t lo = 0, hi = n;  // first index with off_m > c
    while (lo < hi) {
        const uint32_t mid = (lo + hi) / 2;
        if (off_m[mid] <= c) lo = mid + 1;
        else hi = mid;
    }
    const uint32_t j = lo - 1;
    mch_node[c] = j;
    mch_k[c] = c - off_m[j];
}

// A node of at most CAP primitives, from bins to partition, by one
// workgroup of BS threads.
template <int BS, uint32_t CAP>
__global__ __launch_bounds__(BS) void k_single(const NodeRec* nodes, int level, const uint32_t* sch, Slots b, Geo g,
                                               NodeRec* kid, uint32_t* split, uint32_t* err) {
    __shared__ BinLds s;
    __shared__ uint32_t s_rf[CAP];
    const uint32_t j = sch[blockIdx.x];
    const NodeRec nd = nodes[j];
    const uint32_t first = nd.first, count = nd.count;
    const int in = level & 1;
    const float4* ibox = b.box[in];
    constexpr int kPer = (int)(CAP / BS);
    float cen[kPer][3];  // slot centroids, kept for the partition test
    init_bins(s, nd);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const uint32_t r = k * BS + threadIdx.x;
        if (r < count) {
            const float4 lo = ibox[2 * (size_t)(first + r)], hi = ibox[2 * (size_t)(first + r) + 1];
            bin_slot(s, lo, hi);
            cen[k][0] = (lo.x + hi.x) * 0.5f;
            cen[k][1] = (lo.y + hi.y) * 0.5f;
            cen[k][2] = (lo.z + hi.z) * 0.5f;
        }
    }
    __syncthreads();
    decide(s, nd, kid + 2 * (size_t)j, split + j, nullptr, err);
    const Decision d = s.dec;
    if (!d.split) {  // SAH leaf
        for (uint32_t r = threadIdx.x; r < count; r += BS)
            write_record(g, first + r, b.ref[in][first + r], r == count - 1);
        return;
    }
    uint32_t running = 0;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        if ((uint32_t)k * BS < count) {  // block-uniform
            const uint32_t r = k * BS + threadIdx.x;
            const float cv = d.axis == 0 ? cen[k][0] : d.axis == 1 ? cen[k][1] : cen[k][2];
            const uint32_t f = r < count && cv <= d.coord ? 1u : 0u;
            uint32_t tot;
            const uint32_t ex = block_scan<BS>(f, s.w, tot);
            if (r < count) s_rf[r] = (running + ex) | (f << 31);
            running += tot;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0 && running != d.nleft) atomicOr(err, kErrCount);
    const uint32_t nl = d.nleft, nr = count - nl;
    const bool lfin = !active(nl, level + 1), rfin = !active(nr, level + 1);
    for (uint32_t r = threadIdx.x; r < count; r += BS) {
        const uint32_t w = s_rf[r];
        if (w >> 31) {
            const uint32_t dst = w & kRank;
            move_slot(b, in, g, first + r, first + dst, lfin, dst == nl - 1);
        }
        if (r >= nl) {
            uint32_t q = r;
            for (;;) {
                const uint32_t x = s_rf[q];
                if (!(x >> 31) || (x & kRank) >= q) break;
                q = x & kRank;
            }
            move_slot(b, in, g, first + q, first + r, rfin, r == count - 1);
        }
    }
}

// Bins of one chunk of a large node.
__global__ __launch_bounds__(kBlock) void k_mbin(const NodeRec* nodes, int level, const uint32_t* mch_node,
                                                 const uint32_t* mch_k, Slots b, uint32_t* part) {
    __shared__ BinLds s;
    const uint32_t c = blockIdx.x, j = mch_node[c], k = mch_k[c];
    const NodeRec nd = nodes[j];
    const float4* ibox = b.box[level & 1];
    init_bins(s, nd);
    __syncthreads();
    const uint32_t lo = k * kChunk, hi = min(nd.count, lo + kChunk);
    constexpr uint32_t kRun = kChunk / kBlock;
    const uint32_t r0 = lo + threadIdx.x * kRun;
    if (r0 < hi) bin_run(s, ibox, (size_t)nd.first + r0, min(kRun, hi - r0));
    __syncthreads();
    if (threadIdx.x < kBinWords) part[(size_t)c * kBinWords + threadIdx.x] = s.bins[threadIdx.x];
}

__device__ __forceinline__ uint32_t bin_merge(int kind, uint32_t v, uint32_t x) {
    return kind < 3 ? min(v, x) : kind < 6 ? max(v, x) : v + x;
}

// First reduction stage: the block of chunk k (k a multiple of kGroup)
// merges the bins of chunks k .. k + kGroup - 1 of its node into chunk k's.
__global__ __launch_bounds__(kBlock) void k_mreduce(const NodeRec* nodes, const uint32_t* mch_node,
                                                    const uint32_t* mch_k, uint32_t* part) {
    const uint32_t c = blockIdx.x, k = mch_k[c];
    if (k % kGroup != 0) return;
    const uint32_t nch = (nodes[mch_node[c]].count + kChunk - 1) / kChunk;
    const uint32_t end = min(nch, k + kGroup);
    const int t = threadIdx.x;
    if (t >= kBinWords) return;
    const int kind = t % 7;
    uint32_t v = part[(size_t)c * kBinWords + t];
    for (uint32_t q = k + 1; q < end; ++q) v = bin_merge(kind, v, part[(size_t)(c + q - k) * kBinWords + t]);
    part[(size_t)c * kBinWords + t] = v;
}

// Second stage and decision (one block per node, run by the block of its
// first chunk).
__global__ __launch_bounds__(kBlock) void k_mdecide(const NodeRec* nodes, const uint32_t* mch_node,
                                                    const uint32_t* mch_k, const uint32_t* part, NodeRec* kid,
                                                    uint32_t* split, Decision* dec, uint32_t* err) {
    __shared__ BinLds s;
    const uint32_t c = blockIdx.x;
    if (mch_k[c] != 0) return;
    const uint32_t j = mch_node[c];
    const NodeRec nd = nodes[j];
    const uint32_t nch = (nd.count + kChunk - 1) / kChunk;
    init_bins(s, nd);
    const int t = threadIdx.x;
    if (t < kBinWords) {
        const int kind = t % 7;
        uint32_t v = s.bins[t];
        for (uint32_t q = 0; q < nch; q += kGroup) v = bin_merge(kind, v, part[(size_t)(c + q) * kBinWords + t]);
        s.bins[t] = v;
    }
    __syncthreads();
    decide(s, nd, kid + 2 * (size_t)j, split + j, dec + j, err);
}

// Flags and in-chunk ranks of a large node's chunk.
__global__ __launch_bounds__(kBlock) void k_mrank(const NodeRec* nodes, int level, const uint32_t* mch_node,
                                                  const uint32_t* mch_k, Slots b, const Decision* dec, uint32_t* rf,
                                                  uint32_t* chunk_left) {
    __shared__ uint32_t s_w[kBlock / 64];
    const uint32_t c = blockIdx.x, j = mch_node[c], k = mch_k[c];
    const Decision d = dec[j];
    if (!d.split) {
        if (threadIdx.x == 0) chunk_left[c] = 0;
        return;
    }
    const NodeRec nd = nodes[j];
    const float4* ibox = b.box[level & 1];
    const uint32_t lo = k * kChunk, hi = min(nd.count, lo + kChunk);
    uint32_t running = 0;
    for (uint32_t base = lo; base < hi; base += kBlock) {
        const uint32_t r = base + threadIdx.x;
        uint32_t f = 0;
        if (r < hi) {
            const size_t p = nd.first + r;
            f = centroid(ibox[2 * p], ibox[2 * p + 1], d.axis) <= d.coord ? 1u : 0u;
        }
        uint32_t tot;
        const uint32_t ex = block_scan(f, s_w, tot);
        if (r < hi) rf[nd.first + r] = (running + ex) | (f << 31);
        running += tot;
    }
    if (threadIdx.x == 0) chunk_left[c] = running;
}

// Partition moves of a large node's chunk (ranks: chunk prefix + in-chunk rank).
__global__ __launch_bounds__(kBlock) void k_mmove(const NodeRec* nodes, int level, const uint32_t* mch_node,
                                                  const uint32_t* mch_k, Slots b, Geo g, const Decision* dec,
                                                  const uint32_t* rf, const uint32_t* chunk_pre,
                                                  const uint32_t* chunk_left, uint32_t* err, uint32_t* npend,
                                                  uint32_t* pend_dst, uint32_t* pend_q, uint32_t* jflag,
                                                  uint32_t* jmax) {
    const uint32_t c = blockIdx.x, j = mch_node[c], k = mch_k[c];
    const Decision d = dec[j];
    const NodeRec nd = nodes[j];
    const int in = level & 1;
    const uint32_t first = nd.first, count = nd.count;
    const uint32_t lo = k * kChunk, hi = min(count, lo + kChunk);
    if (!d.split) {  // SAH leaf
        for (uint32_t r = lo + threadIdx.x; r < hi; r += kBlock)
            write_record(g, first + r, b.ref[in][first + r], r == count - 1);
        return;
    }
    const uint32_t c0 = c - k, base = chunk_pre[c0];
    if (k == 0 && threadIdx.x == 0) {
        const uint32_t nch = (count + kChunk - 1) / kChunk;
        if (chunk_pre[c0 + nch - 1] - base + chunk_left[c0 + nch - 1] != d.nleft) atomicOr(err, kErrCount);
    }
    const uint32_t nl = d.nleft, nr = count - nl;
    const bool lfin = !active(nl, level + 1), rfin = !active(nr, level + 1);
    for (uint32_t r = lo + threadIdx.x; r < hi; r += kBlock) {
        const uint32_t w = rf[first + r];
        if (w >> 31) {
            const uint32_t dst = chunk_pre[c] - base + (w & kRank);
            move_slot(b, in, g, first + r, first + dst, lfin, dst == nl - 1);
        }
        if (r >= nl) {
            uint32_t q = r;
            bool done = false;
            for (uint32_t steps = 0; steps < kChaseCap; ++steps) {
                const uint32_t x = rf[first + q];
                if (!(x >> 31)) { done = true; break; }
                const uint32_t rk = chunk_pre[c0 + q / kChunk] - base + (x & kRank);
                if (rk >= q) { done = true; break; }
                q = rk;
            }
            if (done) {
                move_slot(b, in, g, first + q, first + r, rfin, r == count - 1);
            } else {  // finished by pointer jumping (k_jump), then k_pending
                if (!jflag[j]) {
                    jflag[j] = 1;  // benign race: every writer stores 1
                    atomicMax(jmax, count);
                }
                const uint32_t i = atomicAdd(npend, 1u);
                pend_dst[i] = (first + r) | (rfin ? 0x80000000u : 0u) | (r == count - 1 ? 0x40000000u : 0u);
                pend_q[i] = first + q;
            }
        }
    }
}

// Pointer jumping over the swap chains of the large split nodes: J[p] is
// the slot p's chain leads to (rank(p) while p holds a left primitive with
// a smaller rank, else p itself).
__global__ __launch_bounds__(kBlock) void k_jinit(const NodeRec* nodes, const uint32_t* mch_node,
                                                  const uint32_t* mch_k, const uint32_t* jflag, const uint32_t* rf,
                                                  const uint32_t* chunk_pre, uint32_t* J) {
    const uint32_t c = blockIdx.x, j = mch_node[c], k = mch_k[c];
    if (!jflag[j]) return;
    const NodeRec nd = nodes[j];
    const uint32_t lo = k * kChunk, hi = min(nd.count, lo + kChunk), base = chunk_pre[c - k];
    for (uint32_t r = lo + threadIdx.x; r < hi; r += kBlock) {
        const uint32_t x = rf[nd.first + r];
        const uint32_t rk = chunk_pre[c] - base + (x & kRank);
        J[nd.first + r] = nd.first + ((x >> 31) && rk < r ? rk : r);
    }
}

// One doubling round over the slots of the nodes with long chains: after
// round m, J[p] is 2^m steps down p's chain (or its end).
__global__ __launch_bounds__(kBlock) void k_jump(const NodeRec* nodes, const uint32_t* mch_node,
                                                 const uint32_t* mch_k, const uint32_t* jflag, const uint32_t* Jin,
                                                 uint32_t* Jout) {
    const uint32_t c = blockIdx.x, j = mch_node[c], k = mch_k[c];
    if (!jflag[j]) return;
    const NodeRec nd = nodes[j];
    const uint32_t lo = k * kChunk, hi = min(nd.count, lo + kChunk);
    for (uint32_t r = lo + threadIdx.x; r < hi; r += kBlock) Jout[nd.first + r] = Jin[Jin[nd.first + r]];
}

__global__ void k_pending(uint32_t n, const uint32_t* pend_dst, const uint32_t* pend_q, const uint32_t* J, Slots b,
                          int level, Geo g) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t d = pend_dst[i];
    move_slot(b, level & 1, g, J[pend_q[i]], d & 0x3FFFFFFFu, (d >> 31) != 0, ((d >> 30) & 1u) != 0);
}

// Children of the split nodes, in level order, appended after the level.
__global__ void k_children(uint32_t n, uint32_t g0, uint32_t next0, const uint32_t* split, const uint32_t* off,
                           const NodeRec* kid, NodeRec* nodes, int32_t* child) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    if (split[j]) {
        const uint32_t o = next0 + 2 * off[j];
        nodes[o] = kid[2 * (size_t)j];
        nodes[o + 1] = kid[2 * (size_t)j + 1];
        child[g0 + j] = (int32_t)o;
    } else {
        child[g0 + j] = -1;
    }
}

// Subtree counts: split nodes (S) and split nodes at even depth (SQ).
__global__ void k_up(uint32_t g0, uint32_t n, int level, const int32_t* child, uint32_t* S, uint32_t* SQ) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint32_t gi = g0 + j;
    const int32_t c = child[gi];
    if (c < 0) {
        S[gi] = 0;
        SQ[gi] = 0;
    } else {
        S[gi] = 1 + S[c] + S[c + 1];
        SQ[gi] = ((level & 1) ? 0u : 1u) + SQ[c] + SQ[c + 1];
    }
}

// Preorder ranks: R among split nodes, Q among split nodes at even depth.
__global__ void k_down(uint32_t g0, uint32_t n, int level, const int32_t* child, const uint32_t* S,
                       const uint32_t* SQ, uint32_t* R, uint32_t* Q) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint32_t gi = g0 + j;
    const int32_t c = child[gi];
    if (c < 0) return;
    R[c] = R[gi] + 1;
    R[c + 1] = R[gi] + 1 + S[c];
    if (level & 1) return;
    uint32_t off = Q[gi] + 1;
    for (int side = 0; side < 2; ++side) {
        const int32_t cc = child[c + side];
        if (cc < 0) continue;
        for (int s2 = 0; s2 < 2; ++s2) {
            const int32_t gc = cc + s2;
            if (child[gc] >= 0) {
                Q[gc] = off;
                off += SQ[gc];
            }
        }
    }
}

__device__ ct_bvh_node out_rec(const NodeRec& r, const int32_t* child, const uint32_t* R, uint32_t gi) {
    ct_bvh_node o;
    for (int k = 0; k < 3; ++k) {
        o.min[k] = r.mn[k];
        o.max[k] = r.mx[k];
    }
    if (child[gi] >= 0) {
        o.left_first = (int32_t)(2 + 2 * R[gi]);
        o.count = -1;
    } else {
        o.left_first = (int32_t)r.first;
        o.count = (int32_t)r.count;
    }
    return o;
}

// Binary nodes: the children pair of the split node of rank r at 2 + 2r.
__global__ void k_emit2(uint32_t g0, uint32_t n, const NodeRec* nodes, const int32_t* child, const uint32_t* R,
                        ct_bvh_node* out) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint32_t gi = g0 + j;
    const int32_t c = child[gi];
    if (gi == 0) {
        out[0] = out_rec(nodes[0], child, R, 0);
        ct_bvh_node pad;
        for (int k = 0; k < 3; ++k) pad.min[k] = pad.max[k] = 0.0f;
        pad.left_first = -1;
        pad.count = -1;
        out[1] = pad;
    }
    if (c < 0) return;
    const uint32_t base = 2 + 2 * R[gi];
    out[base] = out_rec(nodes[c], child, R, (uint32_t)c);
    out[base + 1] = out_rec(nodes[c + 1], child, R, (uint32_t)(c + 1));
}

__device__ void set_slot(ct_mbvh_node& m, int s, const NodeRec& r) {
    m.bminx[s] = r.mn[0];
    m.bminy[s] = r.mn[1];
    m.bminz[s] = r.mn[2];
    m.bmaxx[s] = r.mx[0];
    m.bmaxy[s] = r.mx[1];
    m.bmaxz[s] = r.mx[2];
}

// Quad nodes (GetBVHNodeInfo + MergeNodes): a split node at even depth takes
// its leaf children and the children of its interior children, left first.
__global__ void k_emit4(uint32_t g0, uint32_t n, const NodeRec* nodes, const int32_t* child, const uint32_t* Q,
                        ct_mbvh_node* out) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint32_t gi = g0 + j;
    const int32_t c = child[gi];
    ct_mbvh_node m;
    for (int s = 0; s < 4; ++s) {
        m.bminx[s] = m.bminy[s] = m.bminz[s] = 1e34f;
        m.bmaxx[s] = m.bmaxy[s] = m.bmaxz[s] = -1e34f;
        m.childs[s] = -1;
        m.counts[s] = 0;
    }
    if (c < 0) {
        if (gi != 0) return;
        // leaf root: kept in slot 0 (the host collapse does the same)
        const NodeRec r = nodes[0];
        set_slot(m, 0, r);
        m.childs[0] = (int32_t)r.first;
        m.counts[0] = (int32_t)r.count;
        out[0] = m;
        return;
    }
    int ns = 0;
    for (int side = 0; side < 2; ++side) {
        const int32_t cc = c + side;
        const int32_t gcc = child[cc];
        const int nsub = gcc < 0 ? 1 : 2;
        for (int s2 = 0; s2 < nsub; ++s2) {
            const int32_t x = gcc < 0 ? cc : gcc + s2;
            const NodeRec r = nodes[x];
            set_slot(m, ns, r);
            if (child[x] < 0) {
                m.childs[ns] = (int32_t)r.first;
                m.counts[ns] = (int32_t)r.count;
            } else {
                m.childs[ns] = (int32_t)Q[x];
                m.counts[ns] = -1;
            }
            ++ns;
        }
    }
    out[Q[gi]] = m;
}

// A leaf root (fewer than 4 primitives): records straight from slot order.
__global__ void k_root_leaf(Geo g, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) write_record(g, i, i, i == n - 1);
}

// ------------------------------------------------------------- device scan
constexpr uint32_t kScanTile = 1024;  // 256 threads x 4

__global__ __launch_bounds__(kBlock) void k_scan_sums(const uint32_t* in, uint32_t n, uint32_t* sums) {
    __shared__ uint32_t s_w[kBlock / 64];
    const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * 4;
    uint32_t v = 0;
    for (int k = 0; k < 4; ++k)
        if (base + k < n) v += in[base + k];
    uint32_t tot;
    (void)block_scan(v, s_w, tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// In-place exclusive scan of the tile sums (one block); total to *total.
__global__ __launch_bounds__(kBlock) void k_scan_top(uint32_t* sums, uint32_t nt, uint32_t* total) {
    __shared__ uint32_t s_w[kBlock / 64];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nt; base += kBlock) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < nt ? sums[i] : 0u;
        uint32_t tot;
        const uint32_t ex = block_scan(v, s_w, tot);
        if (i < nt) sums[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(kBlock) void k_scan_out(const uint32_t* in, uint32_t n, const uint32_t* sums,
                                                     uint32_t* out) {
    __shared__ uint32_t s_w[kBlock / 64];
    const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * 4;
    uint32_t v[4], t = 0;
    for (int k = 0; k < 4; ++k) {
        v[k] = base + k < n ? in[base + k] : 0u;
        t += v[k];
    }
    uint32_t tot;
    uint32_t run = sums[blockIdx.x] + block_scan(t, s_w, tot);
    for (int k = 0; k < 4; ++k)
        if (base + k < n) {
            out[base + k] = run;
            run += v[k];
        }
}

inline uint32_t blocks_for(uint64_t n, uint32_t per) { return (uint32_t)std::max<uint64_t>(1, (n + per - 1) / per); }

// Device scratch of one build, released by the destructor.
struct Scratch {
    ct_ctx* c;
    std::vector<void*> bufs;
    explicit Scratch(ct_ctx* cc) : c(cc) {}
    ~Scratch() {
        for (void* p : bufs) ctr::dfree(c, p);
    }
    template <class T>
    int get(T** p, size_t count, const char* what) {
        void* v = nullptr;
        const int r = ctr::dalloc(c, &v, std::max<size_t>(count, 1) * sizeof(T), what);
        if (r != CT_OK) return r;
        bufs.push_back(v);
        *p = static_cast<T*>(v);
        return CT_OK;
    }
};

}  // namespace

namespace ctr {

#define BCHK(expr, what)                                       \
    do {                                                       \
        hipError_t e_ = (expr);                                \
        if (e_ != hipSuccess) return ctr::hip_fail(c, e_, what); \
    } while (0)
#define BLAUNCH(what) BCHK(hipGetLastError(), what)

int device_build(ct_ctx* c, const ct_scene_view* v, DeviceTrees* out, ct_build_stats* st) {
    const auto t0 = std::chrono::steady_clock::now();
    const uint32_t N = v->n_triangles, V = v->n_vertices;
    hipStream_t s = c->ln[0].stream;
    Scratch sc(c);
    const uint64_t W = std::max<uint64_t>(4, (uint64_t)N / 2 + 2);  // nodes of one level
    const uint64_t Nn = 2 * (uint64_t)N + 2;                        // nodes of all levels
    const uint64_t Mc = 2 * (uint64_t)N / kChunk + 2;               // chunks of large nodes
    float* d_vert;
    uint32_t *d_idx, *ref0, *ref1, *rf, *nt, *ns, *nm, *off_t, *off_s, *off_m, *split, *coff, *tch, *sch, *mch_node,
        *mch_k, *part, *J0, *J1, *pend_dst, *pend_q, *jflag,
        *chunk_left, *chunk_pre, *S, *SQ, *R, *Q, *sums, *cnt;
    float4 *box0, *box1;
    NodeRec *nodes, *kid;
    int32_t* child;
    Decision* dec;
    float* rpart;
    const uint32_t nblk_init = blocks_for(N, kBlock);
    int r = CT_OK;
#define GET(p, n, w) \
    if (r == CT_OK) r = sc.get(&p, n, w)
    GET(d_vert, 3 * (uint64_t)V, "build: vertices");
    GET(d_idx, 3 * (uint64_t)N, "build: indices");
    GET(box0, 2 * (uint64_t)N, "build: slot boxes");
    GET(box1, 2 * (uint64_t)N, "build: slot boxes");
    GET(ref0, N, "build: slot ids");
    GET(ref1, N, "build: slot ids");
    GET(rf, N, "build: ranks");
    GET(J0, N, "build: chain jumps");
    GET(J1, N, "build: chain jumps");
    GET(pend_dst, N, "build: long chains");
    GET(pend_q, N, "build: long chains");
    GET(jflag, W, "build: long chains");
    GET(nodes, Nn, "build: nodes");
    GET(child, Nn, "build: children");
    GET(S, Nn, "build: subtree counts");
    GET(SQ, Nn, "build: subtree counts");
    GET(R, Nn, "build: ranks");
    GET(Q, Nn, "build: ranks");
    GET(nt, W, "build: level");
    GET(off_t, W, "build: level");
    GET(tch, W, "build: level");
    GET(ns, W, "build: level");
    GET(nm, W, "build: level");
    GET(off_s, W, "build: level");
    GET(off_m, W, "build: level");
    GET(split, W, "build: level");
    GET(coff, W, "build: level");
    GET(sch, W, "build: level");
    GET(kid, 2 * W, "build: level");
    GET(dec, W, "build: level");
    GET(mch_node, Mc, "build: chunks");
    GET(mch_k, Mc, "build: chunks");
    GET(chunk_left, Mc, "build: chunks");
    GET(chunk_pre, Mc, "build: chunks");
    GET(part, Mc * kBinWords, "build: chunk bins");
    GET(sums, blocks_for(Nn, kScanTile) + 1, "build: scan");
    GET(cnt, 16, "build: counters");
    GET(rpart, (uint64_t)nblk_init * 6, "build: root bounds");
#undef GET
    if (r != CT_OK) return r;
    float4* tris = nullptr;
    {
        void* p = nullptr;
        r = dalloc(c, &p, 3 * (size_t)N * sizeof(float4), "triangle records");
        if (r != CT_OK) return r;
        tris = static_cast<float4*>(p);
        out->tris = tris;
    }
    uint32_t* h = nullptr;  // pinned read-back
    BCHK(hipHostMalloc((void**)&h, 64, hipHostMallocDefault), "build: pinned read-back");
    struct HostFree {
        uint32_t* p;
        ~HostFree() { (void)hipHostFree(p); }
    } hf{h};

    hipEvent_t e0, e1;
    BCHK(hipEventCreate(&e0), "build: event");
    BCHK(hipEventCreate(&e1), "build: event");
    struct EvFree {
        hipEvent_t a, b;
        ~EvFree() {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } ef{e0, e1};

    BCHK(hipMemcpyAsync(d_vert, v->vertices, 12 * (size_t)V, hipMemcpyHostToDevice, s), "build: vertices");
    BCHK(hipMemcpyAsync(d_idx, v->indices, 12 * (size_t)N, hipMemcpyHostToDevice, s), "build: indices");
    BCHK(hipMemsetAsync(cnt, 0, 16 * sizeof(uint32_t), s), "build: counters");
    BCHK(hipEventRecord(e0, s), "build: event");

    const Slots slots{{box0, box1}, {ref0, ref1}};
    const Geo geo{d_vert, d_idx, tris};
    k_init<<<nblk_init, kBlock, 0, s>>>(geo, slots, N, rpart);
    BLAUNCH("build: init");
    k_root<<<1, kBlock, 0, s>>>(rpart, nblk_init, N, nodes);
    BLAUNCH("build: root");

    // exclusive scan of n words into out; total into cnt[slot]
    auto scan = [&](const uint32_t* in, uint32_t* outp, uint32_t n, int slot) -> int {
        const uint32_t nt = blocks_for(n, kScanTile);
        k_scan_sums<<<nt, kBlock, 0, s>>>(in, n, sums);
        k_scan_top<<<1, kBlock, 0, s>>>(sums, nt, cnt + slot);
        k_scan_out<<<nt, kBlock, 0, s>>>(in, n, sums, outp);
        BLAUNCH("build: scan");
        return CT_OK;
    };
    auto readback = [&](int words) -> int {
        BCHK(hipMemcpyAsync(h, cnt, (size_t)words * 4, hipMemcpyDeviceToHost, s), "build: read-back");
        BCHK(hipStreamSynchronize(s), "build: level");
        return CT_OK;
    };

    std::vector<std::pair<uint32_t, uint32_t>> levels;  // [g0, n) per level
    uint32_t long_chains = 0, jump_rounds = 0;
    uint32_t g0 = 0, n = 1;
    const uint32_t tpb = 256;
    for (int level = 0; n > 0; ++level) {
        levels.emplace_back(g0, n);
        const NodeRec* lv = nodes + g0;
        k_plan<<<blocks_for(n, tpb), tpb, 0, s>>>(lv, n, level, nt, ns, nm, split);
        BLAUNCH("build: plan");
        if ((r = scan(ns, off_s, n, 0)) || (r = scan(nm, off_m, n, 1)) || (r = scan(nt, off_t, n, 6)) ||
            (r = readback(7)))
            return r;
        const uint32_t Ts = h[0], Tm = h[1], Tt = h[6];
        if (Ts > W || Tt > W || Tm > Mc) return fail(c, CT_ERR_DEVICE, "build: chunk list overflow");
        k_lists<<<blocks_for(n, tpb), tpb, 0, s>>>(n, nt, off_t, ns, off_s, tch, sch);
        BLAUNCH("build: node lists");
        if (Tt) {
            k_single<64, kTiny><<<Tt, 64, 0, s>>>(lv, level, tch, slots, geo, kid, split, cnt + 4);
            BLAUNCH("build: tiny nodes");
        }
        if (Ts) {
            k_single<kBlock, kChunk><<<Ts, kBlock, 0, s>>>(lv, level, sch, slots, geo, kid, split, cnt + 4);
            BLAUNCH("build: small nodes");
        }
        if (Tm) {
            k_mchunks<<<blocks_for(Tm, tpb), tpb, 0, s>>>(Tm, n, off_m, mch_node, mch_k);
            k_mbin<<<Tm, kBlock, 0, s>>>(lv, level, mch_node, mch_k, slots, part);
            k_mreduce<<<Tm, kBlock, 0, s>>>(lv, mch_node, mch_k, part);
            k_mdecide<<<Tm, kBlock, 0, s>>>(lv, mch_node, mch_k, part, kid, split, dec, cnt + 4);
            k_mrank<<<Tm, kBlock, 0, s>>>(lv, level, mch_node, mch_k, slots, dec, rf, chunk_left);
            BLAUNCH("build: large nodes");
            if ((r = scan(chunk_left, chunk_pre, Tm, 3))) return r;
            BCHK(hipMemsetAsync(cnt + 7, 0, 8, s), "build: counters");
            BCHK(hipMemsetAsync(jflag, 0, (size_t)n * 4, s), "build: counters");
            k_mmove<<<Tm, kBlock, 0, s>>>(lv, level, mch_node, mch_k, slots, geo, dec, rf, chunk_pre, chunk_left,
                                          cnt + 4, cnt + 7, pend_dst, pend_q, jflag, cnt + 8);
            BLAUNCH("build: large nodes");
        }
        if ((r = scan(split, coff, n, 2)) || (r = readback(9))) return r;
        const uint32_t npend = Tm ? h[7] : 0;
        if (npend) {
            // swap chains longer than the in-kernel chase: pointer jumping to
            // their ends in the nodes that have them; a chain is shorter than
            // its node, so ceil(log2(largest such node)) rounds reach every end
            long_chains += npend;
            k_jinit<<<Tm, kBlock, 0, s>>>(lv, mch_node, mch_k, jflag, rf, chunk_pre, J0);
            BLAUNCH("build: chain jumps");
            uint32_t* Jc = J0;
            uint32_t* Jn = J1;
            for (uint32_t span = 1; span < h[8]; span *= 2) {
                k_jump<<<Tm, kBlock, 0, s>>>(lv, mch_node, mch_k, jflag, Jc, Jn);
                BLAUNCH("build: chain jumps");
                std::swap(Jc, Jn);
                ++jump_rounds;
            }
            k_pending<<<blocks_for(npend, tpb), tpb, 0, s>>>(npend, pend_dst, pend_q, Jc, slots, level, geo);
            BLAUNCH("build: long chains");
        }
        if (h[4]) return fail(c, CT_ERR_INVALID_SCENE, "device build: scene extent overflows the SAH cost");
        const uint32_t nsplit = h[2];
        const uint32_t next0 = g0 + n;
        if ((uint64_t)next0 + 2 * (uint64_t)nsplit > Nn || 2 * (uint64_t)nsplit > W)
            return fail(c, CT_ERR_DEVICE, "build: node list overflow");
        k_children<<<blocks_for(n, tpb), tpb, 0, s>>>(n, g0, next0, split, coff, kid, nodes, child);
        BLAUNCH("build: children");
        if (level == 0 && nsplit == 0) {
            k_root_leaf<<<blocks_for(N, tpb), tpb, 0, s>>>(geo, N);
            BLAUNCH("build: leaf root");
        }
        g0 = next0;
        n = 2 * nsplit;
        if (level + 1 >= kMaxDepth) return fail(c, CT_ERR_DEVICE, "build: level limit");
    }
    const uint32_t n_levels = (uint32_t)levels.size();
    // bottom-up subtree counts, top-down ranks
    for (uint32_t L = n_levels; L-- > 0;) {
        k_up<<<blocks_for(levels[L].second, tpb), tpb, 0, s>>>(levels[L].first, levels[L].second, (int)L, child, S, SQ);
        BLAUNCH("build: subtree counts");
    }
    BCHK(hipMemsetAsync(R, 0, 4, s), "build: ranks");
    BCHK(hipMemsetAsync(Q, 0, 4, s), "build: ranks");
    for (uint32_t L = 0; L < n_levels; ++L) {
        k_down<<<blocks_for(levels[L].second, tpb), tpb, 0, s>>>(levels[L].first, levels[L].second, (int)L, child, S,
                                                                  SQ, R, Q);
        BLAUNCH("build: ranks");
    }
    BCHK(hipMemcpyAsync(cnt + 8, S, 4, hipMemcpyDeviceToDevice, s), "build: sizes");
    BCHK(hipMemcpyAsync(cnt + 9, SQ, 4, hipMemcpyDeviceToDevice, s), "build: sizes");
    if ((r = readback(10))) return r;
    if (h[4] & kErrCount) return fail(c, CT_ERR_DEVICE, "device build: partition count mismatch");
    const uint32_t n2 = 2 + 2 * h[8];
    const uint32_t n4 = h[8] ? h[9] : 1;
    void *p2 = nullptr, *p4 = nullptr;
    if ((r = dalloc(c, &p2, (size_t)n2 * sizeof(ct_bvh_node), "binary nodes"))) return r;
    out->nodes2 = static_cast<ct_bvh_node*>(p2);
    if ((r = dalloc(c, &p4, (size_t)n4 * sizeof(ct_mbvh_node), "quad nodes"))) return r;
    out->nodes4 = static_cast<ct_mbvh_node*>(p4);
    for (uint32_t L = 0; L < n_levels; ++L) {
        k_emit2<<<blocks_for(levels[L].second, tpb), tpb, 0, s>>>(levels[L].first, levels[L].second, nodes, child, R,
                                                                   out->nodes2);
        if (!(L & 1))
            k_emit4<<<blocks_for(levels[L].second, tpb), tpb, 0, s>>>(levels[L].first, levels[L].second, nodes, child,
                                                                       Q, out->nodes4);
        BLAUNCH("build: emit");
    }
    BCHK(hipEventRecord(e1, s), "build: event");
    BCHK(hipStreamSynchronize(s), "build: finish");
    float ms = 0.0f;
    BCHK(hipEventElapsedTime(&ms, e0, e1), "build: event");
    out->n2 = n2;
    out->n4 = n4;
    if (st) {
        st->n_bvh_nodes = n2;
        st->n_mbvh_nodes = n4;
        st->levels = n_levels;
        st->max_depth = n_levels + 1;  // the reference's m_MaxDepth: root depth 2
        st->long_chains = long_chains;
        st->jump_rounds = jump_rounds;
        st->device_ms = ms;
        st->total_ms =
            (float)std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    return CT_OK;
}

}  // namespace ctr
