// Device-side types of the wavefront loop (gfx950).
//
// All kernels are compiled with -ffp-contract=off and correctly rounded
// fp32 division/sqrt, and every vector expression below keeps glm's
// evaluation order (dot = (x*x' + y*y') + z*z', normalize = v * (1/sqrt(dot)),
// min(x, y) = (y < x) ? y : x), so the FP32 results equal the CPU oracle's.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../../include/ct_fp.h"

#define CT_DEV __device__ __forceinline__

namespace ctd {

struct f3 {
    float x, y, z;
};
CT_DEV f3 mk3(float a, float b, float c) { return f3{a, b, c}; }
CT_DEV f3 xyz(float4 v) { return f3{v.x, v.y, v.z}; }
CT_DEV f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
CT_DEV f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
CT_DEV f3 operator*(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
CT_DEV f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
CT_DEV f3 operator*(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
CT_DEV f3 operator/(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
CT_DEV f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
CT_DEV float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
CT_DEV f3 cross(f3 x, f3 y) { return f3{x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y}; }
CT_DEV f3 normalize(f3 v) { return v * (1.0f / ct_sqrtf(dot(v, v))); }
CT_DEV float length(f3 v) { return ct_sqrtf(dot(v, v)); }
CT_DEV float gmin(float x, float y) { return (y < x) ? y : x; }
CT_DEV float gmax(float x, float y) { return (x < y) ? y : x; }
CT_DEV f3 vmax0(f3 v) { return f3{gmax(0.0f, v.x), gmax(0.0f, v.y), gmax(0.0f, v.z)}; }
CT_DEV bool isnanf_(float x) { return x != x; }

// Path flags (two high bits of the path word).
enum : uint32_t { LT_SPECULAR = 0u, LT_DIFFUSE = 1u, LT_GLOSSY = 2u };
constexpr uint32_t kSlotMask = 0x3fffffffu;

// Leaf-ordered triangle record: three float4 per triangle, in prim_indices
// order so a leaf is a contiguous run.  .w of the first holds the primitive
// id with bit 31 set on the last triangle of a leaf.
constexpr uint32_t kLastInLeaf = 0x80000000u;

// Stack entries: interior node index, or kLeafBit | first triangle record.
constexpr uint32_t kLeafBit = 0x80000000u;
constexpr int kMaxStack = 128;  // >= 3*32+1 (quad tree of a depth-64 binary tree)

struct DevScene {
    const float4* tris;     // 3 per triangle (leaf order): {p0, id}, {p1-p0}, {p2-p0}
    const float4* nodes4;   // 8 per quad node: bminx,bmaxx,bminy,bmaxy,bminz,bmaxz,childs,counts
    const float4* nodes4f;  // the same nodes for the fast node test: six plane rows, four stack entries, zeros
    const float4* nodes2;   // 2 per binary node: {min, leftFirst}, {max, count}
    const uint4* fnodes;    // 8 per wide compressed node (128 B), layout in host/cnode.h
    const uint4* cnodes;    // 4 per quad node in the compact layout (64 B, cnode4_step in device/traverse.h)
    const float4* sh_prim;  // kPrimRows per primitive, one 128-byte line (layout below)
    const float4* mats;     // 2 per material: {albedo, roughness}, {type bits, ior, alpha, 0}
    const float4* lights;   // 5 per light: v0, v1, v2, centre normal (w = area), emission
    const float4* texels;     // float textures' RGBA texels, concatenated
    const uint32_t* texels8;  // 8-bit textures' RGB texels (byte 3 unused), concatenated
    const uint4* tex_info;    // per texture: first texel (in its buffer), width, height, format (0 float, 1 8-bit)
    uint32_t n_tris, n_nodes4, n_nodes2, n_fnodes, n_mats, n_lights, n_tex;
    int sky_enabled;
    int sky_tex;            // -1: the constant sky colour
    float sky_r, sky_g, sky_b;
    uint32_t plane_flags;   // kPlaneNegZero | kPlaneNaN | kEmptyEntry | kCompactBad over the quad tree
};

// A primitive's shading record (sh_prim), indexed by primitive id: what
// shade reads for a hit, in one 128-byte line (one miss per hit instead of
// the three or four lines of separate position, normal and texture
// coordinate arrays):
//   0 {v0, material index bits}  1 {v1, u0}  2 {v2, v0 tex}
//   3 {n0, u1}  4 {n1, v1 tex}  5 {n2, u2}  6 {centre normal, v2 tex}  7 zero
// (texture coordinates are zero without textures).
constexpr int kPrimRows = 8;

// Properties of the quad tree's box planes, found at upload (k_plane_flags).
constexpr uint32_t kPlaneNegZero = 1u;  // some plane is -0.0
constexpr uint32_t kPlaneNaN = 2u;      // some plane is NaN
constexpr uint32_t kEmptyEntry = 4u;    // some empty slot (count 0) has a child other than 0xFFFFFFFF
constexpr uint32_t kCompactBad = 8u;    // some node's boxes have no compact encoding (not finite, or too large)

// One camera view (plain frame, batch member or served request).
struct ViewDesc {
    float pos[3], w[3], hor[3], ver[3];
    float invw, invh;
    int32_t width, height, spp;
    uint32_t frame_base;
    uint64_t accum_offset;  // float4 index of the view's pixel 0 in the accumulator arena
    uint32_t strip_mul;     // stride of the full pixel strips' order (pixel_of), coprime to their count
    uint32_t pad;
};

// A run of consecutive path slots of one view inside a chunk.
struct Segment {
    uint32_t chunk_offset;  // first slot of the segment in the chunk
    uint32_t view;
    uint32_t first_path;    // path index within the view (pixel * spp + sample)
    uint32_t n;             // paths
    uint32_t pix_offset;    // first resolve task of the segment
    uint32_t pad[3];
};

// Pixel of a view's pixel ordinal (path index / spp).  Ordinals run down
// strips of kTileRows rows column by column, so the 64 primary rays of a
// wave cover a 32 x 2 tile and share more of the tree than a run of one
// row.  On the 7.24M-triangle 4K scene 2 rows ran 1.5% ahead of 4 (which had
// been ahead of 1, 8 and 16); the 259K-triangle atrium is indifferent.
// The n full strips are taken in the order strip = k * mul mod n (mul
// coprime to n, about 0.618 n), so any run of ordinals, such as one
// wavefront chunk, covers strips spread over the whole image: a frame's
// chunks get similar work (with contiguous halves the lower half of the 4K
// scene took its lane twice as long as the upper half's).  The last strip
// has h % kTileRows rows when h is not a multiple, and comes last.  A
// bijection of [0, w * h); seeds and the accumulator use the pixel itself,
// so the order changes no result.
#ifndef CT_TILE_ROWS
#define CT_TILE_ROWS 2
#endif
constexpr uint32_t kTileRows = CT_TILE_ROWS;
CT_DEV uint32_t pixel_of(uint32_t ord, uint32_t w, uint32_t h, uint32_t mul) {
    const uint32_t k = ord / (kTileRows * w);
    const uint32_t r = ord - k * (kTileRows * w);
    const uint32_t nfull = h / kTileRows;
    const uint32_t strip = k < nfull ? (uint32_t)(((uint64_t)k * mul) % nfull) : k;
    const uint32_t y0 = strip * kTileRows;
    const uint32_t rows = min(kTileRows, h - y0);
    const uint32_t x = r / rows;
    return (y0 + (r - x * rows)) * w + x;
}

// Per-chunk sizes, written by the host before the chunk's kernels run (so a
// captured graph of the bounce loop replays unchanged for every chunk).
struct ChunkHdr {
    uint32_t n;        // path slots in the chunk
    uint32_t n_tasks;  // resolve tasks (pixels touched)
    uint32_t nseg;     // segments
    uint32_t gen;      // the chunk's generation (>= 1) in the stamps of the radiance terms
};

struct TraceOpts {
    int accel, traversal;
    float teps;
};

struct ShadeOpts {
    int max_depth, rr_start, bsdf_mode, shadow, indirect;
    float neps, deps, teps;
    int light_pdf;
    int sort;  // block-local material sort before shading
};

// Light-selection factor of the MIS light pdf (x count like Kernel.cu:128,184,
// or / count for uniform selection).
CT_DEV float light_pdf_scale(float pdf, uint32_t n, int mode) { return mode ? pdf / (float)n : pdf * (float)n; }

// Queue buffers of one context (structure of arrays).
struct Queues {
    float4* ray_o[2];   // origin.xyz
    float4* ray_d[2];   // direction.xyz
    float4* thr[2];     // throughput.xyz, lastBsdfPdf
    uint32_t* path[2];  // slot | lastBounceType << 30
    float2* hit;        // t, primitive id bits
    // shadow queue of bounce b at [b % 2]: connect(b) still reads it while
    // shade(b + 1) writes the next one
    float4* sh_o[2];    // shadow origin.xyz, tmax
    float4* sh_d[2];    // shadow direction.xyz, slot bits
    float4* sh_c[2];    // shadow colour.xyz
    float4* lpath;      // per-slot radiance (xyz), .w = bounces whose terms are folded in (bits)
    float4* ebuf[2];    // shade's radiance term of bounce b at [b % 2] (xyz), .w = stamp (bits)
    float4* nbuf[2];    // connect's radiance term of bounce b at [b % 2] (xyz), .w = stamp (bits)
    uint32_t* counters; // see CounterIndex
    uint32_t* spill;    // traversal stack overflow area: per block (max_stack - LDS depth) x blockDim words
    uint32_t* err;      // error flags
    int max_stack = kMaxStack;  // traversal stack entries (LDS + spill); deeper trees report an overflow
};

// counters[] layout: for bounce b in [0, kMaxBounce]
constexpr int kMaxBounce = 64;

// Radiance terms.  A path's radiance is the sum, in this order, of shade's
// term (emission or sky) and connect's term (unoccluded light sample) of
// bounce 0, 1, 2, ...; float addition is not associative, so the order is
// the reference's.  The terms of bounce b are written to ebuf / nbuf[b % 2]
// stamped (gen << 6) | b; shade(b + 2) folds bounce b's pair into lpath and
// k_resolve folds the pairs still open, so shade(b + 1) need not wait for
// connect(b).
CT_DEV uint32_t term_stamp(uint32_t gen, int b) { return (gen << 6) | (uint32_t)b; }
CT_DEV float4 term(f3 v, uint32_t stamp) { return make_float4(v.x, v.y, v.z, __uint_as_float(stamp)); }
CT_DEV void add_term(float4& L, float4 t) {
    L.x = L.x + t.x;
    L.y = L.y + t.y;
    L.z = L.z + t.z;
}
CT_DEV uint32_t* cnt_active(uint32_t* c, int b) { return c + b; }
CT_DEV uint32_t* cnt_shadow(uint32_t* c, int b) { return c + (kMaxBounce + 1) + b; }
// Extension and shadow rays traced by k_tail (a chunk's roulette tail), over
// all the bounces it runs.
CT_DEV uint32_t* cnt_tail_ext(uint32_t* c) { return c + 2 * (kMaxBounce + 1); }
CT_DEV uint32_t* cnt_tail_shadow(uint32_t* c) { return c + 2 * (kMaxBounce + 1) + 1; }
// Work-fetch counters, [queue][range][bounce]: each queue range owns a
// 384-byte slot so ranges never share a cache line (same-line atomics from
// different XCDs serialise).  Four ranges per XCD.
constexpr int kXcds = 8;
constexpr int kFetchRanges = 4 * kXcds;
constexpr int kFetchSlot = 96;
constexpr int kFetchBase = 192;
static_assert(2 * (kMaxBounce + 1) + 2 <= kFetchBase, "counters overlap");
CT_DEV uint32_t* cnt_fetch_ext(uint32_t* c, int b) { return c + kFetchBase + b; }
CT_DEV uint32_t* cnt_fetch_sh(uint32_t* c, int b) { return c + kFetchBase + kFetchRanges * kFetchSlot + b; }
constexpr int kCounterWords = kFetchBase + 2 * kFetchRanges * kFetchSlot + 64;

}  // namespace ctd
