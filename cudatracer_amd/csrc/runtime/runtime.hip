// Context, scene upload and the wavefront frame driver behind the C ABI.
//
// Reference host side replaced here:
//   Params / Params::reset / deallocate     src/CUDA/Kernel.cuh:23-72
//   sample() bounce loop                    src/CUDA/Kernel.cu:272-353
//   launchKernels                           src/CUDA/Kernel.cu:355-372
//   Application::allocateBuffers/Triangles  src/Application.cpp:141-193
//   Application::updateMaterials            src/Application.cpp:222-225
#include "runtime.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <sstream>

using ctd::ChunkHdr;
using ctd::Segment;
using ctd::ViewDesc;

namespace ctr {

int fail(ct_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

int hip_fail(ct_ctx* c, hipError_t e, const char* what) {
    std::ostringstream os;
    os << what << ": " << hipGetErrorString(e);
    return fail(c, CT_ERR_DEVICE, os.str());
}

int dalloc(ct_ctx* c, void** p, size_t bytes, const char* what) {
    *p = nullptr;
    if (bytes == 0) bytes = 16;
    if (c->used + bytes > c->budget) {
        std::ostringstream os;
        os << "device budget exceeded allocating " << what << " (" << bytes << " bytes, " << c->used << " of "
           << c->budget << " in use)";
        return fail(c, CT_ERR_OUT_OF_MEMORY, os.str());
    }
    const hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        std::ostringstream os;
        os << "hipMalloc " << what << " (" << bytes << " bytes): " << hipGetErrorString(e);
        return fail(c, CT_ERR_OUT_OF_MEMORY, os.str());
    }
    c->used += bytes;
    c->allocs.emplace_back(*p, bytes);
    return CT_OK;
}

void dfree(ct_ctx* c, void* p) {
    if (!p) return;
    for (size_t i = 0; i < c->allocs.size(); ++i)
        if (c->allocs[i].first == p) {
            c->used -= c->allocs[i].second;
            c->allocs.erase(c->allocs.begin() + (long)i);
            break;
        }
    (void)hipFree(p);
}

static void drop_lane_graph(Lane& L) {
    if (L.graph_exec) (void)hipGraphExecDestroy(L.graph_exec);
    L.graph_exec = nullptr;
    L.graph_key.clear();
}

// Callers have all lanes idle.
void drop_graphs(ct_ctx* c) {
    for (Lane& L : c->ln) drop_lane_graph(L);
}

void sync_lanes(ct_ctx* c) {
    for (Lane& L : c->ln) {
        if (L.stream) (void)hipStreamSynchronize(L.stream);
        L.pending = false;
    }
}

int check_opts(const ct_ctx* c, const ct_frame_opts& o, const char** why) {
    auto bad = [&](int code, const char* msg) {
        *why = msg;
        return code;
    };
    if (o.max_depth < 1 || o.max_depth > ctd::kMaxBounce) return bad(CT_ERR_INVALID_ARG, "max_depth must be in [1, 64]");
    if (o.spp < 1 || o.spp > (1 << 20)) return bad(CT_ERR_INVALID_ARG, "spp must be in [1, 2^20]");
    if (o.accel != CT_ACCEL_BVH2 && o.accel != CT_ACCEL_BVH4) return bad(CT_ERR_INVALID_ARG, "bad accel");
    if (o.bsdf_mode != CT_BSDF_REFERENCE && o.bsdf_mode != CT_BSDF_TYPED) return bad(CT_ERR_INVALID_ARG, "bad bsdf_mode");
    if (o.traversal != CT_TRAV_EXACT && o.traversal != CT_TRAV_FAST && o.traversal != CT_TRAV_COMPACT)
        return bad(CT_ERR_INVALID_ARG, "bad traversal");
    if (o.light_pdf != CT_LIGHTPDF_REFERENCE && o.light_pdf != CT_LIGHTPDF_UNIFORM)
        return bad(CT_ERR_INVALID_ARG, "bad light_pdf");
    if (!(o.triangle_epsilon >= 0.0f) || !(o.normal_epsilon >= 0.0f) || !(o.dist_epsilon >= 0.0f))
        return bad(CT_ERR_INVALID_ARG, "epsilons must be finite and >= 0");
    if (!c->has_scene) return bad(CT_ERR_NO_SCENE, "no scene uploaded");
    if (o.accel == CT_ACCEL_BVH2 && !c->has_bvh2) return bad(CT_ERR_INVALID_ARG, "scene has no binary tree");
    if (o.accel == CT_ACCEL_BVH4 && !c->has_bvh4) return bad(CT_ERR_INVALID_ARG, "scene has no quad tree");
    if (o.traversal == CT_TRAV_FAST && (c->sc.n_fnodes == 0 || c->wide_stack_bound > (uint32_t)ctd::kMaxStack))
        return bad(CT_ERR_INVALID_ARG, "fast traversal not available for this scene (wide tree too deep)");
    if (o.traversal == CT_TRAV_COMPACT && o.accel != CT_ACCEL_BVH4)
        return bad(CT_ERR_INVALID_ARG, "compact traversal runs on the quad tree (accel CT_ACCEL_BVH4)");
    if (o.traversal == CT_TRAV_COMPACT && (!c->sc.cnodes || (c->sc.plane_flags & ctd::kCompactBad)))
        return bad(CT_ERR_INVALID_ARG, "compact traversal not available for this scene (a box has no 8-bit encoding)");
    return CT_OK;
}

int validate_opts(ct_ctx* c, const ct_frame_opts& o) {
    const char* why = "";
    const int r = check_opts(c, o, &why);
    return r == CT_OK ? CT_OK : fail(c, r, why);
}

bool camera_ok(const ct_camera* cam) {
    if (!cam) return false;
    auto fin3 = [](const float* p) { return std::isfinite(p[0]) && std::isfinite(p[1]) && std::isfinite(p[2]); };
    if (!fin3(cam->position) || !fin3(cam->forward) || !fin3(cam->up) || !std::isfinite(cam->plane_distance))
        return false;
    const float l2 = cam->forward[0] * cam->forward[0] + cam->forward[1] * cam->forward[1] +
                     cam->forward[2] * cam->forward[2];
    return l2 > 0.0f && cam->plane_distance > 0.0f;
}

static void free_paths(ct_ctx* c) {
    sync_lanes(c);
    for (void* p : c->path_bufs) dfree(c, p);
    c->path_bufs.clear();
    for (Lane& L : c->ln) {
        if (L.h_read) (void)hipHostFree(L.h_read);
        L.h_read = nullptr;
        L.q = ctd::Queues{};
        L.wstats = nullptr;
        L.d_hdr = nullptr;
    }
    c->lane_cap = 0;
    drop_graphs(c);
}

int ensure_paths(ct_ctx* c, uint32_t cap) {
    cap = std::max<uint32_t>(cap, 64u);
    const uint32_t threads = (uint32_t)(c->shape.trace_blocks * c->shape.block);
    if (c->lane_cap >= cap && c->spill_threads >= threads && c->ln[0].q.counters) return CT_OK;
    cap = std::max(cap, c->lane_cap);
    free_paths(c);
    auto A = [&](void** p, size_t bytes, const char* w) -> int {
        const int r = dalloc(c, p, bytes, w);
        if (r == CT_OK) c->path_bufs.push_back(*p);
        return r;
    };
    // path buffers shared by the lanes (lane k at offset k * cap)
    const size_t slots = kLanes * (size_t)cap;
    const size_t f4 = slots * sizeof(float4);
    ctd::Queues q{};
    int r = CT_OK;
    for (int k = 0; k < 2 && r == CT_OK; ++k) {
        r = A((void**)&q.ray_o[k], f4, "ray origins");
        if (r == CT_OK) r = A((void**)&q.ray_d[k], f4, "ray directions");
        if (r == CT_OK) r = A((void**)&q.thr[k], f4, "throughput");
        if (r == CT_OK) r = A((void**)&q.path[k], slots * 4, "path words");
    }
    if (r == CT_OK) r = A((void**)&q.hit, slots * sizeof(float2), "hits");
    for (int k = 0; k < 2 && r == CT_OK; ++k) {
        r = A((void**)&q.sh_o[k], f4, "shadow origins");
        if (r == CT_OK) r = A((void**)&q.sh_d[k], f4, "shadow directions");
        if (r == CT_OK) r = A((void**)&q.sh_c[k], f4, "shadow colours");
    }
    if (r == CT_OK) r = A((void**)&q.lpath, f4, "path radiance");
    for (int k = 0; k < 2 && r == CT_OK; ++k) {
        r = A((void**)&q.ebuf[k], f4, "radiance terms");
        if (r == CT_OK) r = A((void**)&q.nbuf[k], f4, "radiance terms");
        // stamp 0 is never a chunk's (generations start at 1)
        if (r == CT_OK && (hipMemsetAsync(q.ebuf[k], 0, f4, c->ln[0].stream) != hipSuccess ||
                           hipMemsetAsync(q.nbuf[k], 0, f4, c->ln[0].stream) != hipSuccess))
            r = fail(c, CT_ERR_DEVICE, "clear radiance terms");
    }
    for (int l = 0; l < kLanes && r == CT_OK; ++l) {
        Lane& L = c->ln[l];
        const size_t o = (size_t)l * cap;
        L.q = q;
        for (int k = 0; k < 2; ++k) {
            L.q.ray_o[k] = q.ray_o[k] + o;
            L.q.ray_d[k] = q.ray_d[k] + o;
            L.q.thr[k] = q.thr[k] + o;
            L.q.path[k] = q.path[k] + o;
        }
        L.q.hit = q.hit + o;
        for (int k = 0; k < 2; ++k) {
            L.q.sh_o[k] = q.sh_o[k] + o;
            L.q.sh_d[k] = q.sh_d[k] + o;
            L.q.sh_c[k] = q.sh_c[k] + o;
        }
        L.q.lpath = q.lpath + o;
        for (int k = 0; k < 2; ++k) {
            L.q.ebuf[k] = q.ebuf[k] + o;
            L.q.nbuf[k] = q.nbuf[k] + o;
        }
        L.gen = 0;
        r = A((void**)&L.q.counters, ctd::kCounterWords * 4, "counters");
        if (r == CT_OK) r = A((void**)&L.q.err, 64, "error flags");
        if (r == CT_OK) r = A((void**)&L.q.spill, (size_t)threads * ctd::kMaxStack * 4, "stack spill");
        if (r == CT_OK) r = A((void**)&L.wstats, 64, "work counters");
        if (r == CT_OK) r = A((void**)&L.d_hdr, sizeof(ChunkHdr), "chunk header");
        if (r == CT_OK && hipHostMalloc((void**)&L.h_read, kReadWords * 4, hipHostMallocDefault) != hipSuccess) {
            L.h_read = nullptr;
            r = fail(c, CT_ERR_OUT_OF_MEMORY, "pinned read-back buffer");
        }
        if (r == CT_OK && hipMemsetAsync(L.q.err, 0, 64, L.stream) != hipSuccess)
            r = fail(c, CT_ERR_DEVICE, "clear error flags");
    }
    if (r == CT_OK && hipStreamSynchronize(c->ln[0].stream) != hipSuccess)
        r = fail(c, CT_ERR_DEVICE, "clear radiance terms");
    if (r != CT_OK) {
        free_paths(c);
        return r;
    }
    c->lane_cap = cap;
    c->spill_threads = threads;
    return CT_OK;
}

// The lane is idle here (its previous chunk has finished); only its own
// captured graph refers to its tables.
int ensure_tables(ct_ctx* c, Lane& L, uint32_t nviews, uint32_t nsegs) {
    if (nviews > L.d_views_cap) {
        dfree(c, L.d_views);
        L.d_views = nullptr;
        const uint32_t n = std::max(nviews, 64u);
        const int r = dalloc(c, (void**)&L.d_views, (size_t)n * sizeof(ViewDesc), "view table");
        if (r != CT_OK) { L.d_views_cap = 0; return r; }
        L.d_views_cap = n;
        drop_lane_graph(L);
    }
    if (nsegs > L.d_segs_cap) {
        dfree(c, L.d_segs);
        L.d_segs = nullptr;
        const uint32_t n = std::max(nsegs, 256u);
        const int r = dalloc(c, (void**)&L.d_segs, (size_t)n * sizeof(Segment), "segment table");
        if (r != CT_OK) { L.d_segs_cap = 0; return r; }
        L.d_segs_cap = n;
        drop_lane_graph(L);
    }
    return CT_OK;
}

// The stride of pixel_of's strip order for n full strips: the integer
// nearest 0.618 n (the golden-ratio stride, which spreads every run of
// consecutive strips evenly over the image) that is coprime to n.
uint32_t strip_stride(uint32_t n) {
    if (n <= 2) return 1;
    auto gcd = [](uint32_t a, uint32_t b) {
        while (b) {
            const uint32_t t = a % b;
            a = b;
            b = t;
        }
        return a;
    };
    const uint32_t target = (uint32_t)std::lround(0.6180339887498949 * n);
    for (uint32_t d = 0; d < n; ++d) {
        if (target + d < n && gcd(target + d, n) == 1) return target + d;
        if (target >= d + 1 && gcd(target - d - 1, n) == 1) return target - d - 1;
    }
    return 1;
}

// Camera basis of sample(), Kernel.cu:274-290 (float, glm evaluation order).
ViewDesc make_view(const ct_camera& cam, int width, int height, int spp, uint32_t frame_base,
                   uint64_t accum_offset) {
    auto dot = [](const float* a, const float* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
    auto cross = [](const float* x, const float* y, float* r) {
        r[0] = x[1] * y[2] - y[1] * x[2];
        r[1] = x[2] * y[0] - y[2] * x[0];
        r[2] = x[0] * y[1] - y[0] * x[1];
    };
    auto normalize = [&](float* v) {
        const float s = 1.0f / std::sqrt(dot(v, v));
        v[0] = v[0] * s;
        v[1] = v[1] * s;
        v[2] = v[2] * s;
    };
    ViewDesc v;
    std::memset(&v, 0, sizeof(v));
    float u[3], vv[3];
    cross(cam.forward, cam.up, u);
    normalize(u);
    cross(u, cam.forward, vv);
    normalize(vv);
    const float dist = cam.plane_distance;
    const float aspect = (float)width / (float)height;
    for (int k = 0; k < 3; ++k) {
        v.pos[k] = cam.position[k];
        v.w[k] = cam.forward[k];
        v.hor[k] = u[k] * dist;
        v.ver[k] = vv[k] * dist;
    }
    for (int k = 0; k < 3; ++k) {
        if (width > height) v.hor[k] = v.hor[k] * aspect;
        else v.ver[k] = v.ver[k] * aspect;
    }
    v.invw = 1.0f / (float)width;
    v.invh = 1.0f / (float)height;
    v.width = width;
    v.height = height;
    v.spp = spp;
    v.frame_base = frame_base;
    v.accum_offset = accum_offset;
    v.strip_mul = strip_stride((uint32_t)height / ctd::kTileRows);
    return v;
}

static ctd::TraceOpts trace_opts(const ct_frame_opts& o) {
    ctd::TraceOpts t;
    t.accel = o.accel;
    t.traversal = o.traversal;
    t.teps = o.triangle_epsilon;
    return t;
}

static ctd::ShadeOpts shade_opts(const ct_frame_opts& o) {
    ctd::ShadeOpts s;
    s.max_depth = o.max_depth;
    s.rr_start = o.rr_start;
    s.bsdf_mode = o.bsdf_mode;
    s.shadow = o.shadow;
    s.indirect = o.indirect;
    s.neps = o.normal_epsilon;
    s.deps = o.dist_epsilon;
    s.teps = o.triangle_epsilon;
    s.light_pdf = o.light_pdf;
    s.sort = o.sort_by_material ? 1 : 0;
    return s;
}

// The kernel sequence of one chunk on a lane, up to the resolve
// (setGlobals -> generate -> depth x (trace, shade) -> trace of the last
// shadow rays), in steps: step 0 is the prologue, step 1 + b is bounce b,
// then finish().  Chunks of several lanes are enqueued step by step in turn,
// so every lane starts at once instead of after the host has enqueued the
// whole of the previous lanes' chunks.  timed records an event pair around
// every kernel.
// The bounce from which a chunk's remaining bounces run in one k_tail
// launch, or -1 (a counting frame: k_tail does not count node and triangle
// tests).
int tail_bounce(const ct_frame_opts& o, bool count) {
    if (count) return -1;
    int b0 = -1;
    if (o.rr_start >= 0) b0 = std::max(0, o.rr_start + kTailAfterRoulette);
    else if (kTailBeforeEnd > 0) b0 = std::max(0, o.max_depth - kTailBeforeEnd);
    return b0 >= 0 && b0 + 1 < o.max_depth ? b0 : -1;
}

struct BounceSeq {
    ct_ctx* c = nullptr;
    Lane* L = nullptr;
    ct_frame_opts o{};
    ctd::TraceOpts to{};
    ctd::ShadeOpts so{};
    unsigned long long* ws = nullptr;
    bool timed = false, nee = false;
    int tail = -1;  // k_tail from this bounce on, or -1
    int e = 0, sel = 0;

    BounceSeq() = default;
    BounceSeq(ct_ctx* c_, Lane* L_, const ct_frame_opts& o_, bool count, bool timed_)
        : c(c_), L(L_), o(o_), to(trace_opts(o_)), so(shade_opts(o_)), ws(count ? L_->wstats : nullptr),
          timed(timed_), nee(o_.shadow && c_->sc.n_lights > 0), tail(tail_bounce(o_, count)) {}
    int steps() const { return 1 + (tail >= 0 ? tail + 1 : o.max_depth); }
    void rec() {
        if (timed) (void)hipEventRecord(L->ev[e++], L->stream);
    }
    hipError_t step(int k) {
        hipStream_t s = L->stream;
        hipError_t r;
        if (k == 0) {
            rec();
            if ((r = ctk::launch_setup(L->q.counters, L->d_hdr, s)) != hipSuccess) return r;
            if ((r = ctk::launch_generate(c->shape, L->d_views, L->d_segs, L->d_hdr, c->lane_cap, L->q, s)) !=
                hipSuccess)
                return r;
            rec();
            return hipSuccess;
        }
        // trace(b): the extension rays of bounce b and the shadow rays of
        // bounce b - 1 (their radiance terms are added by shade(b + 1) or
        // resolve, in order); then shade(b) appends bounce b + 1's extension
        // rays and bounce b's shadow rays, or, at the tail bounce, k_tail
        // runs every remaining bounce
        const int b = k - 1;
        rec();
        if ((r = ctk::launch_trace(c->shape, c->sc, to, L->q, L->d_hdr, sel, b, ws, s)) != hipSuccess) return r;
        rec();
        rec();
        if (b == tail)
            r = ctk::launch_tail(c->shape, c->sc, to, so, L->d_views, L->d_segs, L->d_hdr, L->q, sel, b, s);
        else
            r = ctk::launch_shade(c->shape, c->sc, so, L->d_views, L->d_segs, L->d_hdr, L->q, sel, b, s);
        if (r != hipSuccess) return r;
        rec();
        sel ^= 1;
        return hipSuccess;
    }
    // the last bounce's shadow rays (the launch has no extension rays; k_tail
    // traces its own)
    hipError_t finish() {
        hipError_t r;
        rec();
        if (nee && tail < 0 &&
            (r = ctk::launch_trace(c->shape, c->sc, to, L->q, L->d_hdr, sel, o.max_depth, ws, L->stream)) !=
                hipSuccess)
            return r;
        rec();
        L->n_events = e;
        return hipSuccess;
    }
};

static hipError_t enqueue_bounces(ct_ctx* c, Lane& L, const ct_frame_opts& o, bool count, bool timed) {
    BounceSeq q(c, &L, o, count, timed);
    hipError_t r = hipSuccess;
    for (int k = 0; k < q.steps() && r == hipSuccess; ++k) r = q.step(k);
    return r == hipSuccess ? q.finish() : r;
}

static std::string graph_key(const ct_ctx* c, const Lane& L, const ct_frame_opts& o) {
    std::ostringstream os;
    os << c->lane_cap << ':' << o.max_depth << ':' << o.rr_start << ':' << o.accel << ':' << o.bsdf_mode << ':'
       << o.traversal << ':' << o.shadow << ':' << o.indirect << ':' << o.normal_epsilon << ':' << o.dist_epsilon
       << ':' << o.triangle_epsilon << ':' << o.light_pdf << ':' << o.sort_by_material << ':' << c->shape.block << ':'
       << c->shape.lds_depth << ':' << c->shape.grid2 << ':' << c->shape.grid4 << ':' << c->shape.grid8 << ':' << c->shape.gridc << ':'
       << c->shape.shade_blocks << ':' << (const void*)L.d_views << ':' << (const void*)L.d_segs << ':'
       << (const void*)L.q.counters << ':' << (const void*)c->sc.tris << ':' << (const void*)c->sc.mats;
    return os.str();
}

static constexpr int kEvTotal0 = 2 * (3 * ctd::kMaxBounce + 8) - 2;  // chunk start / end events
static constexpr int kEvResolve = 2 * (3 * ctd::kMaxBounce + 8) - 4;  // resolve pair (timed)

// Chunk launch in three parts.  stage_chunk uploads the chunk's tables;
// then either graph_chunk replays the lane's captured bounce loop or a
// BounceSeq enqueues it; close_chunk queues the resolve after the previous
// chunk's resolve (prev_resolved, may be null) and the read-back.
static int stage_chunk(ct_ctx* c, Lane& L, const std::vector<ViewDesc>& views, const std::vector<Segment>& segs,
                       uint32_t n, uint32_t n_tasks, bool count) {
    int r = ensure_tables(c, L, (uint32_t)views.size(), (uint32_t)segs.size());
    if (r != CT_OK) return r;
    hipStream_t s = L.stream;
    // tables staged in pinned memory: the copies run after this call returns,
    // and the lane's staging is rewritten only once its chunk has finished
    const size_t vb = views.size() * sizeof(ViewDesc), sb = segs.size() * sizeof(Segment);
    const size_t need = 256 + vb + sb;
    if (need > L.h_stage_bytes) {
        if (L.h_stage) (void)hipHostFree(L.h_stage);
        L.h_stage = nullptr;
        L.h_stage_bytes = 0;
        const size_t bytes = std::max<size_t>(need * 2, 1u << 16);
        if (hipHostMalloc((void**)&L.h_stage, bytes, hipHostMallocDefault) != hipSuccess) {
            L.h_stage = nullptr;
            return fail(c, CT_ERR_OUT_OF_MEMORY, "pinned staging of chunk tables");
        }
        L.h_stage_bytes = bytes;
    }
    // the chunk's generation: its radiance terms' stamps differ from any
    // left in the lane's term buffers by earlier chunks (cleared when the
    // 26-bit counter wraps)
    if (++L.gen >= (1u << 26)) {
        L.gen = 1;
        for (int k = 0; k < 2; ++k) {
            HIPCHK(hipMemsetAsync(L.q.ebuf[k], 0, c->lane_cap * sizeof(float4), s), "clear radiance terms");
            HIPCHK(hipMemsetAsync(L.q.nbuf[k], 0, c->lane_cap * sizeof(float4), s), "clear radiance terms");
        }
    }
    const ChunkHdr h{n, n_tasks, (uint32_t)segs.size(), L.gen};
    std::memcpy(L.h_stage, &h, sizeof(h));
    std::memcpy(L.h_stage + 256, views.data(), vb);
    std::memcpy(L.h_stage + 256 + vb, segs.data(), sb);
    HIPCHK(hipMemcpyAsync(L.d_views, L.h_stage + 256, vb, hipMemcpyHostToDevice, s), "upload views");
    HIPCHK(hipMemcpyAsync(L.d_segs, L.h_stage + 256 + vb, sb, hipMemcpyHostToDevice, s), "upload segments");
    HIPCHK(hipMemcpyAsync(L.d_hdr, L.h_stage, sizeof(h), hipMemcpyHostToDevice, s), "upload chunk header");
    if (count) HIPCHK(hipMemsetAsync(L.wstats, 0, 32, s), "clear work counters");
    HIPCHK(hipEventRecord(L.ev[kEvTotal0], s), "event");
    return CT_OK;
}

// The bounce loop from the lane's graph, captured again when the options,
// launch shape or buffers changed.
static int graph_chunk(ct_ctx* c, Lane& L, const ct_frame_opts& o) {
    hipStream_t s = L.stream;
    const std::string key = graph_key(c, L, o);
    if (L.graph_key != key || !L.graph_exec) {
        if (L.graph_exec) (void)hipGraphExecDestroy(L.graph_exec);
        L.graph_exec = nullptr;
        L.graph_key.clear();
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture");
        const hipError_t ce = enqueue_bounces(c, L, o, false, false);
        const hipError_t ee = hipStreamEndCapture(s, &g);
        if (ce != hipSuccess) { if (g) (void)hipGraphDestroy(g); return hip_fail(c, ce, "capture bounce loop"); }
        if (ee != hipSuccess) return hip_fail(c, ee, "end capture");
        const hipError_t ie = hipGraphInstantiate(&L.graph_exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ie != hipSuccess) { L.graph_exec = nullptr; return hip_fail(c, ie, "instantiate graph"); }
        L.graph_key = key;
    }
    HIPCHK(hipGraphLaunch(L.graph_exec, s), "launch graph");
    return CT_OK;
}

static int close_chunk(ct_ctx* c, Lane& L, uint32_t n, float4* accum, const ct_frame_opts& o, bool timed,
                       bool count, hipEvent_t prev_resolved) {
    hipStream_t s = L.stream;
    // accumulator updates in chunk order
    if (prev_resolved) HIPCHK(hipStreamWaitEvent(s, prev_resolved, 0), "order resolves");
    if (timed) HIPCHK(hipEventRecord(L.ev[kEvResolve], s), "event");
    HIPCHK(ctk::launch_resolve(c->shape, L.d_views, L.d_segs, L.d_hdr, c->lane_cap, L.q, accum, s),
           "launch resolve");
    if (timed) HIPCHK(hipEventRecord(L.ev[kEvResolve + 1], s), "event");
    HIPCHK(hipEventRecord(L.resolved, s), "event");
    HIPCHK(hipEventRecord(L.ev[kEvTotal0 + 1], s), "event");
    // queue sizes (ray counts), error flags and work counters
    HIPCHK(hipMemcpyAsync(L.h_read, L.q.counters, kReadCounters * 4, hipMemcpyDeviceToHost, s), "read counters");
    HIPCHK(hipMemcpyAsync(L.h_read + kReadErr, L.q.err, 4, hipMemcpyDeviceToHost, s), "read error flags");
    if (count) HIPCHK(hipMemcpyAsync(L.h_read + kReadWork, L.wstats, 32, hipMemcpyDeviceToHost, s), "read work");
    L.pending = true;
    L.n_paths = n;
    L.max_depth = o.max_depth;
    L.nee = o.shadow && c->sc.n_lights > 0;
    L.tail = tail_bounce(o, count);
    L.timed = timed;
    L.count = count;
    return CT_OK;
}

// Wait for lane L's chunk and add its statistics.
static int finish_chunk(ct_ctx* c, Lane& L, ChunkStats* st) {
    if (!L.pending) return CT_OK;
    L.pending = false;
    HIPCHK(hipStreamSynchronize(L.stream), "chunk");
    const uint32_t* cnt = L.h_read;
    if (cnt[kReadErr]) {
        (void)hipMemsetAsync(L.q.err, 0, 4, L.stream);
        return fail(c, CT_ERR_STACK_OVERFLOW, "traversal stack bound exceeded (malformed tree)");
    }
    if (!st) return CT_OK;
    st->chunks += 1;
    st->primary += cnt[0];
    for (int b = 1; b < L.max_depth; ++b) st->extension += cnt[b];
    for (int b = 0; b < L.max_depth; ++b) st->shadow += cnt[(ctd::kMaxBounce + 1) + b];
    st->extension += cnt[2 * (ctd::kMaxBounce + 1)];  // k_tail's (cnt_tail_ext, cnt_tail_shadow)
    st->shadow += cnt[2 * (ctd::kMaxBounce + 1) + 1];
    st->samples += L.n_paths;
    float ms = 0;
    if (hipEventElapsedTime(&ms, L.ev[kEvTotal0], L.ev[kEvTotal0 + 1]) == hipSuccess) st->total_ms += ms;
    if (L.timed) {
        // event pairs: [generate] then per bounce [trace][shade] (at the tail
        // bounce [trace][k_tail], the last bounce the chunk launches); [trace
        // of the last shadow rays]; [resolve]
        int e = 0;
        auto pair = [&](int i) {
            float m = 0;
            (void)hipEventElapsedTime(&m, L.ev[i], L.ev[i + 1]);
            return (double)m;
        };
        st->other_ms += pair(e);
        e += 2;
        const int nb = L.tail >= 0 ? L.tail + 1 : L.max_depth;
        for (int b = 0; b < nb; ++b) {
            st->extend_ms += pair(e);
            st->extend_launches += 1;
            if (b == L.tail) st->tail_ms += pair(e + 2);
            else st->shade_ms += pair(e + 2);
            e += 4;
        }
        if (L.nee && L.tail < 0) {
            st->connect_ms += pair(e);
            st->extend_launches += 1;
        }
        st->other_ms += pair(kEvResolve);
    }
    if (L.count) {
        unsigned long long w[4];
        std::memcpy(w, cnt + kReadWork, sizeof(w));
        for (int k = 0; k < 4; ++k) st->work[k] += w[k];
    }
    return CT_OK;
}

// After a failed chunk: wait for every lane and clear all lanes' error
// flags, so a chunk of the failed call that was never finished cannot fail
// a later call.  The accumulator may hold part of the failed call's samples.
static void abort_lanes(ct_ctx* c) {
    sync_lanes(c);
    for (Lane& L : c->ln) {
        if (L.q.err) (void)hipMemsetAsync(L.q.err, 0, 4, L.stream);
        if (L.stream) (void)hipStreamSynchronize(L.stream);
    }
}

// Paths of a workload below which it is not split over the lanes.
static constexpr uint64_t kSplitMin = 1u << 18;

int run_paths(ct_ctx* c, const std::vector<ViewDesc>& views, const std::vector<Run>& runs, uint32_t cap,
              float4* accum, const ct_frame_opts& o, int flags, ChunkStats* st) {
    const bool timed = flags & kRunTimed, count = flags & kRunCount, use_graph = flags & kRunGraph,
               serial = flags & kRunSerial;
    uint64_t total = 0;
    for (const Run& run : runs) total += run.n;
    if (total >= kSplitMin && cap > 0) {
        // equal chunks, a multiple of the lane count, so every lane gets the same work
        uint64_t nch = (total + cap - 1) / cap;
        nch = std::max<uint64_t>(kLanes, (nch + kLanes - 1) / kLanes * kLanes);
        cap = (uint32_t)((total + nch - 1) / nch);
    }
    int r = ensure_paths(c, cap);
    if (r != CT_OK) return r;
    cap = std::min(cap, c->lane_cap);
    // work queued on lane 0 by the caller (accumulator clears) comes first
    HIPCHK(hipEventRecord(c->ln[0].resolved, c->ln[0].stream), "event");
    for (int k = 1; k < kLanes; ++k) HIPCHK(hipStreamWaitEvent(c->ln[k].stream, c->ln[0].resolved, 0), "order lanes");
    hipEvent_t prev = nullptr;
    int lane = 0;
    // The first kLanes chunks (one per lane) are launched together, their
    // bounce loops enqueued step by step in turn, so no lane waits for the
    // host to enqueue another lane's whole chunk.  Later chunks start alone,
    // each as soon as its lane's previous chunk has finished.
    int launched = 0;
    struct Staged {
        std::vector<Segment> segs;
        uint32_t fill, tasks;
    };
    std::vector<Staged> group;
    auto launch_group = [&]() -> int {
        const int g = (int)group.size();
        int rr;
        for (int i = 0; i < g; ++i) {
            Lane& L = c->ln[(lane + i) % kLanes];
            if ((rr = finish_chunk(c, L, st)) != CT_OK) return rr;  // the lane's previous chunk
            if ((rr = stage_chunk(c, L, views, group[i].segs, group[i].fill, group[i].tasks, count)) != CT_OK)
                return rr;
        }
        if (use_graph && !count && !timed) {
            for (int i = 0; i < g; ++i)
                if ((rr = graph_chunk(c, c->ln[(lane + i) % kLanes], o)) != CT_OK) return rr;
        } else {
            BounceSeq seq[kLanes];
            for (int i = 0; i < g; ++i) seq[i] = BounceSeq(c, &c->ln[(lane + i) % kLanes], o, count, timed);
            for (int k = 0; k < seq[0].steps(); ++k)
                for (int i = 0; i < g; ++i) HIPCHK(seq[i].step(k), "launch chunk");
            for (int i = 0; i < g; ++i) HIPCHK(seq[i].finish(), "launch chunk");
        }
        for (int i = 0; i < g; ++i) {
            Lane& L = c->ln[(lane + i) % kLanes];
            if ((rr = close_chunk(c, L, group[i].fill, accum, o, timed, count, prev)) != CT_OK) return rr;
            prev = L.resolved;
            if (serial && (rr = finish_chunk(c, L, st)) != CT_OK) return rr;  // done before the next starts
        }
        lane = (lane + g) % kLanes;
        launched += g;
        group.clear();
        return CT_OK;
    };
    auto flush = [&](std::vector<Segment>& chunk_segs, uint32_t chunk_fill, uint32_t chunk_tasks) -> int {
        group.push_back(Staged{chunk_segs, chunk_fill, chunk_tasks});
        return (!serial && launched + (int)group.size() < kLanes) ? CT_OK : launch_group();
    };
    std::vector<Segment> segs;
    uint32_t fill = 0, tasks = 0;
    for (const Run& run : runs) {
        uint32_t first = run.first_path, left = run.n;
        const uint32_t spp = (uint32_t)views[run.view].spp;
        while (left > 0) {
            const uint32_t take = std::min(left, cap - fill);
            Segment sg;
            std::memset(&sg, 0, sizeof(sg));
            sg.chunk_offset = fill;
            sg.view = run.view;
            sg.first_path = first;
            sg.n = take;
            sg.pix_offset = tasks;
            tasks += (first + take - 1) / spp - first / spp + 1;
            segs.push_back(sg);
            fill += take;
            first += take;
            left -= take;
            if (fill == cap) {
                r = flush(segs, fill, tasks);
                if (r != CT_OK) { abort_lanes(c); return r; }
                segs.clear();
                fill = tasks = 0;
            }
        }
    }
    if (fill > 0) {
        r = flush(segs, fill, tasks);
        if (r != CT_OK) { abort_lanes(c); return r; }
    }
    if (!group.empty() && (r = launch_group()) != CT_OK) { abort_lanes(c); return r; }
    for (int k = 0; k < kLanes; ++k) {
        r = finish_chunk(c, c->ln[lane], st);
        lane = (lane + 1) % kLanes;
        if (r != CT_OK) { abort_lanes(c); return r; }
    }
    // the other lanes' work is done: later work on lane 0 needs no further ordering
    return CT_OK;
}

}  // namespace ctr

using namespace ctr;

// ================================================================ C ABI ==
extern "C" {

int ct_abi_version(void) { return CT_ABI_VERSION; }

void ct_frame_opts_default(ct_frame_opts* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->max_depth = 5;
    o->spp = 1;
    o->rr_start = -1;
    o->accel = CT_ACCEL_BVH4;
    o->bsdf_mode = CT_BSDF_REFERENCE;
    o->traversal = CT_TRAV_EXACT;
    o->shadow = 1;
    o->indirect = 1;
    o->normal_epsilon = 1e-5f;
    o->dist_epsilon = 1e-5f;
    o->triangle_epsilon = 1e-3f;
    o->sort_by_material = 0;
    o->light_pdf = CT_LIGHTPDF_REFERENCE;
}

const char* ct_last_error(const ct_ctx* c) { return c ? c->err.c_str() : "null context"; }

void ctr_apply_shape(ct_ctx* c, int block, int lds_depth) {
    c->shape.block = block;
    c->shape.lds_depth = lds_depth;
    if (!ctk::set_trace_grids(c->shape, c->num_cus)) {
        c->shape.grid2 = c->shape.grid4 = c->shape.grid8 = c->shape.gridc = c->shape.full2 = c->shape.full4 =
            c->shape.full8 = c->shape.fullc = c->shape.trace_blocks = c->shape.tail_blocks = std::max(8, c->num_cus * 4);
    }
    c->shape.shade_blocks = std::max(1, c->num_cus * 8);
}

int ct_ctx_create(const ct_device_opts* opts, ct_ctx** out) {
    if (!out) return CT_ERR_INVALID_ARG;
    *out = nullptr;
    ct_ctx* c = new ct_ctx();
    ct_device_opts d;
    std::memset(&d, 0, sizeof(d));
    if (opts) d = *opts;
    const int block = d.block_size > 0 ? d.block_size : 256;
    const int lds = d.stack_lds > 0 ? d.stack_lds : 16;
    if (block % 64 != 0 || block > 1024 || lds > ctd::kMaxStack) {
        delete c;
        return CT_ERR_INVALID_ARG;
    }
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0) {
        delete c;
        return CT_ERR_DEVICE;
    }
    if (d.device < 0 || d.device >= ndev) {
        delete c;
        return CT_ERR_INVALID_ARG;
    }
    c->device = d.device;
    if (hipSetDevice(c->device) != hipSuccess) { delete c; return CT_ERR_DEVICE; }
    hipDeviceProp_t props;
    if (hipGetDeviceProperties(&props, c->device) != hipSuccess) { delete c; return CT_ERR_DEVICE; }
    // the marketing name can be empty; the architecture, CU count and memory
    // identify the part
    std::ostringstream nm;
    nm << (props.name[0] ? props.name : "AMD GPU") << " (" << (props.gcnArchName[0] ? props.gcnArchName : "?") << ", "
       << props.multiProcessorCount << " CUs, " << (props.totalGlobalMem >> 30) << " GiB)";
    c->name = nm.str();
    c->num_cus = props.multiProcessorCount;
    size_t freeb = 0, totalb = 0;
    if (hipMemGetInfo(&freeb, &totalb) != hipSuccess) { delete c; return CT_ERR_DEVICE; }
    c->budget = d.mem_budget ? d.mem_budget : (uint64_t)(0.9 * (double)freeb);
    for (Lane& L : c->ln) {
        bool ok = hipStreamCreate(&L.stream) == hipSuccess &&
                  hipEventCreateWithFlags(&L.resolved, hipEventDisableTiming) == hipSuccess;
        for (auto& ev : L.ev)
            if (ok) ok = hipEventCreate(&ev) == hipSuccess;
        if (!ok) {
            ct_ctx_destroy(c);
            return CT_ERR_DEVICE;
        }
    }
    if (scratch_init(c) != CT_OK) {
        ct_ctx_destroy(c);
        return CT_ERR_DEVICE;
    }
    ctr_apply_shape(c, block, lds);
    if (d.chunk_paths > 0) c->chunk_default = (uint32_t)d.chunk_paths;
    *out = c;
    return CT_OK;
}

void ct_ctx_destroy(ct_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    sync_lanes(c);
    drop_graphs(c);
    for (auto& a : c->allocs) (void)hipFree(a.first);
    c->allocs.clear();
    for (Lane& L : c->ln) {
        if (L.h_read) (void)hipHostFree(L.h_read);
        if (L.h_stage) (void)hipHostFree(L.h_stage);
        for (auto& ev : L.ev)
            if (ev) (void)hipEventDestroy(ev);
        if (L.resolved) (void)hipEventDestroy(L.resolved);
        if (L.stream) (void)hipStreamDestroy(L.stream);
    }
    delete c;
}

int ct_device_name(const ct_ctx* c, char* buf, size_t len) {
    if (!c || !buf || len == 0) return CT_ERR_INVALID_ARG;
    std::snprintf(buf, len, "%s", c->name.c_str());
    return CT_OK;
}

}  // extern "C"
