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

#define HIPCHK(expr, what)                                     \
    do {                                                       \
        hipError_t e_ = (expr);                                \
        if (e_ != hipSuccess) return ctr::hip_fail(c, e_, what); \
    } while (0)

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

static void drop_graph(ct_ctx* c) {
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    c->graph_exec = nullptr;
    c->graph_key.clear();
}

int validate_opts(ct_ctx* c, const ct_frame_opts& o) {
    if (o.max_depth < 1 || o.max_depth > ctd::kMaxBounce)
        return fail(c, CT_ERR_INVALID_ARG, "max_depth must be in [1, 64]");
    if (o.spp < 1 || o.spp > (1 << 20)) return fail(c, CT_ERR_INVALID_ARG, "spp must be in [1, 2^20]");
    if (o.accel != CT_ACCEL_BVH2 && o.accel != CT_ACCEL_BVH4) return fail(c, CT_ERR_INVALID_ARG, "bad accel");
    if (o.bsdf_mode != CT_BSDF_REFERENCE && o.bsdf_mode != CT_BSDF_TYPED)
        return fail(c, CT_ERR_INVALID_ARG, "bad bsdf_mode");
    if (o.traversal != CT_TRAV_EXACT && o.traversal != CT_TRAV_FAST)
        return fail(c, CT_ERR_INVALID_ARG, "bad traversal");
    if (o.light_pdf != CT_LIGHTPDF_REFERENCE && o.light_pdf != CT_LIGHTPDF_UNIFORM)
        return fail(c, CT_ERR_INVALID_ARG, "bad light_pdf");
    if (!(o.triangle_epsilon >= 0.0f) || !(o.normal_epsilon >= 0.0f) || !(o.dist_epsilon >= 0.0f))
        return fail(c, CT_ERR_INVALID_ARG, "epsilons must be finite and >= 0");
    if (!c->has_scene) return fail(c, CT_ERR_NO_SCENE, "no scene uploaded");
    if (o.accel == CT_ACCEL_BVH2 && !c->has_bvh2) return fail(c, CT_ERR_INVALID_ARG, "scene has no binary tree");
    if (o.accel == CT_ACCEL_BVH4 && !c->has_bvh4) return fail(c, CT_ERR_INVALID_ARG, "scene has no quad tree");
    if (o.traversal == CT_TRAV_FAST && (c->sc.n_fnodes == 0 || c->wide_stack_bound > (uint32_t)ctd::kMaxStack))
        return fail(c, CT_ERR_INVALID_ARG, "fast traversal not available for this scene (wide tree too deep)");
    return CT_OK;
}

static void free_paths(ct_ctx* c) {
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->side) (void)hipStreamSynchronize(c->side);
    for (void* p : c->path_bufs) dfree(c, p);
    c->path_bufs.clear();
    c->q = ctd::Queues{};
    c->cap = 0;
    drop_graph(c);
}

int ensure_paths(ct_ctx* c, uint32_t cap) {
    cap = std::max<uint32_t>(cap, 64u);
    const uint32_t threads = (uint32_t)(c->shape.trace_blocks * c->shape.block);
    if (c->cap >= cap && c->spill_threads >= threads && c->q.counters) return CT_OK;
    cap = std::max(cap, c->cap);
    free_paths(c);
    auto A = [&](void** p, size_t bytes, const char* w) -> int {
        const int r = dalloc(c, p, bytes, w);
        if (r == CT_OK) c->path_bufs.push_back(*p);
        return r;
    };
    const size_t f4 = (size_t)cap * sizeof(float4);
    int r = CT_OK;
    for (int k = 0; k < 2 && r == CT_OK; ++k) {
        r = A((void**)&c->q.ray_o[k], f4, "ray origins");
        if (r == CT_OK) r = A((void**)&c->q.ray_d[k], f4, "ray directions");
        if (r == CT_OK) r = A((void**)&c->q.thr[k], f4, "throughput");
        if (r == CT_OK) r = A((void**)&c->q.path[k], (size_t)cap * 4, "path words");
    }
    if (r == CT_OK) r = A((void**)&c->q.hit, (size_t)cap * sizeof(float2), "hits");
    if (r == CT_OK) r = A((void**)&c->q.sh_o, f4, "shadow origins");
    if (r == CT_OK) r = A((void**)&c->q.sh_d, f4, "shadow directions");
    if (r == CT_OK) r = A((void**)&c->q.sh_c, f4, "shadow colours");
    if (r == CT_OK) r = A((void**)&c->q.lpath, f4, "path radiance");
    if (r == CT_OK) r = A((void**)&c->q.counters, ctd::kCounterWords * 4, "counters");
    if (r == CT_OK) r = A((void**)&c->q.err, 64, "error flags");
    if (r == CT_OK) r = A((void**)&c->q.spill, (size_t)threads * ctd::kMaxStack * 4, "stack spill");
    if (r == CT_OK) r = A((void**)&c->spill_side, (size_t)threads * ctd::kMaxStack * 4, "stack spill (side)");
    if (r == CT_OK) r = A((void**)&c->wstats, 64, "work counters");
    if (r == CT_OK) r = A((void**)&c->d_hdr, sizeof(ChunkHdr), "chunk header");
    if (r != CT_OK) {
        free_paths(c);
        return r;
    }
    HIPCHK(hipMemsetAsync(c->q.err, 0, 64, c->stream), "clear error flags");
    c->cap = cap;
    c->spill_threads = threads;
    return CT_OK;
}

int ensure_tables(ct_ctx* c, uint32_t nviews, uint32_t nsegs) {
    if (nviews > c->d_views_cap) {
        dfree(c, c->d_views);
        c->d_views = nullptr;
        const uint32_t n = std::max(nviews, 64u);
        const int r = dalloc(c, (void**)&c->d_views, (size_t)n * sizeof(ViewDesc), "view table");
        if (r != CT_OK) { c->d_views_cap = 0; return r; }
        c->d_views_cap = n;
        drop_graph(c);
    }
    if (nsegs > c->d_segs_cap) {
        dfree(c, c->d_segs);
        c->d_segs = nullptr;
        const uint32_t n = std::max(nsegs, 256u);
        const int r = dalloc(c, (void**)&c->d_segs, (size_t)n * sizeof(Segment), "segment table");
        if (r != CT_OK) { c->d_segs_cap = 0; return r; }
        c->d_segs_cap = n;
        drop_graph(c);
    }
    return CT_OK;
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

// The kernel sequence of one chunk (setGlobals -> generate -> depth x
// (extend, shade, connect) -> resolve).  ev != nullptr records an event pair
// around every kernel.
static hipError_t enqueue_chunk(ct_ctx* c, float4* accum, const ct_frame_opts& o, bool count, bool timed) {
    const ctd::TraceOpts to = trace_opts(o);
    const ctd::ShadeOpts so = shade_opts(o);
    hipStream_t s = c->stream;
    unsigned long long* ws = count ? c->wstats : nullptr;
    int e = 0;
    auto rec = [&]() {
        if (timed) (void)hipEventRecord(c->ev[e++], s);
    };
    hipError_t r;
    rec();
    if ((r = ctk::launch_setup(c->q.counters, c->d_hdr, s)) != hipSuccess) return r;
    if ((r = ctk::launch_generate(c->shape, c->d_views, c->d_segs, c->d_hdr, c->cap, c->q, s)) != hipSuccess)
        return r;
    rec();
    // connect(b) on the side stream overlaps extend(b + 1): both read queues
    // shade(b) wrote, and only connect and shade add to a path's radiance, so
    // shade(b + 1) and resolve wait for connect(b) (the additions keep their
    // order).  Its traversal stacks spill to their own area.
    const bool nee = o.shadow && c->sc.n_lights > 0;
    ctd::Queues qside = c->q;
    qside.spill = c->spill_side;
    int sel = 0;
    for (int b = 0; b < o.max_depth; ++b) {
        rec();
        if ((r = ctk::launch_extend(c->shape, c->sc, to, c->q, sel, b, ws, s)) != hipSuccess) return r;
        rec();
        if (nee && b > 0 && (r = hipStreamWaitEvent(s, c->join[b - 1], 0)) != hipSuccess) return r;
        rec();
        if ((r = ctk::launch_shade(c->shape, c->sc, so, c->d_views, c->d_segs, c->d_hdr, c->q, sel, b, s)) !=
            hipSuccess)
            return r;
        rec();
        if (nee) {
            if ((r = hipEventRecord(c->fork[b], s)) != hipSuccess) return r;
            if ((r = hipStreamWaitEvent(c->side, c->fork[b], 0)) != hipSuccess) return r;
            if (timed) (void)hipEventRecord(c->ev[e], c->side);
            ++e;
            if ((r = ctk::launch_connect(c->shape, c->sc, to, qside, b, ws, c->side)) != hipSuccess) return r;
            if (timed) (void)hipEventRecord(c->ev[e], c->side);
            ++e;
            if ((r = hipEventRecord(c->join[b], c->side)) != hipSuccess) return r;
        } else {
            rec();
            rec();
        }
        sel ^= 1;
    }
    if (nee && (r = hipStreamWaitEvent(s, c->join[o.max_depth - 1], 0)) != hipSuccess) return r;
    rec();
    if ((r = ctk::launch_resolve(c->shape, c->d_views, c->d_segs, c->d_hdr, c->cap, c->q.lpath, accum, s)) !=
        hipSuccess)
        return r;
    rec();
    c->n_events = e;
    return hipSuccess;
}

static std::string graph_key(const ct_ctx* c, const float4* accum, const ct_frame_opts& o) {
    std::ostringstream os;
    os << (const void*)accum << ':' << c->cap << ':' << o.max_depth << ':' << o.rr_start << ':' << o.accel << ':'
       << o.bsdf_mode << ':' << o.traversal << ':' << o.shadow << ':' << o.indirect << ':' << o.normal_epsilon << ':'
       << o.dist_epsilon << ':' << o.triangle_epsilon << ':' << o.light_pdf << ':' << o.sort_by_material << ':' << c->shape.block << ':' << c->shape.lds_depth << ':'
       << c->shape.grid2 << ':' << c->shape.grid4 << ':' << c->shape.grid8 << ':' << c->shape.shade_blocks << ':' << (const void*)c->d_views << ':'
       << (const void*)c->d_segs << ':' << (const void*)c->sc.tris << ':' << (const void*)c->sc.mats;
    return os.str();
}

int run_chunk(ct_ctx* c, const std::vector<ViewDesc>& views, const std::vector<Segment>& segs, uint32_t n,
              uint32_t n_tasks, float4* accum, const ct_frame_opts& o, bool timed, bool count, bool use_graph,
              ChunkStats* st) {
    int r = ensure_tables(c, (uint32_t)views.size(), (uint32_t)segs.size());
    if (r != CT_OK) return r;
    ChunkHdr h{n, n_tasks, (uint32_t)segs.size(), 0u};
    HIPCHK(hipMemcpyAsync(c->d_views, views.data(), views.size() * sizeof(ViewDesc), hipMemcpyHostToDevice,
                          c->stream),
           "upload views");
    HIPCHK(hipMemcpyAsync(c->d_segs, segs.data(), segs.size() * sizeof(Segment), hipMemcpyHostToDevice, c->stream),
           "upload segments");
    HIPCHK(hipMemcpyAsync(c->d_hdr, &h, sizeof(h), hipMemcpyHostToDevice, c->stream), "upload chunk header");
    if (count) HIPCHK(hipMemsetAsync(c->wstats, 0, 32, c->stream), "clear work counters");
    hipEvent_t g0 = c->ev[2 * (3 * ctd::kMaxBounce + 8) - 2], g1 = c->ev[2 * (3 * ctd::kMaxBounce + 8) - 1];
    if (use_graph && !count && !timed) {
        const std::string key = graph_key(c, accum, o);
        if (c->graph_key != key || !c->graph_exec) {
            drop_graph(c);
            hipGraph_t g = nullptr;
            HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "begin capture");
            const hipError_t ce = enqueue_chunk(c, accum, o, false, false);
            const hipError_t ee = hipStreamEndCapture(c->stream, &g);
            if (ce != hipSuccess) { if (g) (void)hipGraphDestroy(g); return hip_fail(c, ce, "capture bounce loop"); }
            if (ee != hipSuccess) return hip_fail(c, ee, "end capture");
            const hipError_t ie = hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ie != hipSuccess) { c->graph_exec = nullptr; return hip_fail(c, ie, "instantiate graph"); }
            c->graph_key = key;
        }
        HIPCHK(hipEventRecord(g0, c->stream), "event");
        HIPCHK(hipGraphLaunch(c->graph_exec, c->stream), "launch graph");
        HIPCHK(hipEventRecord(g1, c->stream), "event");
    } else {
        HIPCHK(hipEventRecord(g0, c->stream), "event");
        HIPCHK(enqueue_chunk(c, accum, o, count, timed), "launch chunk");
        HIPCHK(hipEventRecord(g1, c->stream), "event");
    }
    // queue sizes (ray counts) and error flags
    uint32_t cnt[2 * (ctd::kMaxBounce + 1)];
    uint32_t errflag = 0;
    HIPCHK(hipMemcpyAsync(cnt, c->q.counters, sizeof(cnt), hipMemcpyDeviceToHost, c->stream), "read counters");
    HIPCHK(hipMemcpyAsync(&errflag, c->q.err, 4, hipMemcpyDeviceToHost, c->stream), "read error flags");
    unsigned long long w[4] = {0, 0, 0, 0};
    if (count) HIPCHK(hipMemcpyAsync(w, c->wstats, 32, hipMemcpyDeviceToHost, c->stream), "read work counters");
    HIPCHK(hipStreamSynchronize(c->stream), "chunk");
    if (errflag) {
        (void)hipMemsetAsync(c->q.err, 0, 4, c->stream);
        return fail(c, CT_ERR_STACK_OVERFLOW, "traversal stack bound exceeded (malformed tree)");
    }
    if (st) {
        st->primary += cnt[0];
        for (int b = 1; b < o.max_depth; ++b) st->extension += cnt[b];
        for (int b = 0; b < o.max_depth; ++b) st->shadow += cnt[(ctd::kMaxBounce + 1) + b];
        st->samples += n;
        float ms = 0;
        if (hipEventElapsedTime(&ms, g0, g1) == hipSuccess) st->total_ms += ms;
        if (timed) {
            // event pairs: [generate] then per bounce [extend][shade][connect], then [resolve]
            int e = 0;
            auto pair = [&]() {
                float m = 0;
                (void)hipEventElapsedTime(&m, c->ev[e], c->ev[e + 1]);
                e += 2;
                return (double)m;
            };
            st->other_ms += pair();
            for (int b = 0; b < o.max_depth; ++b) {
                st->extend_ms += pair();
                st->extend_launches += 1;
                st->shade_ms += pair();
                st->connect_ms += pair();
            }
            st->other_ms += pair();
        }
        for (int k = 0; k < 4; ++k) st->work[k] += w[k];
    }
    return CT_OK;
}

int run_paths(ct_ctx* c, const std::vector<ViewDesc>& views, const std::vector<Run>& runs, uint32_t cap,
              float4* accum, const ct_frame_opts& o, bool timed, bool count, bool use_graph, ChunkStats* st) {
    int r = ensure_paths(c, cap);
    if (r != CT_OK) return r;
    cap = std::min(cap, c->cap);
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
                r = run_chunk(c, views, segs, fill, tasks, accum, o, timed, count, use_graph, st);
                if (r != CT_OK) return r;
                segs.clear();
                fill = tasks = 0;
            }
        }
    }
    if (fill > 0) {
        r = run_chunk(c, views, segs, fill, tasks, accum, o, timed, count, use_graph, st);
        if (r != CT_OK) return r;
    }
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
        c->shape.grid2 = c->shape.grid4 = c->shape.grid8 = c->shape.trace_blocks = std::max(8, c->num_cus * 4);
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
    if (hipStreamCreate(&c->stream) != hipSuccess) { delete c; return CT_ERR_DEVICE; }
    if (hipStreamCreate(&c->side) != hipSuccess) { (void)hipStreamDestroy(c->stream); delete c; return CT_ERR_DEVICE; }
    for (int b = 0; b < ctd::kMaxBounce; ++b)
        if (hipEventCreateWithFlags(&c->fork[b], hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&c->join[b], hipEventDisableTiming) != hipSuccess) {
            delete c;
            return CT_ERR_DEVICE;
        }
    for (auto& ev : c->ev)
        if (hipEventCreate(&ev) != hipSuccess) { delete c; return CT_ERR_DEVICE; }
    const int block = d.block_size > 0 ? d.block_size : 256;
    const int lds = d.stack_lds > 0 ? d.stack_lds : 16;
    if (block % 64 != 0 || block > 1024 || lds > ctd::kMaxStack) { delete c; return CT_ERR_INVALID_ARG; }
    ctr_apply_shape(c, block, lds);
    if (d.chunk_paths > 0) c->chunk_default = (uint32_t)d.chunk_paths;
    *out = c;
    return CT_OK;
}

void ct_ctx_destroy(ct_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->side) (void)hipStreamSynchronize(c->side);
    drop_graph(c);
    for (auto& a : c->allocs) (void)hipFree(a.first);
    c->allocs.clear();
    for (auto& ev : c->ev)
        if (ev) (void)hipEventDestroy(ev);
    for (int b = 0; b < ctd::kMaxBounce; ++b) {
        if (c->fork[b]) (void)hipEventDestroy(c->fork[b]);
        if (c->join[b]) (void)hipEventDestroy(c->join[b]);
    }
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int ct_device_name(const ct_ctx* c, char* buf, size_t len) {
    if (!c || !buf || len == 0) return CT_ERR_INVALID_ARG;
    std::snprintf(buf, len, "%s", c->name.c_str());
    return CT_OK;
}

}  // extern "C"
