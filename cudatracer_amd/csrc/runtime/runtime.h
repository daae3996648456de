// Internal runtime of the C ABI: the context, its device buffers and the
// chunk driver shared by the plain frame loop, the batch API and the server.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../../include/cudatracer.h"
#include "../device/dev_common.h"
#include "../device/launch.h"

struct ChunkStats {
    uint64_t primary = 0, extension = 0, shadow = 0, samples = 0;
    double extend_ms = 0, shade_ms = 0, connect_ms = 0, other_ms = 0, total_ms = 0, tail_ms = 0;
    uint64_t work[4] = {0, 0, 0, 0};  // closest-hit node/tri tests, shadow node/tri tests
    int extend_launches = 0;
    int chunks = 0;
};

#ifndef CT_LANES
#define CT_LANES 2
#endif
// Pipelines the chunk driver rotates through.
constexpr int kLanes = CT_LANES;
// The bounce from which a chunk's remaining bounces run in one k_tail
// launch (tail_bounce): with roulette, rr_start + CT_TAIL_AFTER_ROULETTE;
// without, max_depth - CT_TAIL_BEFORE_END (0: no tail).
#ifndef CT_TAIL_AFTER_ROULETTE
#define CT_TAIL_AFTER_ROULETTE 0
#endif
#ifndef CT_TAIL_BEFORE_END
#define CT_TAIL_BEFORE_END 0
#endif
constexpr int kTailAfterRoulette = CT_TAIL_AFTER_ROULETTE;
constexpr int kTailBeforeEnd = CT_TAIL_BEFORE_END;
static_assert(kLanes >= 2, "at least two lanes");

// One of the pipelines the chunk driver rotates through.  Each lane
// has its own queue views (a half of the context's path buffers), counters,
// chunk tables, streams, events and captured bounce loop, so the chunks of
// the two lanes run concurrently and fill each other's launch tails.
struct Lane {
    ctd::Queues q{};
    hipStream_t stream = nullptr;
    hipEvent_t resolved = nullptr;   // after this lane's latest resolve
    unsigned long long* wstats = nullptr;
    ctd::ViewDesc* d_views = nullptr;
    uint32_t d_views_cap = 0;
    ctd::Segment* d_segs = nullptr;
    uint32_t d_segs_cap = 0;
    ctd::ChunkHdr* d_hdr = nullptr;
    uint32_t* h_read = nullptr;      // pinned read-back: counters, error flag, work counters
    uint8_t* h_stage = nullptr;      // pinned upload staging of the chunk tables
    size_t h_stage_bytes = 0;
    // captured bounce loop
    hipGraphExec_t graph_exec = nullptr;
    std::string graph_key;
    hipEvent_t ev[2 * (3 * ctd::kMaxBounce + 8)] = {};
    int n_events = 0;
    uint32_t gen = 0;  // generation of the lane's latest chunk (ChunkHdr::gen)
    // the chunk in flight (launched, not yet finished)
    bool pending = false;
    uint32_t n_paths = 0;
    int max_depth = 0;
    bool timed = false, count = false;
    bool nee = false;  // the chunk ends with a launch for the last bounce's shadow rays
    int tail = -1;     // the chunk's k_tail bounce, or -1
};
// h_read layout (uint32 words)
constexpr int kReadCounters = 2 * (ctd::kMaxBounce + 1) + 2;  // per-bounce counts, then k_tail's two
constexpr int kReadErr = kReadCounters;
constexpr int kReadWork = kReadCounters + 2;  // 4 x uint64, 8-byte aligned
constexpr int kReadWords = kReadWork + 8;

struct ct_ctx {
    int device = 0;
    Lane ln[kLanes];
    std::string err;
    std::string name;
    int num_cus = 0;
    uint64_t budget = 0, used = 0;
    std::vector<std::pair<void*, size_t>> allocs;

    ctk::LaunchShape shape;
    uint32_t chunk_default = 1u << 22;  // paths per chunk

    // scene
    bool has_scene = false;
    bool has_bvh2 = false, has_bvh4 = false;
    uint32_t wide_stack_bound = 0;  // worst-case stack of the wide tree (fast mode)
    ctd::DevScene sc{};
    std::vector<void*> scene_bufs;
    std::vector<ct_material> h_mats;
    std::vector<uint32_t> h_mat_idx, h_light_idx;
    std::vector<float> h_light_rows;  // 20 floats per light

    // plain frame accumulator
    int width = 0, height = 0;
    int32_t samples = 0;
    float4* accum = nullptr;
    void* image = nullptr;  // tonemapped image of ct_read_image[_rgba8], kept between calls
    size_t image_bytes = 0;

    // path state: buffers of kLanes * lane_cap slots, lane k uses [k * lane_cap, (k + 1) * lane_cap)
    uint32_t lane_cap = 0;
    std::vector<void*> path_bufs;
    uint32_t spill_threads = 0;

    // this device's copy of the allocation-free entry point's scratch array
    // (its device address; runtime/noalloc.hip)
    void* scratch = nullptr;
};

// Return the failure of a HIP call as CT_ERR_DEVICE on context `c`.
#define HIPCHK(expr, what)                                       \
    do {                                                         \
        hipError_t e_ = (expr);                                  \
        if (e_ != hipSuccess) return ctr::hip_fail(c, e_, what); \
    } while (0)

namespace ctr {

int fail(ct_ctx* c, int code, const std::string& msg);
int hip_fail(ct_ctx* c, hipError_t e, const char* what);
int dalloc(ct_ctx* c, void** p, size_t bytes, const char* what);
void dfree(ct_ctx* c, void* p);

int validate_opts(ct_ctx* c, const ct_frame_opts& o);
// Option checks of validate_opts without touching the context's error string
// (for callers on other threads than the context's owner); *why gets the
// message of a failure.
int check_opts(const ct_ctx* c, const ct_frame_opts& o, const char** why);
// Finite position, forward and up, non-zero forward, plane distance > 0.
bool camera_ok(const ct_camera* cam);
// Device bytes of one path slot of the path-state buffers (every lane holds
// lane_cap slots): ray origin, direction and throughput (two sets), path
// words (two), hit, shadow origin, direction and colour (two sets), path
// radiance and the radiance terms (two sets each of shade's and connect's).
constexpr size_t kPathSlotBytes = 6 * 16 + 2 * 4 + 8 + 6 * 16 + 16 + 4 * 16;
// Path buffers for chunks of up to cap paths on each lane.
int ensure_paths(ct_ctx* c, uint32_t cap);
int ensure_tables(ct_ctx* c, Lane& L, uint32_t nviews, uint32_t nsegs);
void drop_graphs(ct_ctx* c);
// Wait for every lane's stream.
void sync_lanes(ct_ctx* c);
ctd::ViewDesc make_view(const ct_camera& cam, int width, int height, int spp, uint32_t frame_base,
                        uint64_t accum_offset);
// Stride of the pixel-strip order for n full strips (see ctd::pixel_of).
uint32_t strip_stride(uint32_t n);

// Split a list of (view, first_path, n) runs into chunks of at most cap paths
// and run them in order, alternating between the two lanes (a workload that
// fits one chunk is split in two so both lanes work on it).  Work the caller
// queued on lane 0's stream before the call is ordered before all of it; both
// lanes are idle when it returns.  Accumulator updates keep chunk order.
struct Run {
    uint32_t view, first_path, n;
};
// How run_paths runs the chunks: kRunTimed records an event pair around every
// kernel (per-stage times in ChunkStats), kRunCount counts node / triangle
// tests, kRunGraph replays each lane's captured bounce loop, kRunSerial
// runs one chunk at a time (nothing overlaps: per-kernel times are the
// kernels' own, for measurement).
enum RunFlags { kRunTimed = 1, kRunCount = 2, kRunGraph = 4, kRunSerial = 8 };
int run_paths(ct_ctx* c, const std::vector<ctd::ViewDesc>& views, const std::vector<Run>& runs, uint32_t cap,
              float4* accum, const ct_frame_opts& o, int flags, ChunkStats* st);

}  // namespace ctr

namespace ctr {
// Trees and triangle records built on the device (allocated with dalloc;
// the caller owns them, also when the build fails part-way).
struct DeviceTrees {
    float4* tris = nullptr;
    ct_bvh_node* nodes2 = nullptr;
    ct_mbvh_node* nodes4 = nullptr;
    uint32_t n2 = 0, n4 = 0;
};
// Build the binary and quad trees of v's triangles on lane 0's stream; the
// result equals the host build byte for byte.
int device_build(ct_ctx* c, const ct_scene_view* v, DeviceTrees* out, ct_build_stats* st);
}  // namespace ctr

namespace ctr {
// Look up the device address of this device's scratch array (c->device
// current); once per device, kept for the process.
int scratch_init(ct_ctx* c);
}  // namespace ctr

// Set the launch shape (block, LDS stack depth) and size the persistent grid.
extern "C" void ctr_apply_shape(ct_ctx* c, int block, int lds_depth);
