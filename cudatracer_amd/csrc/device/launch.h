// Host-side launchers of the wavefront kernels (kernels.hip).
#pragma once

#include <algorithm>

#include "../../../include/cudatracer.h"
#include "dev_common.h"

namespace ctk {

// Launch geometry chosen by the runtime (and the autotuner).
struct LaunchShape {
    int block = 256;          // threads per block for extend / connect
    int lds_depth = 16;       // per-lane LDS stack entries
    int trace_blocks = 1024;  // largest persistent grid at full fill (sizes the stack spill area)
    int tail_blocks = 1024;   // k_tail's resident grid (<= trace_blocks)
    int grid2 = 1024, grid4 = 1024, grid8 = 1024, gridc = 1024;  // persistent grids per tree (c: compact quad)
    // Eighths of the resident blocks that the frame loop's persistent
    // traversal grids (grid2 ... gridc) fill.  Below 8 the other lane's shade
    // kernel runs beside a traversal launch instead of waiting for its waves
    // to retire: 5 gave the shortest frames both on the 259K-triangle atrium
    // (cache-resident; +4% over 8) and on the 7.24M-triangle sphere (bound by
    // memory latency; +1%).  The ray-level launches always fill every slot
    // (full2 ...).
    int trace_fill = 5;
    int full2 = 1024, full4 = 1024, full8 = 1024, fullc = 1024;
    int shade_blocks = 1024;  // grid-stride grid of shade
    int max_blocks = 65536;   // cap for the simple grid-stride kernels
};

// Persistent grids of the traversal kernels of each tree for ls.block and
// ls.lds_depth (resident blocks per CU x CUs); false if a kernel cannot run
// with that shape.
bool set_trace_grids(LaunchShape& ls, int num_cus);

hipError_t launch_setup(uint32_t* counters, const ctd::ChunkHdr* hdr, hipStream_t s);
// Counters for a ray-level launch of n rays (error flags, 16 words, cleared).
hipError_t launch_setup_rays(uint32_t* counters, uint32_t* err, uint32_t n, hipStream_t s);
hipError_t launch_generate(const LaunchShape& ls, const ctd::ViewDesc* views, const ctd::Segment* segs,
                           const ctd::ChunkHdr* hdr, uint32_t max_n, const ctd::Queues& q, hipStream_t s);
// Closest hit for the rays [0, active[bounce]) of q.ray_o/ray_d[sel] into
// q.hit (the ray-level entry points).
hipError_t launch_extend(const LaunchShape& ls, const ctd::DevScene& sc, const ctd::TraceOpts& to,
                         const ctd::Queues& q, int sel, int bounce, hipStream_t s);
// The frame loop's traversal of bounce b: closest hit for the extension rays
// of bounce b and any hit for the shadow rays of bounce b - 1 in one launch
// (bounce = max_depth: the last bounce's shadow rays).  wstats (may be null)
// accumulates closest-hit node / triangle tests and any-hit node / triangle
// tests.
hipError_t launch_trace(const LaunchShape& ls, const ctd::DevScene& sc, const ctd::TraceOpts& to,
                        const ctd::Queues& q, const ctd::ChunkHdr* hdr, int sel, int bounce,
                        unsigned long long* wstats, hipStream_t s);
// The roulette tail of a chunk from bounce bounce0 on, after trace(bounce0):
// each lane runs the remaining bounces of one path of bounce0's extension
// queue (k_tail).  Grid: ls.tail_blocks blocks of ls.block threads, the
// blocks that fit the GPU at once (within the stack spill area's
// trace_blocks).
hipError_t launch_tail(const LaunchShape& ls, const ctd::DevScene& sc, const ctd::TraceOpts& to,
                       const ctd::ShadeOpts& so, const ctd::ViewDesc* views, const ctd::Segment* segs,
                       const ctd::ChunkHdr* hdr, const ctd::Queues& q, int sel, int bounce0, hipStream_t s);
hipError_t launch_shade(const LaunchShape& ls, const ctd::DevScene& sc, const ctd::ShadeOpts& so,
                        const ctd::ViewDesc* views, const ctd::Segment* segs, const ctd::ChunkHdr* hdr,
                        const ctd::Queues& q, int sel, int bounce, hipStream_t s);
hipError_t launch_resolve(const LaunchShape& ls, const ctd::ViewDesc* views, const ctd::Segment* segs,
                          const ctd::ChunkHdr* hdr, uint32_t max_tasks, const ctd::Queues& q, float4* accum,
                          hipStream_t s);
hipError_t launch_tonemap(const float4* accum, uint32_t n, float samples, float4* out, hipStream_t s);
hipError_t launch_tonemap8(const float4* accum, uint32_t n, float samples, uint32_t* out, hipStream_t s);
// The fast node test's copy of a quad tree (8 float4 per node), its compact
// layout (4 uint4 per node) and the tree's plane / slot flags (*flags
// cleared first).
hipError_t launch_prepare_quad(const float4* nodes4, uint32_t n4, float4* fast, uint4* compact, uint32_t* flags,
                               hipStream_t s);
hipError_t launch_shadow_flags(const LaunchShape& ls, const ctd::DevScene& sc, const ctd::TraceOpts& to,
                               const ctd::Queues& q, uint32_t n, uint8_t* out, hipStream_t s);

}  // namespace ctk
