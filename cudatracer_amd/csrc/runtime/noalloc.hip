// The allocation-free closest-hit entry point (ct_trace_rays_device_noalloc).
//
// Callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc) get the closest-hit
// kernel of ct_trace_rays_device with every buffer beyond their rays and
// hits in one file-scope device array, `scratch`: the work counters and
// error flags (kScratchCtl bytes) and the traversal stack's spill region
// (kScratchSpill bytes).  Each GPU the library runs on holds its own copy of
// the array, at its own address, until the process ends; ct_ctx_create looks
// that address up once per device (hipGetSymbolAddress) and every later use
// goes through it.  All calls on one GPU share its array: calls that overlap
// on one GPU (two contexts, two threads) race and must be serialised by the
// caller.  The setup kernel of each call clears what must start at zero
// (counters, error flags), so a call does not depend on the previous one.
#include <mutex>

#include "runtime.h"

namespace {

constexpr size_t kScratchCtl = 64u << 10;     // counters (kCounterWords words) + error flags
constexpr size_t kScratchSpill = 32u << 20;   // traversal stack spill
constexpr size_t kScratchBytes = kScratchCtl + kScratchSpill;
constexpr size_t kScratchErr = kScratchCtl - 256;  // error flags: the last 256 bytes of the control part
static_assert((size_t)ctd::kCounterWords * 4 <= kScratchErr, "counters do not fit the control part");
// Stack bound of the allocation-free launch (LDS + spill).  The reference's
// traversal stack holds 32 entries (MBVHNode.cuh:128); 48 covers every tree
// of the configuration scenes (deepest quad tree: 16 levels) with room, and
// deeper stacks are reported as CT_ERR_STACK_OVERFLOW, not undefined.
constexpr int kNoAllocStack = 48;
constexpr int kMaxDevices = 64;

std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

}  // namespace

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace ctr {

// With c->device current: the device address of this device's `scratch`
// (looked up on the first context of the device, kept for the process).
int scratch_init(ct_ctx* c) {
    if (c->device < 0 || c->device >= kMaxDevices) return fail(c, CT_ERR_INVALID_ARG, "device ordinal too large");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[c->device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return hip_fail(c, e != hipSuccess ? e : hipErrorInvalidSymbol, "scratch address");
        g_scratch[c->device] = p;
    }
    c->scratch = g_scratch[c->device];
    return CT_OK;
}

}  // namespace ctr

using namespace ctr;

extern "C" int ct_trace_rays_device_noalloc(ct_ctx* c, const void* d_rays_o, const void* d_rays_d, uint32_t n,
                                            const ct_frame_opts* opts, void* d_hits, float* kernel_ms) {
    if (!c || (n && (!d_rays_o || !d_rays_d || !d_hits))) return CT_ERR_INVALID_ARG;
    ct_frame_opts o;
    if (opts) o = *opts;
    else ct_frame_opts_default(&o);
    if (o.spp < 1) o.spp = 1;
    if (o.max_depth < 1) o.max_depth = 1;
    int r = validate_opts(c, o);
    if (r != CT_OK) return r;
    if (!c->scratch) return fail(c, CT_ERR_DEVICE, "no scratch array on this device");
    if (kernel_ms) *kernel_ms = 0.0f;
    if (n == 0) return CT_OK;
    (void)hipSetDevice(c->device);
    // the persistent grid, cut to the blocks whose spill regions fit the array
    ctk::LaunchShape ls = c->shape;
    const int lds = std::min(ls.lds_depth, kNoAllocStack);
    const size_t per_block = (size_t)(kNoAllocStack - lds) * (size_t)ls.block * 4;
    const size_t fit = per_block ? kScratchSpill / per_block : (size_t)1 << 30;
    if (fit < (size_t)ctd::kFetchRanges)
        return fail(c, CT_ERR_INVALID_ARG, "launch shape (block size x stack depth) needs more than the 32 MiB "
                                           "scratch array; use a smaller block or ct_trace_rays_device");
    const int cap = (int)std::min<size_t>(fit, (size_t)1 << 30);
    ls.full2 = std::min(ls.full2, cap);
    ls.full4 = std::min(ls.full4, cap);
    ls.full8 = std::min(ls.full8, cap);
    ls.fullc = std::min(ls.fullc, cap);
    uint8_t* base = static_cast<uint8_t*>(c->scratch);  // the device address, never the symbol itself
    ctd::Queues q{};
    q.ray_o[0] = (float4*)d_rays_o;
    q.ray_d[0] = (float4*)d_rays_d;
    q.hit = (float2*)d_hits;
    q.counters = reinterpret_cast<uint32_t*>(base);
    q.err = reinterpret_cast<uint32_t*>(base + kScratchErr);
    q.spill = reinterpret_cast<uint32_t*>(base + kScratchCtl);
    q.max_stack = kNoAllocStack;
    const ctd::TraceOpts to{o.accel, o.traversal, o.triangle_epsilon};
    hipStream_t s = c->ln[0].stream;
    HIPCHK(ctk::launch_setup_rays(q.counters, q.err, n, s), "setup");
    HIPCHK(hipEventRecord(c->ln[0].ev[0], s), "event");
    HIPCHK(ctk::launch_extend(ls, c->sc, to, q, 0, 0, s), "extend");
    HIPCHK(hipEventRecord(c->ln[0].ev[1], s), "event");
    uint32_t errflag = 0;
    HIPCHK(hipMemcpyAsync(&errflag, q.err, 4, hipMemcpyDeviceToHost, s), "flags");
    HIPCHK(hipStreamSynchronize(s), "trace rays");
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, c->ln[0].ev[0], c->ln[0].ev[1]), "elapsed");
    if (kernel_ms) *kernel_ms = ms;
    if (errflag)
        return fail(c, CT_ERR_STACK_OVERFLOW, "traversal stack bound (48 entries without allocation) exceeded; "
                                              "the rays that needed more are reported as misses");
    return CT_OK;
}
