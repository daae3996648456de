// Wavefront path-tracing kernels for gfx950 (MI355X).
//
// Reference stages (src/CUDA/Kernel.cu):        here:
//   setGlobals          :75-86                  k_setup (all bounce counters zeroed once per chunk)
//   generatePrimaryRays :208-234                k_generate
//   extend              :236-245                k_trace    (persistent, wave-fetched; bounce b's
//   connect             :247-256                            extension rays with bounce b - 1's
//                                                           shadow rays in one launch)
//   shade               :88-206                 k_shade    (ballot + block prefix-sum queue append)
//   accumulation        :100,134,255            k_resolve  (per-path radiance summed in sample order)
//   draw_framebuffer    :258-270                k_tonemap
//
// Path state is structure-of-arrays (float4 rows, 16-byte coalesced loads).
// Queue lengths live in device counters that the next kernel reads, so the
// bounce loop needs no host read-back (the reference copies two counters to
// the host after every bounce, Kernel.cu:309-310,335-336).
#include "launch.h"

#include "bsdf.h"
#include "traverse.h"

namespace ctd {

CT_DEV int find_seg(const Segment* __restrict__ segs, int nseg, uint32_t slot) {
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (segs[mid].chunk_offset <= slot) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

struct SlotInfo {
    uint32_t pixel, frame, view;
};
CT_DEV SlotInfo decode_slot(const ViewDesc* __restrict__ views, const Segment* __restrict__ segs, int nseg,
                            uint32_t slot) {
    const Segment sg = segs[find_seg(segs, nseg, slot)];
    const ViewDesc& v = views[sg.view];
    const uint32_t local = sg.first_path + (slot - sg.chunk_offset);
    const uint32_t spp = (uint32_t)v.spp;
    const uint32_t ord = local / spp;
    SlotInfo si;
    si.pixel = pixel_of(ord, (uint32_t)v.width, (uint32_t)v.height, v.strip_mul);
    si.frame = v.frame_base + (local - ord * spp);
    si.view = sg.view;
    return si;
}

// ---------------------------------------------------------------- setup --
__global__ void k_setup(uint32_t* counters, const ChunkHdr* hdr) {
    for (int i = threadIdx.x; i < kCounterWords; i += blockDim.x) counters[i] = 0u;
    __syncthreads();
    if (threadIdx.x == 0) *cnt_active(counters, 0) = hdr->n;
}

// The same for a ray-level launch of n rays, with the size as an argument
// (no header in device memory) and the error flags cleared too.
__global__ void k_setup_rays(uint32_t* counters, uint32_t* err, uint32_t n) {
    for (int i = threadIdx.x; i < kCounterWords; i += blockDim.x) counters[i] = 0u;
    if (threadIdx.x < 16) err[threadIdx.x] = 0u;
    __syncthreads();
    if (threadIdx.x == 0) *cnt_active(counters, 0) = n;
}

// ------------------------------------------------------------- generate --
// generatePrimaryRays + Ray::generate (Kernel.cu:208-234, Ray.cuh:53-61,85-95)
__global__ void k_generate(const ViewDesc* __restrict__ views, const Segment* __restrict__ segs,
                           const ChunkHdr* hdr, Queues q) {
    const uint32_t n = hdr->n;
    const int nseg = (int)hdr->nseg;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const Segment sg = segs[find_seg(segs, nseg, i)];
        const ViewDesc& v = views[sg.view];
        const uint32_t local = sg.first_path + (i - sg.chunk_offset);
        const uint32_t spp = (uint32_t)v.spp;
        const uint32_t ord = local / spp;
        const uint32_t frame = v.frame_base + (local - ord * spp);
        const uint32_t pixel = pixel_of(ord, (uint32_t)v.width, (uint32_t)v.height, v.strip_mul);
        const uint32_t x = pixel % (uint32_t)v.width, y = pixel / (uint32_t)v.width;
        uint32_t seed = ct_primary_seed(pixel, frame);
        const float jx = ct_rand_float(&seed);
        const float jy = ct_rand_float(&seed);
        const float px = (float)x + jx - 0.5f;
        const float py = (float)y + jy - 0.5f;
        const float sx = 2.0f * (px * v.invw) - 1.0f;
        const float sy = 1.0f - 2.0f * (py * v.invh);
        const f3 w = mk3(v.w[0], v.w[1], v.w[2]);
        const f3 hor = mk3(v.hor[0], v.hor[1], v.hor[2]);
        const f3 ver = mk3(v.ver[0], v.ver[1], v.ver[2]);
        const f3 d = normalize(w + hor * sx + ver * sy);
        q.ray_o[0][i] = make_float4(v.pos[0], v.pos[1], v.pos[2], 0.0f);
        q.ray_d[0][i] = make_float4(d.x, d.y, d.z, 0.0f);
        q.thr[0][i] = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
        q.path[0][i] = i | (LT_SPECULAR << 30);
        q.lpath[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}

// --------------------------------------------------------------- extend --
// Persistent closest-hit over a queue of rays (see trace_persistent): the
// ray-level entry points.
template <int ACCEL>
CT_DEV void extend_impl(const DevScene& sc, const TraceOpts& to, const Queues& q, int sel, int bounce,
                        LaneStack& st) {
    const uint32_t n = *cnt_active(q.counters, bounce);
    const float4* __restrict__ ro = q.ray_o[sel];
    const float4* __restrict__ rd = q.ray_d[sel];
    float2* hit = q.hit;
    TravCount tc{0u, 0u, 0u, 0u};
    trace_persistent<ACCEL, kClosest, false>(
        sc, to.teps, n, 0u, cnt_fetch_ext(q.counters, bounce), st, tc, q.err,
        [&](uint32_t i, bool) {
            const float4 o4 = ro[i], d4 = rd[i];
            return RayIn{xyz(o4), xyz(d4), 1e34f};
        },
        [&](uint32_t i, bool, float t, int prim) {
            if (prim == -2) { t = 1e34f; prim = -1; }
            hit[i] = make_float2(t, __int_as_float(prim));
        });
}

// GFX9 buffer resource word 3 (raw 32-bit access).
constexpr int kBufferWord3 = 0x00020000;

CT_DEV LaneStack make_stack(const Queues& q, int lds_depth) {
    LaneStack st;
    // the block size comes from a 16-bit vector load of the dispatch packet:
    // made scalar here, so the stride and the spill resource stay in SGPRs
    const uint32_t bdim = __builtin_amdgcn_readfirstlane(blockDim.x);
    st.stride4 = 4u * bdim;
    // positions [0, lds_depth) in LDS, [lds_depth, max_stack) in the spill
    // region (none when the LDS part holds them all)
    const int depth = lds_depth < q.max_stack ? lds_depth : q.max_stack;
    st.lds_end = (uint32_t)depth * st.stride4;
    st.max_end = (uint32_t)q.max_stack * st.stride4;
    const uint32_t region = (uint32_t)(q.max_stack - depth) * st.stride4;
    char* base = reinterpret_cast<char*>(q.spill) + (size_t)blockIdx.x * region;
    st.spill = __builtin_amdgcn_make_buffer_rsrc(base, 0, (int)region, kBufferWord3);
    st.reset();
    return st;
}

// One kernel per tree (and counting mode), so each gets its own register
// budget.  Eight waves per SIMD (64 VGPRs) is the occupancy the traversal
// needs to hide its memory latency; asked for explicitly, the scheduler
// reaches it without spills (the counting copies spill a register).
template <int TREE>
__global__ void __attribute__((amdgpu_waves_per_eu(8)))
k_extend(DevScene sc, TraceOpts to, Queues q, int sel, int bounce, int lds_depth) {
    LaneStack st = make_stack(q, lds_depth);
    extend_impl<TREE>(sc, to, q, sel, bounce, st);
}

// ---------------------------------------------------------------- trace --
// The frame loop's traversal launch of bounce b: the extension rays of
// bounce b (extend, Kernel.cu:236-245: closest hit into q.hit) and the
// shadow rays of bounce b - 1 (connect, Kernel.cu:247-256: any hit), one
// persistent launch for both queues.  An unoccluded shadow ray's colour is
// its path's radiance term of bounce b - 1, stored for shade(b + 1) or
// k_resolve to add in order.  Bounce 0 has no shadow queue; the launch after
// the last bounce (b = max_depth) has no extension rays.
template <int ACCEL, bool COUNT>
CT_DEV void trace_impl(const DevScene& sc, const TraceOpts& to, const Queues& q, const ChunkHdr* hdr, int sel,
                       int bounce, LaneStack& st, TravCount& tc) {
    const uint32_t n = *cnt_active(q.counters, bounce);
    const int sb = bounce - 1;  // the shadow rays' bounce
    const uint32_t n2 = bounce > 0 ? *cnt_shadow(q.counters, sb) : 0u;
    const uint32_t stamp = term_stamp(hdr->gen, sb);
    const float4* __restrict__ ro = q.ray_o[sel];
    const float4* __restrict__ rd = q.ray_d[sel];
    const float4* __restrict__ so = q.sh_o[sb & 1];
    const float4* __restrict__ sd = q.sh_d[sb & 1];
    trace_persistent<ACCEL, kDual, COUNT>(
        sc, to.teps, n, n2, cnt_fetch_ext(q.counters, bounce), st, tc, q.err,
        [&](uint32_t i, bool any) {
            const float4 o4 = any ? so[i] : ro[i], d4 = any ? sd[i] : rd[i];
            return RayIn{xyz(o4), xyz(d4), any ? o4.w : 1e34f};
        },
        [&](uint32_t i, bool any, float t, int prim) {
            if (!any) {
                if (prim == -2) { t = 1e34f; prim = -1; }
                q.hit[i] = make_float2(t, __int_as_float(prim));
                return;
            }
            if (prim != -1) return;  // occluded (or stack overflow, flagged)
            const uint32_t slot = __float_as_uint(sd[i].w);
            const float4 c = q.sh_c[sb & 1][i];
            q.nbuf[sb & 1][slot] = make_float4(c.x, c.y, c.z, __uint_as_float(stamp));
        });
}

template <int TREE, bool COUNT>
__global__ void __attribute__((amdgpu_waves_per_eu(8)))
k_trace(DevScene sc, TraceOpts to, Queues q, const ChunkHdr* hdr, int sel, int bounce, int lds_depth,
        unsigned long long* wstats) {
    LaneStack st = make_stack(q, lds_depth);
    TravCount tc{0u, 0u, 0u, 0u};
    trace_impl<TREE, COUNT>(sc, to, q, hdr, sel, bounce, st, tc);
    if (COUNT) {
        atomicAdd(&wstats[0], (unsigned long long)tc.nodes);
        atomicAdd(&wstats[1], (unsigned long long)tc.tris);
        atomicAdd(&wstats[2], (unsigned long long)tc.snodes);
        atomicAdd(&wstats[3], (unsigned long long)tc.stris);
    }
}

// ---------------------------------------------------------------- shade --
constexpr int kShadeBlock = 256;
constexpr uint32_t kSortKeys = 16;

// Material key of queue entry i for the shading sort: 0 miss, 1 + material
// type for hits, kSortKeys - 1 past the end of the queue.
CT_DEV uint32_t material_key(const DevScene& sc, const Queues& q, uint32_t i, uint32_t n) {
    if (i >= n) return kSortKeys - 1;
    const float2 h = q.hit[i];
    if (!(h.x < 1e34f)) return 0;
    const uint32_t mi = __float_as_uint(sc.sh_prim[kPrimRows * (size_t)__float_as_int(h.y)].w);
    const int mtype = __float_as_int(sc.mats[2 * (size_t)mi + 1].x);
    return 1u + (uint32_t)min(max(mtype, 0), (int)kSortKeys - 3);
}

// Block-local counting sort of the block's kShadeBlock queue entries by
// material key (LDS histogram, one-wave scan, scatter of indices), so each
// wave shades mostly one BSDF.  Returns the entry this thread shades.  The
// order inside a key is arbitrary: a path's radiance and the rays it emits do
// not depend on which lane shades it.
CT_DEV uint32_t sorted_entry(const DevScene& sc, const Queues& q, uint32_t first, uint32_t n, uint32_t* hist,
                             uint32_t* perm) {
    __syncthreads();  // previous iteration's readers are done
    if (threadIdx.x < kSortKeys) hist[threadIdx.x] = 0u;
    __syncthreads();
    const uint32_t mine = first + threadIdx.x;
    const uint32_t key = material_key(sc, q, mine, n);
    const uint32_t rank = atomicAdd(&hist[key], 1u);
    __syncthreads();
    if (threadIdx.x < 64) {
        const uint32_t c = threadIdx.x < kSortKeys ? hist[threadIdx.x] : 0u;
        uint32_t v = c;
        for (int d = 1; d < (int)kSortKeys; d <<= 1) {
            const uint32_t u = __shfl_up(v, d);
            if ((int)lane_id() >= d) v += u;
        }
        if (threadIdx.x < kSortKeys) hist[threadIdx.x] = v - c;
    }
    __syncthreads();
    perm[hist[key] + rank] = mine;
    __syncthreads();
    return perm[threadIdx.x];
}

// Appends to the extension and shadow queues for a block of kShadeBlock
// threads: LDS prefix over the block's 4 waves and one atomic per block and
// queue.  Every thread of the block must call it (two barriers).  One atomic
// per wave instead, all on the same two counters from every XCD, measured
// 2.2x slower shading (the atomics serialise).  The LDS slots are safe to
// reuse in the next call: a wave writes only its own count slot before the
// first barrier, and thread 0 rewrites the shared ones only after it.
CT_DEV void block_append2(bool pe, bool ps, uint32_t* ce, uint32_t* cs, uint32_t* lds, uint32_t& je, uint32_t& js) {
    const uint64_t me = __ballot(pe ? 1 : 0), ms = __ballot(ps ? 1 : 0);
    const uint32_t w = threadIdx.x >> 6, lane = lane_id();
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64u - lane));
    if (lane == 0) {
        lds[w] = (uint32_t)__popcll(me);
        lds[4 + w] = (uint32_t)__popcll(ms);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t te = 0, ts = 0;
        for (int k = 0; k < 4; ++k) {
            const uint32_t a = lds[k], b = lds[4 + k];
            lds[k] = te;
            lds[4 + k] = ts;
            te += a;
            ts += b;
        }
        lds[8] = te ? atomicAdd(ce, te) : 0u;
        lds[9] = ts ? atomicAdd(cs, ts) : 0u;
    }
    __syncthreads();
    je = lds[8] + lds[w] + (uint32_t)__popcll(me & lt);
    js = lds[9] + lds[4 + w] + (uint32_t)__popcll(ms & lt);
}

// One queue entry of shade (Kernel.cu:88-206; the oracle's path_sample
// restates the same steps): the hit (t, prim) of the ray (org0, dir) of path
// slot `slot` at `bounce`, with throughput thr, the last bounce's pdf and
// type.  Out: the hit's radiance term (add, when contrib), the light
// sample's shadow ray (push_sh) and the continuation ray (push_ext).  k_shade
// and k_tail both shade through it.
struct ShadeOut {
    f3 add, s_o, s_d, s_c, e_o, e_d, e_thr;
    float s_tmax, e_pdf;
    uint32_t e_type;
    bool contrib, push_sh, push_ext;
};

CT_DEV void shade_one(const DevScene& sc, const ShadeOpts& so, const ViewDesc* __restrict__ views,
                      const Segment* __restrict__ segs, int nseg, int bounce, uint32_t slot, uint32_t lastType,
                      const f3 org0, f3 dir, f3 thr, const float lastPdf, const float t, const int prim,
                      ShadeOut& out) {
    out.add = out.s_o = out.s_d = out.s_c = out.e_o = out.e_d = out.e_thr = mk3(0, 0, 0);
    out.s_tmax = out.e_pdf = 0.0f;
    out.e_type = 0;
    out.contrib = out.push_sh = out.push_ext = false;
    f3 &add = out.add, &s_o = out.s_o, &s_d = out.s_d, &s_c = out.s_c, &e_o = out.e_o, &e_d = out.e_d,
       &e_thr = out.e_thr;
    float &s_tmax = out.s_tmax, &e_pdf = out.e_pdf;
    uint32_t& e_type = out.e_type;
    bool &contrib = out.contrib, &push_sh = out.push_sh, &push_ext = out.push_ext;
    if (!(t < 1e34f)) {
        if (sc.sky_enabled) {
            f3 sky = mk3(sc.sky_r, sc.sky_g, sc.sky_b);
            if (sc.sky_tex >= 0) {  // Kernel.cu:96-100
                float u, v;
                ct_sky_uv(dir.x, dir.y, dir.z, &u, &v);
                const float4 c = tex_fetch(sc.texels, sc.texels8, sc.tex_info, sc.sky_tex, u, v);
                sky = mk3(c.x, c.y, c.z);
            }
            add = (thr / lastPdf) * sky;
            contrib = true;
        }
    } else {
        const float4* rec = sc.sh_prim + kPrimRows * (size_t)prim;
        const float4 P0 = rec[0], P1 = rec[1], P2 = rec[2];
        const uint32_t mi = __float_as_uint(P0.w);
        const float4 M0 = sc.mats[2 * (size_t)mi], M1 = sc.mats[2 * (size_t)mi + 1];
        const int mtype = __float_as_int(M1.x);
        const f3 p0 = xyz(P0), p1 = xyz(P1), p2 = xyz(P2);
        const f3 hitp = org0 + t * dir;
        const float4 N0 = rec[3], N1 = rec[4], N2 = rec[5], NC = rec[6];
        f3 N = xyz(NC);
        const f3 bary = bary_coords(hitp, N, p0, p1, p2);
        f3 iN = bary.x * xyz(N0) + bary.y * xyz(N1) + bary.z * xyz(N2);
        // textures: diffuse colour and normal map (Kernel.cu:106-112,146)
        const uint32_t tpack = __float_as_uint(M1.w);
        const int dtex = (int)(tpack & 0xFFFFu) - 1, ntex = (int)(tpack >> 16) - 1;
        float tu = 0.0f, tv = 0.0f;
        if (dtex >= 0 || ntex >= 0) {
            tu = (bary.x * P1.w + bary.y * N0.w) + bary.z * N2.w;
            tv = (bary.x * P2.w + bary.y * N1.w) + bary.z * NC.w;
        }
        if (ntex >= 0) {
            const float4 c = tex_fetch(sc.texels, sc.texels8, sc.tex_info, ntex, tu, tv);
            iN = sample_to_world(normalize(mk3(c.x, c.y, c.z)), iN);
        }
        if (mtype == MAT_LIGHT) {
            if (so.indirect) {
                const float DdotNL = -dot(dir, iN);
                const f3 Le = xyz(M0);
                f3 c;
                if (lastType == LT_SPECULAR) {
                    c = thr * Le;
                } else {
                    const float lightPdf = light_pdf_scale(t * t / (DdotNL * tri_area(p0, p1, p2)), sc.n_lights, so.light_pdf);
                    c = thr * Le * (1.0f / (lastPdf + lightPdf));
                }
                add = vmax0(c);
                contrib = true;
            }
        } else {
            const SlotInfo si = decode_slot(views, segs, nseg, slot);
            f3 org = hitp;
            thr = vmax0(thr / lastPdf);
            const bool flipped = dot(dir, N) > 0.0f;
            if (flipped) { N = -N; iN = -iN; }
            uint32_t seed = ct_shade_seed(si.pixel, si.frame, ct_path_length_key(bounce));
            const float r3 = ct_rand_float(&seed);
            const float r4 = ct_rand_float(&seed);
            f3 color = xyz(M0);
            if (dtex >= 0) {
                const float4 c = tex_fetch(sc.texels, sc.texels8, sc.tex_info, dtex, tu, tv);
                color = mk3(c.x, c.y, c.z);
            }
            const f3 wiw = -dir;
            f3 wo = mk3(0, 0, 0), f = mk3(0, 0, 0);
            float pdf = 0.0f;
            uint32_t type = LT_DIFFUSE;
            bool valid = false, transmit = false;
            const bool typed = so.bsdf_mode == CT_BSDF_TYPED;
            const float roughness = M0.w;
            const float alpha = gmax(roughness, 1e-3f);
            // `type` is the material's lobe, set before sampling:
            // next-event estimation runs for a diffuse or glossy
            // lobe also when the sample is rejected (the path then
            // ends after it), which keeps the MIS sum unbiased
            if ((!typed && roughness < 0.1f) || (typed && mtype == MAT_SPECULAR)) {
                type = LT_SPECULAR;
                wo = reflect(dir, iN);
                if (dot(N, wo) > 0.0f) {
                    pdf = 1.0f;
                    f = color * (1.0f / ct_fabsf(dot(iN, wo)));
                    valid = true;
                }
            } else if (typed && is_dielectric(mtype)) {
                const float ior = M1.y;
                const float etaI = flipped ? ior : 1.0f;
                const float etaT = flipped ? 1.0f : ior;
                const float cosI = dot(wiw, iN);
                const float F = fresnel_dielectric(cosI, etaT / etaI);
                type = LT_SPECULAR;
                pdf = 1.0f;
                if (r3 < F || F >= 1.0f) {
                    wo = reflect(dir, iN);
                    valid = dot(N, wo) > 0.0f;
                } else {
                    const float eta = etaI / etaT;
                    const float dv = dot(iN, dir);
                    const float k = 1.0f - eta * eta * (1.0f - dv * dv);
                    wo = eta * dir - (eta * dv + ct_sqrtf(k)) * iN;
                    valid = k >= 0.0f && dot(N, wo) < 0.0f;
                    transmit = true;
                }
                if (valid) f = color * (1.0f / ct_fabsf(dot(iN, wo)));
            } else if (typed && is_microfacet(mtype)) {
                type = LT_GLOSSY;
                f3 Tv, Bv;
                tangent_frame(iN, Tv, Bv);
                const f3 wol = mk3(dot(wiw, Tv), dot(wiw, Bv), dot(wiw, iN));
                if (wol.z > 0.0f) {
                    const f3 wm = sample_vndf(mtype, wol, alpha, r3, r4);
                    const f3 wil = reflect(-wol, wm);
                    if (wil.z > 0.0f) {
                        wo = normalize(wil.x * Tv + wil.y * Bv + wil.z * iN);
                        if (dot(N, wo) > 0.0f) {
                            mf_eval(mtype, alpha, color, wol, wil, &f, &pdf);
                            valid = pdf > 0.0f;
                        }
                    }
                }
            } else {
                type = LT_DIFFUSE;
                wo = normalize(tangent_to_world(cos_weighted(r3, r4), iN));
                if (dot(N, wo) > 0.0f) {
                    pdf = gmax(0.0f, dot(wo, iN)) * CT_INV_PI_F;
                    f = color * CT_INV_PI_F;
                    valid = true;
                }
            }
            {
                // next-event estimation (Kernel.cu:161-195)
                if (type != LT_SPECULAR && so.shadow && sc.n_lights > 0) {
                    int light = ct_rand_int_max(&seed, (int)sc.n_lights - 1);
                    if (light > (int)sc.n_lights - 1) light = (int)sc.n_lights - 1;
                    const float r1 = ct_rand_float(&seed);
                    const float r2 = ct_rand_float(&seed);
                    const float4* lr = sc.lights + 5 * (size_t)light;
                    const float4 L0 = lr[0], L1 = lr[1], L2 = lr[2], LN = lr[3], LE = lr[4];
                    const f3 lightPos = random_point(xyz(L0), xyz(L1), xyz(L2), r1, r2);
                    f3 Lv = lightPos - org;
                    const float sqDist = dot(Lv, Lv);
                    const float distance = ct_sqrtf(sqDist);
                    Lv = Lv / distance;
                    const f3 lN = xyz(LN);
                    const float NdotL = dot(iN, Lv);
                    const float LNdotL = -dot(lN, Lv);
                    if (NdotL > 0.0f && LNdotL > 0.0f) {
                        const float ar = LN.w;
                        const f3 emission = xyz(LE);
                        f3 fe;
                        float shadowPdf;
                        if (type == LT_DIFFUSE) {
                            shadowPdf = ct_fabsf(dot(Lv, iN)) * CT_INV_PI_F;
                            fe = color * CT_INV_PI_F;
                        } else {
                            f3 Tv, Bv;
                            tangent_frame(iN, Tv, Bv);
                            const f3 wol = mk3(dot(wiw, Tv), dot(wiw, Bv), dot(wiw, iN));
                            const f3 wil = mk3(dot(Lv, Tv), dot(Lv, Bv), dot(Lv, iN));
                            mf_eval(mtype, alpha, color, wol, wil, &fe, &shadowPdf);
                        }
                        if (shadowPdf > 0.0f) {
                            const float lightPdf = light_pdf_scale(sqDist / (LNdotL * ar), sc.n_lights, so.light_pdf);
                            s_c = thr * emission * fe * (NdotL / (shadowPdf + lightPdf));
                            s_o = org + so.neps * Lv;
                            s_d = Lv;
                            s_tmax = distance - 2.0f * so.deps;
                            push_sh = true;
                        }
                    }
                }
                // continuation (Kernel.cu:197-205); a rejected sample ends the path
                if (valid && !(pdf < 1e-6f || isnanf_(pdf) || isnanf_(thr.x) || isnanf_(thr.y) || isnanf_(thr.z) ||
                      (thr.x <= 0.0f && thr.y <= 0.0f && thr.z <= 0.0f))) {
                    thr = vmax0(thr * f * ct_fabsf(dot(iN, wo)));
                    org = transmit ? org - N * so.neps : org + N * so.neps;
                    bool alive = bounce + 1 < so.max_depth;
                    if (alive && so.rr_start >= 0 && bounce >= so.rr_start) {
                        const f3 tt = thr / pdf;
                        const float qs = gmin(1.0f, gmax(tt.x, gmax(tt.y, tt.z)));
                        const float rr = ct_rand_float(&seed);
                        if (!(rr < qs)) alive = false;
                        else thr = thr / qs;
                    }
                    if (alive) {
                        push_ext = true;
                        e_o = org;
                        e_d = wo;
                        e_thr = thr;
                        e_pdf = pdf;
                        e_type = type;
                    }
                }
            }
        }
    }
}

// shade (Kernel.cu:88-206) for queue entries [0, active[bounce]); see the
// oracle's path_sample for the same steps on the CPU.  SORT: shade in
// block-local material order (ct_frame_opts.sort_by_material).  Register
// budgets, each the largest occupancy without spills: sorted, five waves per
// SIMD (at most 96 VGPRs; 1.3% faster per frame than the four the
// unconstrained allocation leaves, and six spill); unsorted, six (at most 80
// VGPRs; on the 7.24M-triangle 4K scene 0.8% faster per frame than five).
template <bool SORT>
__global__ void __launch_bounds__(kShadeBlock) __attribute__((amdgpu_waves_per_eu(SORT ? 5 : 6)))
k_shade(DevScene sc, ShadeOpts so, const ViewDesc* __restrict__ views, const Segment* __restrict__ segs,
        const ChunkHdr* hdr, Queues q, int sel, int bounce) {
    __shared__ uint32_t s_hist[kSortKeys];
    __shared__ uint32_t s_app[10];
    __shared__ uint32_t s_perm[SORT ? kShadeBlock : 1];
    const uint32_t n = *cnt_active(q.counters, bounce);
    const int nseg = (int)hdr->nseg;
    const uint32_t gen = hdr->gen;
    const int nsel = sel ^ 1;
    const uint32_t stride = gridDim.x * kShadeBlock;
    // block-uniform loop: all threads of a block iterate together
    for (uint32_t first = blockIdx.x * kShadeBlock; first < n; first += stride) {
        const uint32_t i = SORT ? sorted_entry(sc, q, first, n, s_hist, s_perm) : first + threadIdx.x;
        const bool active = i < n;
        bool push_ext = false, push_sh = false;
        f3 e_o = mk3(0, 0, 0), e_d = mk3(0, 0, 0), e_thr = mk3(0, 0, 0);
        float e_pdf = 0.0f;
        uint32_t e_type = 0, slot = 0;
        f3 s_o = mk3(0, 0, 0), s_d = mk3(0, 0, 0), s_c = mk3(0, 0, 0);
        float s_tmax = 0.0f;
        ShadeOut sh;
        if (active) {
            const float4 o4 = q.ray_o[sel][i], d4 = q.ray_d[sel][i], th4 = q.thr[sel][i];
            const uint32_t pw = q.path[sel][i];
            const float2 h = q.hit[i];
            slot = pw & kSlotMask;
            shade_one(sc, so, views, segs, nseg, bounce, slot, pw >> 30, xyz(o4), xyz(d4), xyz(th4), th4.w, h.x,
                      __float_as_int(h.y), sh);
            push_ext = sh.push_ext;
            e_o = sh.e_o;
            e_d = sh.e_d;
            e_thr = sh.e_thr;
            e_pdf = sh.e_pdf;
            e_type = sh.e_type;
            push_sh = sh.push_sh;
            s_o = sh.s_o;
            s_d = sh.s_d;
            s_c = sh.s_c;
            s_tmax = sh.s_tmax;
            // bounce - 2's pair of terms into the radiance (shade's, then
            // connect's), then this bounce's term for shade(bounce + 2)
            if (bounce >= 2) {
                const uint32_t st = term_stamp(gen, bounce - 2);
                const float4 e = q.ebuf[bounce & 1][slot], nn = q.nbuf[bounce & 1][slot];
                const bool he = __float_as_uint(e.w) == st, hn = __float_as_uint(nn.w) == st;
                if (he || hn) {
                    float4 L = q.lpath[slot];
                    if (he) add_term(L, e);
                    if (hn) add_term(L, nn);
                    L.w = __uint_as_float((uint32_t)bounce - 1u);
                    q.lpath[slot] = L;
                }
            }
            if (sh.contrib) q.ebuf[bounce & 1][slot] = term(sh.add, term_stamp(gen, bounce));
        }
        // block-aggregated queue appends (every thread of the block reaches here)
        uint32_t je = 0, js = 0;
        block_append2(push_ext, push_sh, cnt_active(q.counters, bounce + 1), cnt_shadow(q.counters, bounce), s_app, je,
                      js);
        if (push_ext) {
            q.ray_o[nsel][je] = make_float4(e_o.x, e_o.y, e_o.z, 0.0f);
            q.ray_d[nsel][je] = make_float4(e_d.x, e_d.y, e_d.z, 0.0f);
            q.thr[nsel][je] = make_float4(e_thr.x, e_thr.y, e_thr.z, e_pdf);
            q.path[nsel][je] = slot | (e_type << 30);
        }
        if (push_sh) {
            q.sh_o[bounce & 1][js] = make_float4(s_o.x, s_o.y, s_o.z, s_tmax);
            q.sh_d[bounce & 1][js] = make_float4(s_d.x, s_d.y, s_d.z, __uint_as_float(slot));
            q.sh_c[bounce & 1][js] = make_float4(s_c.x, s_c.y, s_c.z, 0.0f);
        }
    }
}

// ----------------------------------------------------------------- tail --
// A chunk's roulette tail in one launch.  Once roulette has thinned the
// paths, a bounce's trace and shade launches hold few rays each and last as
// long as their slowest ray, plus a dependent-launch gap each.  From bounce
// b0 on, k_tail instead gives each pair of lanes one path of bounce b0's
// extension queue (its hit already traced by trace(b0)) and runs the path's
// remaining bounces itself: the even lane shades (shade_one), then traces
// the continuation ray while the odd lane traces the light sample's shadow
// ray, both in one trace_single loop, so a bounce costs the longer of the
// two traversals rather than their sum.  Per path the steps and their
// arithmetic are those of the queue pipeline, so the results are the same;
// the radiance terms are added in the same order (bounce b0 - 2's and
// b0 - 1's pairs first, which shade(b0) and shade(b0 + 1) would fold, then
// each bounce's shade term and light sample in turn), and the path's sum is
// marked fully folded for k_resolve.  The rays it traces are counted in
// cnt_tail_ext / cnt_tail_shadow.  Both lanes of a pair make the same loop
// decisions (exchanged with lane shuffles), so each shuffle reads an active
// partner.
template <int ACCEL>
__global__ void __launch_bounds__(1024)
k_tail(DevScene sc, TraceOpts to, ShadeOpts so, const ViewDesc* __restrict__ views, const Segment* __restrict__ segs,
       const ChunkHdr* hdr, Queues q, int sel, int bounce0, int lds_depth) {
    LaneStack st = make_stack(q, lds_depth);
    const uint32_t n = *cnt_active(q.counters, bounce0);
    const int nseg = (int)hdr->nseg;
    const uint32_t gen = hdr->gen;
    const bool shader = (__lane_id() & 1u) == 0u;  // the even lane of a pair
    const int partner = (int)(__lane_id() ^ 1u);
    uint32_t n_ext = 0, n_sh = 0;
    const uint32_t pairs = (gridDim.x * blockDim.x) / 2u;
    for (uint32_t i = (blockIdx.x * blockDim.x + threadIdx.x) / 2u; i < n; i += pairs) {
        uint32_t slot = 0, type = 0;
        float4 L = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        f3 org = mk3(0, 0, 0), dir = mk3(0, 0, 0), thr = mk3(0, 0, 0);
        float pdf = 0.0f, t = 0.0f;
        int prim = -1;
        if (shader) {
            const float4 o4 = q.ray_o[sel][i], d4 = q.ray_d[sel][i], th4 = q.thr[sel][i];
            const uint32_t pw = q.path[sel][i];
            const float2 h = q.hit[i];
            slot = pw & kSlotMask;
            type = pw >> 30;
            L = q.lpath[slot];
            for (int b = bounce0 - 2; b < bounce0; ++b) {
                if (b < 0) continue;
                const uint32_t stamp = term_stamp(gen, b);
                const float4 e = q.ebuf[b & 1][slot], nn = q.nbuf[b & 1][slot];
                if (__float_as_uint(e.w) == stamp) add_term(L, e);
                if (__float_as_uint(nn.w) == stamp) add_term(L, nn);
            }
            org = xyz(o4);
            dir = xyz(d4);
            thr = xyz(th4);
            pdf = th4.w;
            t = h.x;
            prim = __float_as_int(h.y);
        }
        for (int b = bounce0;; ++b) {
            // zeroed: the odd lane's copy is only read through the shuffles
            ShadeOut sh{};
            if (shader) {
                shade_one(sc, so, views, segs, nseg, b, slot, type, org, dir, thr, pdf, t, prim, sh);
                if (sh.contrib) add_term(L, term(sh.add, 0u));
            }
            // the shadow ray to the odd lane; the even lane traces the
            // continuation, the odd lane the shadow ray, in one loop
            const float sx = __shfl(sh.s_o.x, partner), sy = __shfl(sh.s_o.y, partner),
                        sz = __shfl(sh.s_o.z, partner);
            const float dx = __shfl(sh.s_d.x, partner), dy = __shfl(sh.s_d.y, partner),
                        dz = __shfl(sh.s_d.z, partner);
            const float tmax = __shfl(sh.s_tmax, partner);
            const bool has_sh = __shfl((int)sh.push_sh, partner) != 0;
            f3 ro, rd;
            float rt;
            bool go;
            if (shader) {
                ro = sh.e_o;
                rd = sh.e_d;
                rt = 1e34f;
                go = sh.push_ext;
            } else {
                ro = mk3(sx, sy, sz);
                rd = mk3(dx, dy, dz);
                rt = tmax;
                go = has_sh;
            }
            int hit = -1;
            if (go) hit = trace_single<ACCEL>(sc, to.teps, ro, rd, rt, !shader, st, q.err);
            const int occ = __shfl(hit, partner);
            if (shader) {
                if (sh.push_sh) {
                    ++n_sh;
                    if (occ == -1) add_term(L, term(sh.s_c, 0u));
                }
                if (sh.push_ext) {
                    ++n_ext;
                    org = sh.e_o;
                    dir = sh.e_d;
                    thr = sh.e_thr;
                    pdf = sh.e_pdf;
                    type = sh.e_type;
                    t = rt;
                    prim = hit;
                    if (prim == -2) {
                        t = 1e34f;
                        prim = -1;
                    }
                }
            }
            const bool more = __shfl((int)sh.push_ext, (int)(__lane_id() & ~1u)) != 0;
            if (!more) break;
        }
        if (shader) {
            L.w = __uint_as_float((uint32_t)kMaxBounce);  // every term folded
            q.lpath[slot] = L;
        }
    }
    // one atomic per wave and counter
    for (int off = 32; off > 0; off >>= 1) {
        n_ext += __shfl_xor(n_ext, off);
        n_sh += __shfl_xor(n_sh, off);
    }
    if (__lane_id() == 0) {
        if (n_ext) atomicAdd(cnt_tail_ext(q.counters), n_ext);
        if (n_sh) atomicAdd(cnt_tail_shadow(q.counters), n_sh);
    }
}

// -------------------------------------------------------------- resolve --
// The radiance terms a path's shades have not folded in: stamped with this
// chunk and a bounce >= the folded count, which leaves at most the path's
// last two bounces (one per parity).  Added in bounce order, shade's term
// before connect's.
CT_DEV void fold_open_terms(float4& L, const Queues& q, uint32_t s, uint32_t gen) {
    const uint32_t kf = __float_as_uint(L.w);
    const float4 e0 = q.ebuf[0][s], n0 = q.nbuf[0][s], e1 = q.ebuf[1][s], n1 = q.nbuf[1][s];
    auto open = [&](float4 v) -> int {
        const uint32_t st = __float_as_uint(v.w);
        return ((st >> 6) == gen && (st & 63u) >= kf) ? (int)(st & 63u) : -1;
    };
    const int ke0 = open(e0), kn0 = open(n0), ke1 = open(e1), kn1 = open(n1);
    const int k0 = ke0 > kn0 ? ke0 : kn0, k1 = ke1 > kn1 ? ke1 : kn1;
    const bool one_first = k1 >= 0 && (k0 < 0 || k1 < k0);
    for (int pass = 0; pass < 2; ++pass) {
        if ((pass == 0) != one_first) {
            if (ke0 >= 0) add_term(L, e0);
            if (kn0 >= 0) add_term(L, n0);
        } else {
            if (ke1 >= 0) add_term(L, e1);
            if (kn1 >= 0) add_term(L, n1);
        }
    }
}

// accum[pixel] += per-sample radiance, samples in increasing order.
__global__ void k_resolve(const ViewDesc* __restrict__ views, const Segment* __restrict__ segs,
                          const ChunkHdr* hdr, Queues q, float4* accum) {
    const uint32_t n_tasks = hdr->n_tasks;
    const int nseg = (int)hdr->nseg;
    const uint32_t gen = hdr->gen;
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n_tasks; j += gridDim.x * blockDim.x) {
        int lo = 0, hi = nseg - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (segs[mid].pix_offset <= j) lo = mid;
            else hi = mid - 1;
        }
        const Segment sg = segs[lo];
        const ViewDesc& v = views[sg.view];
        const uint32_t spp = (uint32_t)v.spp;
        const uint32_t p = sg.first_path / spp + (j - sg.pix_offset);
        const uint32_t g0 = max(p * spp, sg.first_path);
        const uint32_t g1 = min((p + 1) * spp, sg.first_path + sg.n);
        float4* a = accum + v.accum_offset + pixel_of(p, (uint32_t)v.width, (uint32_t)v.height, v.strip_mul);
        float4 acc = *a;
        for (uint32_t g = g0; g < g1; ++g) {
            const uint32_t s = sg.chunk_offset + (g - sg.first_path);
            float4 L = q.lpath[s];
            fold_open_terms(L, q, s, gen);
            acc.x = acc.x + L.x;
            acc.y = acc.y + L.y;
            acc.z = acc.z + L.z;
            acc.w = acc.w + 1.0f;
        }
        *a = acc;
    }
}

// draw_framebuffer (Kernel.cu:258-270): rgb / samples, then x^(1/2.2).
__global__ void k_tonemap(const float4* __restrict__ accum, uint32_t n, float samples, float4* out) {
    const float e = 1.0f / 2.2f;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float4 a = accum[i];
        const f3 c = mk3(a.x, a.y, a.z) / samples;
        out[i] = make_float4(ct_powf(c.x, e), ct_powf(c.y, e), ct_powf(c.z, e), 1.0f);
    }
}

// Ray-level any-hit: occluded flag per ray (rays in sh_o / sh_d layout).
template <int ACCEL>
CT_DEV void flags_impl(const DevScene& sc, const TraceOpts& to, const Queues& q, uint32_t n, LaneStack& st,
                       uint8_t* out) {
    TravCount tc{0u, 0u, 0u, 0u};
    trace_persistent<ACCEL, kAny, false>(
        sc, to.teps, n, 0u, cnt_fetch_sh(q.counters, 0), st, tc, q.err,
        [&](uint32_t i, bool) {
            const float4 o4 = q.sh_o[0][i], d4 = q.sh_d[0][i];
            return RayIn{xyz(o4), xyz(d4), o4.w};
        },
        [&](uint32_t i, bool, float, int prim) { out[i] = prim >= 0 ? 1 : 0; });
}

template <int TREE>
__global__ void k_shadow_flags(DevScene sc, TraceOpts to, Queues q, uint32_t n, int lds_depth, uint8_t* out) {
    LaneStack st = make_stack(q, lds_depth);
    flags_impl<TREE>(sc, to, q, n, st, out);
}

}  // namespace ctd

// ------------------------------------------------------------ launchers --
namespace ctk {

using namespace ctd;

hipError_t launch_setup(uint32_t* counters, const ChunkHdr* hdr, hipStream_t s) {
    hipLaunchKernelGGL(k_setup, dim3(1), dim3(256), 0, s, counters, hdr);
    return hipGetLastError();
}
hipError_t launch_setup_rays(uint32_t* counters, uint32_t* err, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_setup_rays, dim3(1), dim3(256), 0, s, counters, err, n);
    return hipGetLastError();
}
hipError_t launch_generate(const LaunchShape& ls, const ViewDesc* views, const Segment* segs,
                           const ChunkHdr* hdr, uint32_t max_n, const Queues& q, hipStream_t s) {
    const uint32_t blocks = std::max<uint32_t>(1u, std::min<uint32_t>((max_n + 255) / 256, (uint32_t)ls.max_blocks));
    hipLaunchKernelGGL(k_generate, dim3(blocks), dim3(256), 0, s, views, segs, hdr, q);
    return hipGetLastError();
}
int tree_of(const TraceOpts& to) {
    if (to.traversal == CT_TRAV_FAST) return 8;
    if (to.traversal == CT_TRAV_COMPACT) return kTreeCompact;
    return to.accel == CT_ACCEL_BVH2 ? 2 : 4;
}
int grid_of(const LaunchShape& ls, int tree) {
    return tree == 2 ? ls.grid2 : (tree == 4 ? ls.grid4 : (tree == kTreeCompact ? ls.gridc : ls.grid8));
}
int full_of(const LaunchShape& ls, int tree) {
    return tree == 2 ? ls.full2 : (tree == 4 ? ls.full4 : (tree == kTreeCompact ? ls.fullc : ls.full8));
}

template <int TREE>
void extend_t(const LaunchShape& ls, const DevScene& sc, const TraceOpts& to, const Queues& q, int sel, int bounce,
              hipStream_t s) {
    const size_t lds = (size_t)ls.lds_depth * ls.block * 4;
    const dim3 grid(full_of(ls, TREE)), block(ls.block);
    hipLaunchKernelGGL((k_extend<TREE>), grid, block, lds, s, sc, to, q, sel, bounce, ls.lds_depth);
}
template <int TREE>
void trace_t(const LaunchShape& ls, const DevScene& sc, const TraceOpts& to, const Queues& q, const ChunkHdr* hdr,
             int sel, int bounce, unsigned long long* wstats, hipStream_t s) {
    const size_t lds = (size_t)ls.lds_depth * ls.block * 4;
    const dim3 grid(grid_of(ls, TREE)), block(ls.block);
    if (wstats)
        hipLaunchKernelGGL((k_trace<TREE, true>), grid, block, lds, s, sc, to, q, hdr, sel, bounce, ls.lds_depth,
                           wstats);
    else
        hipLaunchKernelGGL((k_trace<TREE, false>), grid, block, lds, s, sc, to, q, hdr, sel, bounce, ls.lds_depth,
                           wstats);
}

hipError_t launch_extend(const LaunchShape& ls, const DevScene& sc, const TraceOpts& to, const Queues& q, int sel,
                         int bounce, hipStream_t s) {
    switch (tree_of(to)) {
        case 2: extend_t<2>(ls, sc, to, q, sel, bounce, s); break;
        case 4: extend_t<4>(ls, sc, to, q, sel, bounce, s); break;
        case kTreeCompact: extend_t<kTreeCompact>(ls, sc, to, q, sel, bounce, s); break;
        default: extend_t<8>(ls, sc, to, q, sel, bounce, s); break;
    }
    return hipGetLastError();
}
hipError_t launch_trace(const LaunchShape& ls, const DevScene& sc, const TraceOpts& to, const Queues& q,
                        const ChunkHdr* hdr, int sel, int bounce, unsigned long long* wstats, hipStream_t s) {
    switch (tree_of(to)) {
        case 2: trace_t<2>(ls, sc, to, q, hdr, sel, bounce, wstats, s); break;
        case 4: trace_t<4>(ls, sc, to, q, hdr, sel, bounce, wstats, s); break;
        case kTreeCompact: trace_t<kTreeCompact>(ls, sc, to, q, hdr, sel, bounce, wstats, s); break;
        default: trace_t<8>(ls, sc, to, q, hdr, sel, bounce, wstats, s); break;
    }
    return hipGetLastError();
}
hipError_t launch_tail(const LaunchShape& ls, const DevScene& sc, const TraceOpts& to, const ShadeOpts& so,
                       const ViewDesc* views, const Segment* segs, const ChunkHdr* hdr, const Queues& q, int sel,
                       int bounce0, hipStream_t s) {
    const size_t lds = (size_t)ls.lds_depth * ls.block * 4;
    const dim3 grid(ls.tail_blocks), block(ls.block);
    switch (tree_of(to)) {
        case 2:
            hipLaunchKernelGGL(k_tail<2>, grid, block, lds, s, sc, to, so, views, segs, hdr, q, sel, bounce0,
                               ls.lds_depth);
            break;
        case 4:
            hipLaunchKernelGGL(k_tail<4>, grid, block, lds, s, sc, to, so, views, segs, hdr, q, sel, bounce0,
                               ls.lds_depth);
            break;
        case kTreeCompact:
            hipLaunchKernelGGL(k_tail<kTreeCompact>, grid, block, lds, s, sc, to, so, views, segs, hdr, q, sel,
                               bounce0, ls.lds_depth);
            break;
        default:
            hipLaunchKernelGGL(k_tail<8>, grid, block, lds, s, sc, to, so, views, segs, hdr, q, sel, bounce0,
                               ls.lds_depth);
            break;
    }
    return hipGetLastError();
}
hipError_t launch_shade(const LaunchShape& ls, const DevScene& sc, const ShadeOpts& so, const ViewDesc* views,
                        const Segment* segs, const ChunkHdr* hdr, const Queues& q, int sel, int bounce,
                        hipStream_t s) {
    if (so.sort)
        hipLaunchKernelGGL(k_shade<true>, dim3(ls.shade_blocks), dim3(kShadeBlock), 0, s, sc, so, views, segs, hdr, q,
                           sel, bounce);
    else
        hipLaunchKernelGGL(k_shade<false>, dim3(ls.shade_blocks), dim3(kShadeBlock), 0, s, sc, so, views, segs, hdr,
                           q, sel, bounce);
    return hipGetLastError();
}
hipError_t launch_resolve(const LaunchShape& ls, const ViewDesc* views, const Segment* segs, const ChunkHdr* hdr,
                          uint32_t max_tasks, const Queues& q, float4* accum, hipStream_t s) {
    const uint32_t blocks =
        std::max<uint32_t>(1u, std::min<uint32_t>((max_tasks + 255) / 256, (uint32_t)ls.max_blocks));
    hipLaunchKernelGGL(k_resolve, dim3(blocks), dim3(256), 0, s, views, segs, hdr, q, accum);
    return hipGetLastError();
}
// The tonemapped image as 8-bit RGBA (alpha 255): byte = clamp(v, 0, 1) * 255 + 0.5, truncated.
__global__ void k_tonemap8(const float4* __restrict__ accum, uint32_t n, float samples, uint32_t* out) {
    const float e = 1.0f / 2.2f;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float4 a = accum[i];
        const f3 c = mk3(a.x, a.y, a.z) / samples;
        const float v[3] = {ct_powf(c.x, e), ct_powf(c.y, e), ct_powf(c.z, e)};
        uint32_t p = 0xFF000000u;
        for (int k = 0; k < 3; ++k) {
            const float q = v[k] > 0.0f ? (v[k] < 1.0f ? v[k] : 1.0f) : 0.0f;
            p |= (uint32_t)(q * 255.0f + 0.5f) << (8 * k);
        }
        out[i] = p;
    }
}
hipError_t launch_tonemap8(const float4* accum, uint32_t n, float samples, uint32_t* out, hipStream_t s) {
    const uint32_t blocks = std::min<uint32_t>((n + 255) / 256, 65536u);
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(k_tonemap8, dim3(blocks), dim3(256), 0, s, accum, n, samples, out);
    return hipGetLastError();
}
// One axis of a compact quad node (cnode4_step, device/traverse.h): the
// children's [lo, hi] (slots in `valid`) as 8-bit offsets on the grid
// p + q 2^k with p = the smallest lo.  The smallest step whose 251 points
// span the range first, coarser ones until every decoded plane
// fmaf(q, 2^k, p) contains its box in float; offsets round outwards, with no
// further margin.  A decoded box can then differ from the node's box test by
// the rounding of the slab values, a few ulps of |p - o| / |d| along the
// ray.  That changes nothing for trees whose primitive boxes are padded (the
// builders pad by 1e-3, scene.cpp:99): a ray whose slab interval is that
// close to empty, or whose entry is that close to the current hit, passes
// within ulps of a box face and cannot reach a triangle, which lies 1e-3
// inside every box that holds it.  (The wide nodes' extra margin step cost
// C2 9% of the compact mode's speed: 11% more node tests against 2%.)
// False if no step up to 2^kCompactMaxExp works or a plane is not finite.
CT_DEV bool compact_axis(const float lo[4], const float hi[4], uint32_t valid, float& p_out, uint32_t& k_biased,
                         uint32_t& wlo, uint32_t& whi) {
    float mn = INFINITY, mx = -INFINITY;
    for (int i = 0; i < 4; ++i) {
        if (!((valid >> i) & 1u)) continue;
        if (!isfinite(lo[i]) || !isfinite(hi[i])) return false;
        mn = fminf(mn, lo[i]);
        mx = fmaxf(mx, hi[i]);
    }
    p_out = 0.0f;
    k_biased = 127u;
    wlo = whi = 0u;
    if (!valid) return true;
    const double ext = (double)mx - (double)mn;
    int k = -126;
    if (ext > 0.0) k = max(-126, (int)ceil(log2(ext / 250.0)));
    for (; k <= kCompactMaxExp; ++k) {
        const double s = ldexp(1.0, k);
        const float sf = (float)s;  // exact: -126 <= k <= 60
        const float p = mn;
        bool ok = true;
        uint32_t a = 0u, b = 0u;
        for (int i = 0; i < 4 && ok; ++i) {
            if (!((valid >> i) & 1u)) continue;
            const double ql = floor(((double)lo[i] - (double)p) / s);
            const double qh = ceil(((double)hi[i] - (double)p) / s);
            if (ql < 0.0 || qh > 255.0 || ql > qh) { ok = false; break; }
            if (!(fmaf((float)ql, sf, p) <= lo[i]) || !(fmaf((float)qh, sf, p) >= hi[i])) { ok = false; break; }
            a |= (uint32_t)ql << (8 * i);
            b |= (uint32_t)qh << (8 * i);
        }
        if (ok) {
            p_out = p;
            k_biased = (uint32_t)(k + 127);
            wlo = a;
            whi = b;
            return true;
        }
    }
    return false;
}

// Per quad node: the fast node test's layout (the six plane rows, the four
// stack entries kLeafBit | first triangle record or the child node, a zero
// row), the compact layout of cnode4_step (same entries, kNoEntry for an
// empty slot) and the tree's kPlaneNegZero / kPlaneNaN / kEmptyEntry /
// kCompactBad flags (flags cleared by the caller; one atomic per wave that
// found something).
__global__ void k_prepare_quad(const float4* __restrict__ nodes4, uint32_t n4, float4* __restrict__ fast,
                               uint4* __restrict__ compact, uint32_t* flags) {
    uint32_t f = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += gridDim.x * blockDim.x) {
        const float4* n = nodes4 + 8 * (size_t)i;
        float4* o = fast + 8 * (size_t)i;
        float pl[6][4];
        for (int r = 0; r < 6; ++r) {
            const float4 v = n[r];
            o[r] = v;
            const float c[4] = {v.x, v.y, v.z, v.w};
            for (int k = 0; k < 4; ++k) {
                pl[r][k] = c[k];
                if (__float_as_uint(c[k]) == 0x80000000u) f |= ctd::kPlaneNegZero;
                if (isnan(c[k])) f |= ctd::kPlaneNaN;
            }
        }
        const float4 ch = n[6], cn = n[7];
        const uint32_t child[4] = {__float_as_uint(ch.x), __float_as_uint(ch.y), __float_as_uint(ch.z),
                                   __float_as_uint(ch.w)};
        const uint32_t count[4] = {__float_as_uint(cn.x), __float_as_uint(cn.y), __float_as_uint(cn.z),
                                   __float_as_uint(cn.w)};
        uint32_t e[4], ce[4], valid = 0u;
        for (int k = 0; k < 4; ++k) {
            e[k] = count[k] == 0xFFFFFFFFu ? child[k] : (ctd::kLeafBit | child[k]);
            if (count[k] == 0u && child[k] != 0xFFFFFFFFu) f |= ctd::kEmptyEntry;
            ce[k] = count[k] == 0u ? ctd::kNoEntry : e[k];
            if (count[k] != 0u) valid |= 1u << k;
        }
        o[6] = make_float4(__uint_as_float(e[0]), __uint_as_float(e[1]), __uint_as_float(e[2]), __uint_as_float(e[3]));
        o[7] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        float p[3];
        uint32_t kb[3], wl[3], wh[3];
        for (int a = 0; a < 3; ++a)
            if (!compact_axis(pl[2 * a], pl[2 * a + 1], valid, p[a], kb[a], wl[a], wh[a])) f |= ctd::kCompactBad;
        uint4* w = compact + 4 * (size_t)i;
        w[0] = make_uint4(__float_as_uint(p[0]), __float_as_uint(p[1]), __float_as_uint(p[2]),
                          kb[0] | (kb[1] << 8) | (kb[2] << 16));
        w[1] = make_uint4(wl[0], wh[0], wl[1], wh[1]);
        w[2] = make_uint4(wl[2], wh[2], 0u, 0u);
        w[3] = make_uint4(ce[0], ce[1], ce[2], ce[3]);
    }
    uint32_t w = 0;
    for (uint32_t bit : {ctd::kPlaneNegZero, ctd::kPlaneNaN, ctd::kEmptyEntry, ctd::kCompactBad})
        if (__ballot(f & bit)) w |= bit;
    if (__lane_id() == 0 && w) atomicOr(flags, w);
}
hipError_t launch_prepare_quad(const float4* nodes4, uint32_t n4, float4* fast, uint4* compact, uint32_t* flags,
                               hipStream_t s) {
    const uint32_t blocks = std::min<uint32_t>((n4 + 255) / 256, 4096u);
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(k_prepare_quad, dim3(blocks), dim3(256), 0, s, nodes4, n4, fast, compact, flags);
    return hipGetLastError();
}
hipError_t launch_tonemap(const float4* accum, uint32_t n, float samples, float4* out, hipStream_t s) {
    const uint32_t blocks = std::min<uint32_t>((n + 255) / 256, 65536u);
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(k_tonemap, dim3(blocks), dim3(256), 0, s, accum, n, samples, out);
    return hipGetLastError();
}
hipError_t launch_shadow_flags(const LaunchShape& ls, const DevScene& sc, const TraceOpts& to, const Queues& q,
                               uint32_t n, uint8_t* out, hipStream_t s) {
    const size_t lds = (size_t)ls.lds_depth * ls.block * 4;
    const int tree = tree_of(to);
    const dim3 grid(full_of(ls, tree)), block(ls.block);
    if (tree == 2) hipLaunchKernelGGL(k_shadow_flags<2>, grid, block, lds, s, sc, to, q, n, ls.lds_depth, out);
    else if (tree == 4) hipLaunchKernelGGL(k_shadow_flags<4>, grid, block, lds, s, sc, to, q, n, ls.lds_depth, out);
    else if (tree == kTreeCompact)
        hipLaunchKernelGGL(k_shadow_flags<kTreeCompact>, grid, block, lds, s, sc, to, q, n, ls.lds_depth, out);
    else hipLaunchKernelGGL(k_shadow_flags<8>, grid, block, lds, s, sc, to, q, n, ls.lds_depth, out);
    return hipGetLastError();
}

}  // namespace ctk

namespace ctk {
// Resident blocks per CU of the persistent traversal kernels of one tree
// (extend and trace, the smaller of the two) for a block size and LDS depth.
template <int TREE>
static int per_cu(int block, size_t lds) {
    int nb = 0, nc = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void*>(&ctd::k_extend<TREE>),
                                                     block, lds) != hipSuccess)
        return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nc, reinterpret_cast<const void*>(&ctd::k_trace<TREE, false>),
                                                     block, lds) != hipSuccess)
        return 0;
    return nb < nc ? nb : nc;
}

bool set_trace_grids(LaunchShape& ls, int num_cus) {
    const size_t lds = (size_t)ls.lds_depth * ls.block * 4;
    const int p2 = per_cu<2>(ls.block, lds), p4 = per_cu<4>(ls.block, lds), p8 = per_cu<8>(ls.block, lds);
    const int pc = per_cu<ctd::kTreeCompact>(ls.block, lds);
    if (p2 <= 0 || p4 <= 0 || p8 <= 0 || pc <= 0) return false;
    // at least one block per fetch range (the persistent kernels drain ranges
    // by block id)
    const int lo = ctd::kFetchRanges;
    const int fill = std::min(8, std::max(1, ls.trace_fill));
    ls.full2 = std::max(lo, num_cus * p2);
    ls.full4 = std::max(lo, num_cus * p4);
    ls.full8 = std::max(lo, num_cus * p8);
    ls.fullc = std::max(lo, num_cus * pc);
    ls.grid2 = std::max(lo, ls.full2 * fill / 8);
    ls.grid4 = std::max(lo, ls.full4 * fill / 8);
    ls.grid8 = std::max(lo, ls.full8 * fill / 8);
    ls.gridc = std::max(lo, ls.fullc * fill / 8);
    ls.trace_blocks = std::max(std::max(ls.full2, ls.full4), std::max(ls.full8, ls.fullc));
    // k_tail: the blocks of its most register-hungry tree that fit at once
    int pt = 1 << 30;
    for (const void* f : {reinterpret_cast<const void*>(&ctd::k_tail<2>), reinterpret_cast<const void*>(&ctd::k_tail<4>),
                          reinterpret_cast<const void*>(&ctd::k_tail<ctd::kTreeCompact>),
                          reinterpret_cast<const void*>(&ctd::k_tail<8>)}) {
        int nb = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, ls.block, lds) != hipSuccess || nb <= 0) return false;
        pt = std::min(pt, nb);
    }
    ls.tail_blocks = std::min(ls.trace_blocks, std::max(1, num_cus * pt));
    return true;
}
}  // namespace ctk
