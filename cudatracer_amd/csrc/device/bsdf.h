// Shading math on gfx950: triangle helpers (src/Core/Triangle.cuh:46-96),
// the cosine lobe and tangent frame (src/Core/Tools.h:110-169), the
// reference's Lambert/mirror BSDF (src/Core/BSDF.h:26-63) and the material
// extensions the configurations need: smooth dielectric (Fresnel after
// Tools.h:96-108) and GGX/Beckmann microfacet reflection with visible-normal
// sampling from src/Core/Microfacet.cuh.  Transcendentals come from
// include/ct_fp.h so the FP32 oracle reproduces every bit.
#pragma once

#include "dev_common.h"

namespace ctd {

// ---------------------------------------------------------- triangles --
CT_DEV f3 bary_coords(f3 p, f3 n, f3 p0, f3 p1, f3 p2) {
    const float abc = dot(n, cross(p1 - p0, p2 - p0));
    const float pbc = dot(n, cross(p1 - p, p2 - p));
    const float pca = dot(n, cross(p2 - p, p0 - p));
    const float alpha = pbc / abc;
    const float beta = pca / abc;
    const float gamma = 1.0f - alpha - beta;
    return mk3(alpha, beta, gamma);
}
CT_DEV f3 random_point(f3 p0, f3 p1, f3 p2, float r1, float r2) {
    if (r1 + r2 > 1.0f) { r1 = 1.0f - r1; r2 = 1.0f - r2; }
    const float a = 1.0f - r1 - r2;
    return a * p0 + r1 * p1 + r2 * p2;
}
CT_DEV float tri_area(f3 p0, f3 p1, f3 p2) {
    const float a = length(p0 - p1), b = length(p1 - p2), c = length(p2 - p0);
    const float s = (a + b + c) / 2.0f;
    return ct_sqrtf(s * (s - a) * (s - b) * (s - c));
}

// --------------------------------------------------------------- frames --
CT_DEV void tangent_frame(f3 N, f3& T, f3& B) {
    if (N.z < 0.0f) {
        const float a = 1.0f / (1.0f - N.z);
        const float b = N.x * N.y * a;
        B = mk3(1.0f - N.x * N.x * a, -b, N.x);
        T = mk3(b, N.y * N.y * a - 1.0f, -N.y);
    } else {
        const float a = 1.0f / (1.0f + N.z);
        const float b = -N.x * N.y * a;
        B = mk3(1.0f - N.x * N.x * a, b, -N.x);
        T = mk3(b, 1.0f - N.y * N.y * a, -N.y);
    }
}
// convertToLocalSpace + sampleToWorld (Random.cuh:37-53): the frame the
// reference rotates normal-map samples with.
CT_DEV f3 sample_to_world(f3 s, f3 N) {
    const f3 W = ct_fabsf(N.x) > 0.99f ? mk3(0.0f, 1.0f, 0.0f) : mk3(1.0f, 0.0f, 0.0f);
    const f3 T = normalize(cross(N, W));
    const f3 B = cross(N, T);
    return s.x * T + s.y * B + s.z * N;
}
// Nearest texel of texture t at (u, v) (SceneData::getTextureColor,
// SceneData.cuh:81-97).  info[t] = {first texel, width, height, format}:
// format 0 texels are float4; format 1 texels are 8-bit RGB (alpha 1) that
// decode to c / 255.0f, the floats the reference's Surface holds for 8-bit
// images, in a quarter of the bytes.  Non-temporal loads: texels are
// fetched at random during shading and should not evict the tree from L2.
CT_DEV float4 tex_fetch(const float4* __restrict__ texels, const uint32_t* __restrict__ texels8,
                        const uint4* __restrict__ info, int t, float u, float v) {
    const uint4 ti = info[t];
    const uint32_t i = ti.x + ct_texel_index(ti.y, ti.z, u, v);
    if (ti.w) {
        const uint32_t p = __builtin_nontemporal_load(texels8 + i);
        return make_float4((float)(p & 255u) / 255.0f, (float)((p >> 8) & 255u) / 255.0f,
                           (float)((p >> 16) & 255u) / 255.0f, 1.0f);
    }
    const float* f = reinterpret_cast<const float*>(texels + i);
    return make_float4(__builtin_nontemporal_load(f), __builtin_nontemporal_load(f + 1),
                       __builtin_nontemporal_load(f + 2), __builtin_nontemporal_load(f + 3));
}
CT_DEV f3 tangent_to_world(f3 V, f3 N) {
    f3 T, B;
    tangent_frame(N, T, B);
    return V.x * T + V.y * B + V.z * N;
}
// DiffuseReflectionCosWeighted (Tools.h:162-169): sqrt(1.0 - r1) in double.
CT_DEV f3 cos_weighted(float r0, float r1) {
    const float term1 = CT_TWO_PI_F * r0;
    const float term2 = (float)ct_sqrt(1.0 - (double)r1);
    const float s = ct_sinf(term1);
    const float c = ct_cosf(term1);
    return mk3(c * term2, s * term2, ct_sqrtf(r1));
}
CT_DEV f3 reflect(f3 I, f3 N) { return I - N * dot(N, I) * 2.0f; }

// ------------------------------------------------------------ microfacet --
CT_DEV float cos2t(f3 w) { return w.z * w.z; }
CT_DEV float sin2t(f3 w) { return gmax(0.0f, 1.0f - cos2t(w)); }
CT_DEV float sint(f3 w) { return ct_sqrtf(sin2t(w)); }
CT_DEV float tant(f3 w) { return sint(w) / w.z; }
CT_DEV float tan2t(f3 w) { return sin2t(w) / cos2t(w); }
CT_DEV float clamp11(float v) { return v < -1.0f ? -1.0f : (v > 1.0f ? 1.0f : v); }
CT_DEV float cosphi(f3 w) { const float st = sint(w); return (st == 0.0f) ? 1.0f : clamp11(w.x / st); }
CT_DEV float sinphi(f3 w) { const float st = sint(w); return (st == 0.0f) ? 0.0f : clamp11(w.y / st); }

enum : int { MAT_LIGHT = 0, MAT_SPECULAR = 1, MAT_FRESNEL = 2, MAT_LAMBERT = 3, MAT_BECKMANN = 4, MAT_GGX = 5,
             MAT_TROWBRIDGE = 6, MAT_FRESNEL_BECKMANN = 7, MAT_FRESNEL_GGX = 8, MAT_FRESNEL_TROWBRIDGE = 9 };
CT_DEV bool is_beckmann(int t) { return t == MAT_BECKMANN || t == MAT_FRESNEL_BECKMANN; }
CT_DEV bool is_microfacet(int t) { return t == MAT_BECKMANN || t == MAT_GGX || t == MAT_TROWBRIDGE; }
CT_DEV bool is_dielectric(int t) {
    return t == MAT_FRESNEL || t == MAT_FRESNEL_BECKMANN || t == MAT_FRESNEL_GGX || t == MAT_FRESNEL_TROWBRIDGE;
}

// Beckmann D, Microfacet.cuh:129-139 (isotropic)
CT_DEV float D_beckmann(f3 wh, float alpha) {
    const float tan2 = tan2t(wh);
    if ((2.0f * tan2) == tan2) return 0.0f;
    const float cos4 = cos2t(wh) * cos2t(wh);
    const float cp = cosphi(wh), sp = sinphi(wh);
    return ct_expf(-tan2 * (cp * cp / (alpha * alpha) + sp * sp / (alpha * alpha))) /
           (CT_PI_F * alpha * alpha * cos4);
}
// GGX D (PBRT-v3 TrowbridgeReitzDistribution::D, isotropic)
CT_DEV float D_ggx(f3 wh, float alpha) {
    const float tan2 = tan2t(wh);
    if ((2.0f * tan2) == tan2) return 0.0f;
    const float cos4 = cos2t(wh) * cos2t(wh);
    const float e = tan2 / (alpha * alpha);
    return 1.0f / (CT_PI_F * alpha * alpha * cos4 * (1.0f + e) * (1.0f + e));
}
// lambda_ggx, Microfacet.cuh:417-426
CT_DEV float lambda_ggx(f3 w, float alpha) {
    const float absTan = ct_fabsf(tant(w));
    if (2.0f * absTan == absTan) return 0.0f;
    const float cp = cosphi(w), sp = sinphi(w);
    const float a = ct_sqrtf(cp * cp * alpha * alpha + sp * sp * alpha * alpha);
    const float a2t2 = a * absTan * a * absTan;
    return (-1.0f + ct_sqrtf(1.0f + a2t2)) / 2.0f;
}
// lambda_beckmann, Microfacet.cuh:500-513
CT_DEV float lambda_beckmann(f3 w, float alpha) {
    const float absTan = ct_fabsf(tant(w));
    if ((2.0f * absTan) == absTan) return 0.0f;
    const float cp = cosphi(w), sp = sinphi(w);
    const float al = ct_sqrtf(cp * cp * alpha * alpha + sp * sp * alpha * alpha);
    const float a = 1.0f / (al * absTan);
    if (a >= 1.6f) return 0.0f;
    return (1.0f - 1.259f * a + 0.396f * a * a) / (3.535f * a + 2.181f * a * a);
}
// sampleGGX_P22_11, Microfacet.cuh:312-351
CT_DEV void ggx_p22_11(float cosThetaI, float* sx, float* sy, float r1, float r2) {
    if (cosThetaI > 0.9999f) {
        const float r = ct_sqrtf(r1 / (1.0f - r1));
        const float sinPhi = ct_sinf(CT_TWO_PI_F * r2);
        const float cosPhi = ct_cosf(CT_TWO_PI_F * r2);
        *sx = r * cosPhi;
        *sy = r * sinPhi;
        return;
    }
    const float sinThetaI = ct_sqrtf(gmax(0.0f, 1.0f - cosThetaI * cosThetaI));
    const float tanThetaI = sinThetaI / cosThetaI;
    const float a = 1.0f / tanThetaI;
    const float G1 = 2.0f / (1.0f + ct_sqrtf(1.0f + 1.0f / (a * a)));
    const float A = 2.0f * r1 / G1 - 1.0f;
    const float B = tanThetaI;
    const float tmp = gmin(1.0f / (A * A - 1.0f), 1.0e12f);
    const float D = ct_sqrtf(B * B * tmp * tmp - (A * A - B * B) * tmp);
    const float sx1 = B * tmp - D;
    const float sx2 = B * tmp + D;
    *sx = (A < 0.0f || sx2 > 1.0f / tanThetaI) ? sx1 : sx2;
    float S;
    if (r2 > 0.5f) { S = 1.0f; r2 = 2.0f * (r2 - 0.5f); }
    else { S = -1.0f; r2 = 2.0f * (0.5f - r2); }
    const float z = (r2 * (r2 * (r2 * 0.27385f - 0.73369f) + 0.46341f)) /
                    (r2 * (r2 * (r2 * 0.093073f + 0.309420f) - 1.000000f) + 0.597999f);
    *sy = S * z * ct_sqrtf(1.0f + (*sx) * (*sx));
}
// BeckmannSample11, Microfacet.cuh:141-209
CT_DEV void beckmann_11(float cosThetaI, float r1, float r2, float* sx, float* sy) {
    if (cosThetaI > 0.9999f) {
        const float r = ct_sqrtf(-ct_logf(1.0f - r1));
        const float sinPhi = ct_sinf(2.0f * CT_PI_F * r2);
        const float cosPhi = ct_cosf(2.0f * CT_PI_F * r2);
        *sx = r * cosPhi;
        *sy = r * sinPhi;
        return;
    }
    const float sinThetaI = ct_sqrtf(gmax(0.0f, 1.0f - cosThetaI * cosThetaI));
    const float tanThetaI = sinThetaI / cosThetaI;
    const float cotThetaI = 1.0f / tanThetaI;
    float a = -1.0f;
    float c = ct_erff(cotThetaI);
    const float sample_x = gmax(r1, 1e-6f);
    const float thetaI = ct_acosf(cosThetaI);
    const float fit = 1.0f + thetaI * (-0.876f + thetaI * (0.4265f - 0.0594f * thetaI));
    float b = c - (1.0f + c) * ct_powf(1.0f - sample_x, fit);
    const float normalization =
        1.0f / (1.0f + c + 1.0f / CT_ROOT_PI_F * tanThetaI * ct_expf(-cotThetaI * cotThetaI));
    int it = 0;
    while (++it < 10) {
        if (!(b >= a && b <= c)) b = 0.5f * (a + c);
        const float invErf = ct_erfinvf(b);
        const float value =
            normalization * (1.0f + b + 1.0f / ct_sqrtf(CT_PI_F) * tanThetaI * ct_expf(-invErf * invErf)) - sample_x;
        const float derivative = normalization * (1.0f - invErf * tanThetaI);
        if (ct_fabsf(value) < 1e-5f) break;
        if (value > 0.0f) c = b;
        else a = b;
        b -= value / derivative;
    }
    *sx = ct_erfinvf(b);
    *sy = ct_erfinvf(2.0f * gmax(r2, 1e-6f) - 1.0f);
}
// sampleGGX / BeckmannSample with the sample_ggx / sample_beckmann flips.
CT_DEV f3 sample_vndf(int type, f3 wo, float alpha, float r1, float r2) {
    const bool flip = wo.z < 0.0f;
    const f3 wi = flip ? -wo : wo;
    const f3 ws = normalize(mk3(alpha * wi.x, alpha * wi.y, wi.z));
    float sx, sy;
    if (is_beckmann(type)) beckmann_11(ws.z, r1, r2, &sx, &sy);
    else ggx_p22_11(ws.z, &sx, &sy, r1, r2);
    const float cp = cosphi(ws), sp = sinphi(ws);
    const float tmp = cp * sx - sp * sy;
    sy = sp * sx + cp * sy;
    sx = tmp;
    sx = alpha * sx;
    sy = alpha * sy;
    f3 wm = normalize(mk3(-sx, -sy, 1.0f));
    if (is_beckmann(type)) {
        if (flip) wm = -wm;
    } else if (wm.z < 0.0f) {
        wm = -wm;
    }
    return wm;
}
// f = F D G2 / (4 cos_o cos_i), pdf = D G1(wo) / (4 cos_o); Schlick F0 = albedo.
CT_DEV void mf_eval(int type, float alpha, f3 albedo, f3 wo, f3 wi, f3* f, float* pdf) {
    if (wo.z <= 0.0f || wi.z <= 0.0f) { *f = mk3(0, 0, 0); *pdf = 0.0f; return; }
    const f3 wm = normalize(wo + wi);
    const bool bk = is_beckmann(type);
    const float D = bk ? D_beckmann(wm, alpha) : D_ggx(wm, alpha);
    const float Lo = bk ? lambda_beckmann(wo, alpha) : lambda_ggx(wo, alpha);
    const float Li = bk ? lambda_beckmann(wi, alpha) : lambda_ggx(wi, alpha);
    const float G1o = 1.0f / (1.0f + Lo);
    const float G2 = 1.0f / (1.0f + Lo + Li);
    const float c = gmax(0.0f, dot(wo, wm));
    const float m = 1.0f - c;
    const float m5 = m * m * m * m * m;
    const f3 F = albedo + (mk3(1.0f, 1.0f, 1.0f) - albedo) * m5;
    *pdf = D * G1o / (4.0f * wo.z);
    *f = F * (D * G2 / (4.0f * wo.z * wi.z));
}
// Dielectric Fresnel, relative IOR eta = n_t / n_i (Tools.h:96-108, eta < 1 allowed)
CT_DEV float fresnel_dielectric(float cosThetaI, float eta) {
    const float scale = 1.0f / eta;
    const float cosThetaTSqr = 1.0f - (1.0f - cosThetaI * cosThetaI) * (scale * scale);
    if (cosThetaTSqr <= 0.0f) return 1.0f;
    const float cosThetaT = ct_sqrtf(cosThetaTSqr);
    const float Rs = (cosThetaI - eta * cosThetaT) / (cosThetaI + eta * cosThetaT);
    const float Rp = (eta * cosThetaI - cosThetaT) / (eta * cosThetaI + cosThetaT);
    return 0.5f * (Rs * Rs + Rp * Rp);
}

}  // namespace ctd
