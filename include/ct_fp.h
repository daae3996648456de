/*
 * ct_fp.h — the numeric contract of the path-tracing loop.
 *
 * Everything here is compiled identically for the host (g++/clang, SSE2,
 * -ffp-contract=off) and for gfx950 (hipcc, -ffp-contract=off, correctly
 * rounded fp32 divide/sqrt), so that the FP32 CPU oracle and the HIP kernels
 * produce bit-identical values from the same inputs:
 *
 *   - the counter-based RNG of the reference (WangHash + xorshift32),
 *     src/Core/Random.cuh:10-34;
 *   - float transcendental functions written from published methods
 *     (Cody & Waite argument reduction, Taylor polynomials evaluated with
 *     fused multiply-add, Giles' erfinv and Abramowitz & Stegun 7.1.26 erf
 *     as used in src/Core/Microfacet.cuh:36-82).  The reference calls the
 *     device math library for these; its exact outputs are not reproducible,
 *     so this header defines them for both sides.  Accuracy against libm
 *     (double) is checked by the test-suite (max error a few ulp).
 *
 * No state, no allocation; plain C++ (usable from C++ host code and HIP).
 */
#ifndef CT_FP_H
#define CT_FP_H

#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define CT_HD __host__ __device__ inline
#else
#include <math.h>
#define CT_HD inline
#endif

/* ---------------------------------------------------------------- bits --- */
CT_HD uint32_t ct_f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
CT_HD float ct_u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

CT_HD float ct_fmaf(float a, float b, float c) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_fmaf(a, b, c);
#else
    return fmaf(a, b, c);
#endif
}
CT_HD float ct_sqrtf(float x) {  /* IEEE correctly rounded on both sides */
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_sqrtf(x);
#else
    return sqrtf(x);
#endif
}
CT_HD double ct_sqrt(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_sqrt(x);
#else
    return sqrt(x);
#endif
}
/* round-half-even to integral value */
CT_HD float ct_rintf(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_rintf(x);
#else
    return rintf(x);
#endif
}
CT_HD float ct_fabsf(float x) { return ct_u2f(ct_f2u(x) & 0x7fffffffu); }

/* ----------------------------------------------------------------- RNG --- */
/* src/Core/Random.cuh:10-17 */
CT_HD uint32_t ct_wang_hash(uint32_t s) {
    s = (s ^ 61u) ^ (s >> 16);
    s *= 9u;
    s = s ^ (s >> 4);
    s *= 0x27d4eb2du;
    s = s ^ (s >> 15);
    return s;
}
/* xorshift32 (13, 17, 5), src/Core/Random.cuh:19-25 */
CT_HD uint32_t ct_rand_u32(uint32_t* seed) {
    uint32_t s = *seed;
    s ^= s << 13;
    s ^= s << 17;
    s ^= s << 5;
    *seed = s;
    return s;
}
/* src/Core/Random.cuh:27-30: float(u) * 2.3283064365387e-10f (may return 1.0) */
CT_HD float ct_rand_float(uint32_t* seed) {
    return (float)ct_rand_u32(seed) * 2.3283064365387e-10f;
}
/* src/Core/Random.cuh:32-34 */
CT_HD int ct_rand_int_max(uint32_t* seed, int max) {
    return (int)(ct_rand_float(seed) * ((float)max + 0.99999f));
}
/* Primary-ray jitter seed, src/CUDA/Kernel.cu:225 (thread index == pixel). */
CT_HD uint32_t ct_primary_seed(uint32_t pixel, uint32_t frame) {
    return (pixel + frame * 147565741u) * 720898027u;
}
/* Shading seed, src/CUDA/Kernel.cu:144, computed in uint32 (wraps). */
CT_HD uint32_t ct_shade_seed(uint32_t pixel, uint32_t frame, uint32_t path_length) {
    return ct_wang_hash(pixel * 16789u + frame * 1791u + path_length * 720898027u);
}
/* The reference numbers its shade passes 0,2,3,4,5 (Kernel.cu:306,330): the
 * first bounce uses 0 and bounce b>0 uses b+1.  Kept for every depth. */
CT_HD uint32_t ct_path_length_key(int bounce) {
    return bounce == 0 ? 0u : (uint32_t)(bounce + 1);
}

/* --------------------------------------------------------- constants --- */
#define CT_PI_F        0x1.921fb6p+1f   /* float(pi)        = glm::pi<float>() */
#define CT_TWO_PI_F    0x1.921fb6p+2f   /* float(2 pi)      = glm::two_pi<float>() */
#define CT_INV_PI_F    0x1.45f306p-2f   /* float(1/pi)      = glm::one_over_pi<float>() */
#define CT_ROOT_PI_F   0x1.c5bf8ap+0f   /* float(sqrt(pi))  = glm::root_pi<float>() */

/* Cody–Waite splits (hi has trailing zero bits so k*hi is exact for |k|<2^8). */
#define CT_PIO2_HI   0x1.921fb6p+0f
#define CT_PIO2_MID -0x1.777a5cp-25f
#define CT_PIO2_LO  -0x1.ee59dap-50f
#define CT_TWO_OVER_PI 0x1.45f306p-1f
#define CT_LN2_HI    0x1.62e430p-1f
#define CT_LN2_LO   -0x1.05c610p-29f
#define CT_LOG2E     0x1.715476p+0f

/* --------------------------------------------------------- sin / cos --- */
/* Kernel polynomials on |r| <= pi/4: Taylor series to r^9 (sin) and r^10 (cos),
 * Horner with fma.  Truncation error < 3e-9 relative. */
CT_HD float ct_sin_poly(float r) {
    const float z = r * r;
    float p = 0x1.71de3ap-19f;                 /*  1/9!  */
    p = ct_fmaf(p, z, -0x1.a01a02p-13f);       /* -1/7!  */
    p = ct_fmaf(p, z,  0x1.111112p-7f);        /*  1/5!  */
    p = ct_fmaf(p, z, -0x1.555556p-3f);        /* -1/3!  */
    return ct_fmaf(p * z, r, r);
}
CT_HD float ct_cos_poly(float r) {
    const float z = r * r;
    float p = -0x1.27e4fcp-22f;                /* -1/10! */
    p = ct_fmaf(p, z,  0x1.a01a02p-16f);       /*  1/8!  */
    p = ct_fmaf(p, z, -0x1.6c16c2p-10f);       /* -1/6!  */
    p = ct_fmaf(p, z,  0x1.555556p-5f);        /*  1/4!  */
    p = ct_fmaf(p, z, -0.5f);                  /* -1/2!  */
    return ct_fmaf(p, z, 1.0f);
}
/* Reduce x = k*pi/2 + r.  Valid for |x| < 2^15 (plenty for path tracing
 * angles, which lie in [-2pi, 2pi]). */
CT_HD float ct_reduce_pio2(float x, int* q) {
    const float k = ct_rintf(x * CT_TWO_OVER_PI);
    float r = ct_fmaf(-k, CT_PIO2_HI, x);
    r = ct_fmaf(-k, CT_PIO2_MID, r);
    r = ct_fmaf(-k, CT_PIO2_LO, r);
    *q = (int)k;
    return r;
}
CT_HD float ct_sinf(float x) {
    int q;
    const float r = ct_reduce_pio2(x, &q);
    const float s = ct_sin_poly(r), c = ct_cos_poly(r);
    switch (q & 3) {
        case 0: return s;
        case 1: return c;
        case 2: return -s;
        default: return -c;
    }
}
CT_HD float ct_cosf(float x) {
    int q;
    const float r = ct_reduce_pio2(x, &q);
    const float s = ct_sin_poly(r), c = ct_cos_poly(r);
    switch (q & 3) {
        case 0: return c;
        case 1: return -s;
        case 2: return -c;
        default: return s;
    }
}

/* ----------------------------------------------------------------- exp --- */
/* 2^k for integer k in [-149, 127] (exact) */
CT_HD float ct_exp2i(int k) {
    if (k >= -126) return ct_u2f((uint32_t)(k + 127) << 23);
    return ct_u2f(1u << (k + 149));
}
/* exp(x): x = k ln2 + r, |r| <= ln2/2, Taylor to r^7. */
CT_HD float ct_expf(float x) {
    if (x != x) return x;
    if (x > 88.72284f) return ct_u2f(0x7f800000u);
    if (x < -103.97208f) return 0.0f;
    const float k = ct_rintf(x * CT_LOG2E);
    float r = ct_fmaf(-k, CT_LN2_HI, x);
    r = ct_fmaf(-k, CT_LN2_LO, r);
    float p = 0x1.a01a02p-13f;                 /* 1/7! */
    p = ct_fmaf(p, r, 0x1.6c16c2p-10f);        /* 1/6! */
    p = ct_fmaf(p, r, 0x1.111112p-7f);         /* 1/5! */
    p = ct_fmaf(p, r, 0x1.555556p-5f);         /* 1/4! */
    p = ct_fmaf(p, r, 0x1.555556p-3f);         /* 1/3! */
    p = ct_fmaf(p, r, 0.5f);
    p = ct_fmaf(p, r, 1.0f);
    p = ct_fmaf(p, r, 1.0f);
    const int ki = (int)k;
    /* scale in two steps so that results near the denormal range stay exact */
    if (ki > 127) return p * ct_exp2i(127) * ct_exp2i(ki - 127);
    if (ki < -126) return p * ct_exp2i(ki + 64) * ct_exp2i(-64);
    return p * ct_exp2i(ki);
}

/* ----------------------------------------------------------------- log --- */
/* log(x) = e ln2 + 2 atanh(s), s = (m-1)/(m+1), m in [sqrt(1/2), sqrt(2)). */
CT_HD float ct_logf(float x) {
    if (x != x) return x;
    if (x < 0.0f) return ct_u2f(0x7fc00000u);
    if (x == 0.0f) return ct_u2f(0xff800000u);
    if (ct_f2u(x) == 0x7f800000u) return x;
    uint32_t u = ct_f2u(x);
    int e = 0;
    if (u < 0x00800000u) {          /* denormal: scale by 2^24 */
        x = x * 16777216.0f;
        u = ct_f2u(x);
        e = -24;
    }
    e += (int)(u >> 23) - 127;
    float m = ct_u2f((u & 0x007fffffu) | 0x3f800000u);   /* [1, 2) */
    if (m > 1.41421354f) { m = m * 0.5f; e += 1; }        /* exact */
    const float s = (m - 1.0f) / (m + 1.0f);
    const float z = s * s;
    float p = 0x1.745d18p-4f;                  /* 1/11 */
    p = ct_fmaf(p, z, 0x1.c71c72p-4f);         /* 1/9  */
    p = ct_fmaf(p, z, 0x1.24924ap-3f);         /* 1/7  */
    p = ct_fmaf(p, z, 0x1.99999ap-3f);         /* 1/5  */
    p = ct_fmaf(p, z, 0x1.555556p-2f);         /* 1/3  */
    const float s2 = s + s;                    /* exact */
    const float lm = ct_fmaf(s2 * z, p, s2);   /* 2 atanh(s) */
    const float fe = (float)e;
    return ct_fmaf(fe, CT_LN2_HI, ct_fmaf(fe, CT_LN2_LO, lm));
}

/* pow(x, y) for x >= 0: exp(y log x).  Used only where the reference calls
 * powf with moderate exponents (Microfacet.cuh:173). */
CT_HD float ct_powf(float x, float y) {
    if (y == 0.0f) return 1.0f;
    if (x == 0.0f) return y > 0.0f ? 0.0f : ct_u2f(0x7f800000u);
    return ct_expf(y * ct_logf(x));
}

/* ---------------------------------------------------------------- acos --- */
/* asin(t) for |t| <= 0.5: Taylor series to t^21. */
CT_HD float ct_asin_poly(float t) {
    const float z = t * t;
    float p = 0x1.12ef3cp-7f;                  /* c10 */
    p = ct_fmaf(p, z, 0x1.3fde50p-7f);         /* c9  */
    p = ct_fmaf(p, z, 0x1.7a8788p-7f);         /* c8  */
    p = ct_fmaf(p, z, 0x1.c9999ap-7f);         /* c7  */
    p = ct_fmaf(p, z, 0x1.1c4ec4p-6f);         /* c6  */
    p = ct_fmaf(p, z, 0x1.6e8ba2p-6f);         /* c5  */
    p = ct_fmaf(p, z, 0x1.f1c71cp-6f);         /* c4  */
    p = ct_fmaf(p, z, 0x1.6db6dcp-5f);         /* c3  */
    p = ct_fmaf(p, z, 0x1.333334p-4f);         /* c2  */
    p = ct_fmaf(p, z, 0x1.555556p-3f);         /* c1  */
    return ct_fmaf(p * z, t, t);
}
CT_HD float ct_acosf(float x) {
    if (x != x || x > 1.0f || x < -1.0f) return ct_u2f(0x7fc00000u);
    if (ct_fabsf(x) <= 0.5f)
        return (CT_PIO2_HI - ct_asin_poly(x)) + CT_PIO2_MID;
    if (x > 0.0f) {
        const float t = ct_sqrtf((1.0f - x) * 0.5f);
        return 2.0f * ct_asin_poly(t);
    }
    const float t = ct_sqrtf((1.0f + x) * 0.5f);
    return (2.0f * CT_PIO2_HI - 2.0f * ct_asin_poly(t)) + 2.0f * CT_PIO2_MID;
}

/* --------------------------------------------------------------- atan2 --- */
/* atan(t) for |t| <= 1/2: Taylor series t - t^3/3 + ... to t^27
 * (truncation below 2e-9 relative). */
CT_HD float ct_atan_poly(float t) {
    const float z = t * t;
    float p = -0x1.2f684cp-5f;                 /* -1/27 */
    p = ct_fmaf(p, z,  0x1.47ae14p-5f);        /*  1/25 */
    p = ct_fmaf(p, z, -0x1.642c86p-5f);        /* -1/23 */
    p = ct_fmaf(p, z,  0x1.861862p-5f);        /*  1/21 */
    p = ct_fmaf(p, z, -0x1.af286cp-5f);        /* -1/19 */
    p = ct_fmaf(p, z,  0x1.e1e1e2p-5f);        /*  1/17 */
    p = ct_fmaf(p, z, -0x1.111112p-4f);        /* -1/15 */
    p = ct_fmaf(p, z,  0x1.3b13b2p-4f);        /*  1/13 */
    p = ct_fmaf(p, z, -0x1.745d18p-4f);        /* -1/11 */
    p = ct_fmaf(p, z,  0x1.c71c72p-4f);        /*  1/9  */
    p = ct_fmaf(p, z, -0x1.24924ap-3f);        /* -1/7  */
    p = ct_fmaf(p, z,  0x1.99999ap-3f);        /*  1/5  */
    p = ct_fmaf(p, z, -0x1.555556p-2f);        /* -1/3  */
    return ct_fmaf(p * z, t, t);
}
/* atan(a) for a in [0, 1]: above 1/2, atan(a) = pi/4 + atan((a-1)/(a+1)),
 * where a - 1 is exact (Sterbenz) and |(a-1)/(a+1)| <= 1/3. */
CT_HD float ct_atan01(float a) {
    if (a <= 0.5f) return ct_atan_poly(a);
    const float t = (a - 1.0f) / (a + 1.0f);
    return (0x1.921fb6p-1f + ct_atan_poly(t)) + -0x1.777a5cp-26f;
}
/* atan2(y, x) for finite arguments: octant reduction to [0, 1], then the
 * quadrant from the signs (atan2(+-0, -0) = +-pi, atan2(+-0, +0) = +-0). */
CT_HD float ct_atan2f(float y, float x) {
    const float ax = ct_fabsf(x), ay = ct_fabsf(y);
    float r;
    if (ay == 0.0f && ax == 0.0f) r = 0.0f;
    else if (ay <= ax) r = ct_atan01(ay / ax);
    else r = (CT_PIO2_HI - ct_atan01(ax / ay)) + CT_PIO2_MID;
    if (ct_f2u(x) >> 31) r = (0x1.921fb6p+1f - r) + -0x1.777a5cp-24f;
    return (ct_f2u(y) >> 31) ? -r : r;
}

/* ------------------------------------------------------------ textures --- */
/* Nearest-texel address of SceneData::getTextureColor (SceneData.cuh:81-97):
 * wrap each coordinate into [0, 1) by fmod (x - trunc(x): exact, and the
 * sign of a zero result does not matter below), add 1 if negative, then
 * uint(x * (width - 1)).  Returns the texel index inside the texture.  The
 * normal-map lookup (SceneData.cuh:99-114) skips the fmod, which reads
 * outside the texture for coordinates above 1; it shares this wrap here. */
CT_HD float ct_wrap01(float x) {
    const float a = ct_fabsf(x);
    float f;
    if (a < 8388608.0f) f = x - (float)(int32_t)x;  /* |x| < 2^23: trunc via int */
    else f = 0.0f;                                  /* already an integer */
    if (f < 0.0f) f += 1.0f;
    return f;
}
CT_HD uint32_t ct_texel_index(uint32_t width, uint32_t height, float u, float v) {
    const float x = ct_wrap01(u), y = ct_wrap01(v);
    const uint32_t ix = (uint32_t)(x * (float)(width - 1));
    const uint32_t iy = (uint32_t)(y * (float)(height - 1));
    return ix + iy * width;
}
/* Sky texture coordinates of a miss direction (Kernel.cu:96-100). */
CT_HD void ct_sky_uv(float dx, float dy, float dz, float* u, float* v) {
    *u = 1.0f + ct_atan2f(dx, -dz) * CT_INV_PI_F * 0.5f;
    *v = 1.0f - ct_acosf(dy) * CT_INV_PI_F;
}

/* ------------------------------------------------------- erf / erfinv --- */
/* Abramowitz & Stegun 7.1.26, in the evaluation order of Microfacet.cuh:69-82. */
CT_HD float ct_erff(float x) {
    float sign = 1.0f;
    if (x < 0.0f) sign = -1.0f;
    x = ct_fabsf(x);
    const float t = 1.0f / (1.0f + 0.3275911f * x);
    const float y = 1.0f -
        (((((1.061405429f * t + -1.453152027f) * t) + 1.421413741f) * t + -0.284496736f) * t + 0.254829592f) *
        t * ct_expf(-x * x);
    return sign * y;
}
/* M. Giles, "Approximating the erfinv function" (GPU Computing Gems, 2011),
 * single-precision branch pair, as used at Microfacet.cuh:36-67. */
CT_HD float ct_erfinvf(float x) {
    float w, p;
    x = x < -0.999f ? -0.999f : (x > 0.999f ? 0.999f : x);
    w = -ct_logf((1.0f - x) * (1.0f + x));
    if (w < 5.0f) {
        w = w - 2.5f;
        p = 2.81022636e-08f;
        p = 3.43273939e-07f + p * w;
        p = -3.5233877e-06f + p * w;
        p = -4.39150654e-06f + p * w;
        p = 0.00021858087f + p * w;
        p = -0.00125372503f + p * w;
        p = -0.00417768164f + p * w;
        p = 0.246640727f + p * w;
        p = 1.50140941f + p * w;
    } else {
        w = ct_sqrtf(w) - 3.f;
        p = -0.000200214257f;
        p = 0.000100950558f + p * w;
        p = 0.00134934322f + p * w;
        p = -0.00367342844f + p * w;
        p = 0.00573950773f + p * w;
        p = -0.0076224613f + p * w;
        p = 0.00943887047f + p * w;
        p = 1.00167406f + p * w;
        p = 2.83297682f + p * w;
    }
    return p * x;
}

#endif /* CT_FP_H */
