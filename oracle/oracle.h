/*
 * oracle.h — CPU oracle of the wavefront path-tracing loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: the tests, the
 * smoke test and the benchmark's CPU-baseline leg load it to compute the
 * expected values.  The product library (cudatracer_amd) never includes,
 * links or calls anything here.
 *
 * It restates, in scalar C++17 templated on float and double, the reference's
 * traversal, builder and shading (file:line citations in oracle.cpp).  The
 * float instantiation uses the hand-written transcendental functions of
 * include/ct_fp.h so that it is bit-comparable with the HIP kernels; the
 * double instantiation uses libm and is the accuracy reference (rel-MSE).
 */
#ifndef CT_ORACLE_H
#define CT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "../include/cudatracer.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_scene or_scene;

/* Scene from TriangleList arrays (vertices are indexed through indices). */
or_scene* or_scene_create(uint32_t n_vertices, const float* vertices, const float* normals, uint32_t n_tris,
                          const uint32_t* indices, const float* center_normals, const uint32_t* mat_idx,
                          uint32_t n_lights, const uint32_t* light_idx, uint32_t n_materials,
                          const ct_material* materials, int32_t sky_enabled, const float* sky);
void or_scene_destroy(or_scene* s);
/* Texture coordinates (2 per vertex, NULL = zeros), textures (copied) and the
 * sky texture (-1 = constant sky).  Returns 0, or 1 for an index out of range. */
int or_scene_set_textures(or_scene* s, const float* texcoords, uint32_t n_textures, const ct_texture* textures,
                          int32_t sky_texture);
/* The texel SceneData::getTextureColor returns for (u, v). */
int or_texel(const or_scene* s, int32_t texture, float u, float v, float* rgba);
float or_atan2f(float y, float x);
void or_sky_uv(float dx, float dy, float dz, float* u, float* v);

/* Build the binary SAH tree by the reference's 30-plane rescan and collapse
 * it to the 4-wide tree.  Returns 0 on success. */
int or_scene_build(or_scene* s);
/* Use the given trees instead of building (the trees of an upload, which
 * may have been edited; traversed as given).  prim_indices: n_triangles
 * entries.  Returns 0, or 1 for a missing array or an index out of range. */
int or_scene_set_trees(or_scene* s, uint32_t n_bvh, const ct_bvh_node* bvh, uint32_t n_mbvh, const ct_mbvh_node* mbvh,
                       const uint32_t* prim_indices);
/* Copy the trees out (sizes first with NULL buffers). */
int or_scene_tree_sizes(const or_scene* s, uint32_t* n_bvh, uint32_t* n_mbvh, uint32_t* n_prims);
int or_scene_tree_export(const or_scene* s, ct_bvh_node* bvh, ct_mbvh_node* mbvh, uint32_t* prim_indices);

/* accel: 0 = binary tree, 1 = 4-wide tree, 2 = brute force over all triangles.
 * precision: 32 or 64.  t_out gets the hit distance as double (float values
 * widened exactly for precision 32).  counters (may be NULL): per ray
 * [node tests, triangle tests]. */
int or_trace(const or_scene* s, int accel, int precision, uint32_t n, const float* origins, const float* dirs,
             float tri_eps, double* t_out, int32_t* prim_out, uint32_t* counters, int threads);
int or_trace_shadow(const or_scene* s, int accel, int precision, uint32_t n, const float* origins,
                    const float* dirs, const float* tmax, float tri_eps, uint8_t* occluded_out,
                    uint32_t* counters, int threads);

typedef struct or_render_stats {
    uint64_t primary_rays, extension_rays, shadow_rays, samples;
    uint64_t node_tests_closest, tri_tests_closest, node_tests_shadow, tri_tests_shadow;
} or_render_stats;

/* Render rows [row_begin, row_end) of a width x height image: spp samples per
 * pixel with frame indices frame_base + s, added to accum (4 per pixel, full
 * image layout; .w counts samples).  precision 32 accumulates in float
 * (accum32), 64 in double (accum64); the other pointer may be NULL. */
int or_render(const or_scene* s, int precision, const ct_camera* cam, const ct_frame_opts* opts, int32_t width,
              int32_t height, uint32_t frame_base, int32_t spp, int32_t row_begin, int32_t row_end,
              float* accum32, double* accum64, int threads, or_render_stats* stats);

/* Individual numeric routines, exported for known-answer tests. */
float or_ct_sinf(float x);
float or_ct_cosf(float x);
float or_ct_expf(float x);
float or_ct_logf(float x);
float or_ct_acosf(float x);
float or_ct_powf(float x, float y);
float or_ct_erff(float x);
float or_ct_erfinvf(float x);
uint32_t or_wang_hash(uint32_t s);
uint32_t or_xorshift(uint32_t s);
/* Microfacet helpers (precision 64, isotropic alpha): D(wh), Lambda(w),
 * sample the visible normal wm for wo; type CT_MAT_GGX or CT_MAT_BECKMANN. */
double or_mf_D(int type, double alpha, const double* wh);
double or_mf_lambda(int type, double alpha, const double* w);
void or_mf_sample_vndf32(int type, float alpha, const float* wo, float r1, float r2, float* wm_out);
/* Fresnel for a dielectric, relative IOR eta = n_t / n_i. */
double or_fresnel_dielectric(double cos_i, double eta);
/* glm::refract(I, N, eta) as the dielectric branch uses it (precision 32 or
 * 64): writes the direction, returns 1 unless total internal reflection. */
int or_refract(int precision, const float* I, const float* N, float eta, double* out);
/* BSDF evaluate/sample in the local frame for tests: returns f (rgb=1 albedo)
 * and pdf.  mode 0 evaluate(wo, wi), mode 1 sample(wo, r1, r2) -> wi. */
int or_bsdf_local(int type, float roughness, const float* wo, const float* wi_in, float r1, float r2,
                  float* wi_out, float* f_out, float* pdf_out);

#ifdef __cplusplus
}
#endif
#endif
