/*
 * cudatracer.h — C ABI of the MI355X (gfx950) wavefront path-tracing loop.
 *
 * This is the drop-in boundary that replaces the reference's per-frame entry
 *   cudaError_t launchKernels(cudaArray_const_t, Params&, int rayBufferSize)
 *   (src/CUDA/Kernel.cuh:74, defined src/CUDA/Kernel.cu:355-372)
 * and the device-memory ownership of Application::allocate{Buffers,Triangles}
 * (src/Application.cpp:141-193), Application::updateMaterials
 * (src/Application.cpp:222-225) and Params::reset (src/CUDA/Kernel.cuh:66-71).
 *
 * Conventions
 *   - plain C types, host pointers and sizes only; no HIP or torch types;
 *   - every entry point returns an int status (CT_OK == 0) and never exits the
 *     process (the reference's cuda() macro calls exit(), CudaAssert.h:7-21);
 *     ct_last_error() gives a message for the last failure on a context;
 *   - one context is owned by one host thread at a time (the reference's UI
 *     thread races with launchKernels, Application.cpp:114-120; here callers
 *     serialise, and the render server owns its context on its own thread).
 */
#ifndef CUDATRACER_H
#define CUDATRACER_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CT_ABI_VERSION 5

/* ---------------------------------------------------------------- status */
enum {
    CT_OK = 0,
    CT_ERR_INVALID_ARG = 1,    /* bad pointer, size or option                 */
    CT_ERR_OUT_OF_MEMORY = 2,  /* device allocation failed or budget exceeded  */
    CT_ERR_DEVICE = 3,         /* HIP runtime error                           */
    CT_ERR_NO_SCENE = 4,       /* trace before ct_scene_upload                */
    CT_ERR_ADMISSION = 5,      /* render request can never fit the budget     */
    CT_ERR_NOT_FOUND = 6,      /* unknown request id                          */
    CT_ERR_STACK_OVERFLOW = 7, /* traversal stack bound exceeded (never with the
                                  builder's depth limit; reported, not UB)    */
    CT_ERR_CANCELLED = 8,      /* request cancelled before completion         */
    CT_ERR_INVALID_SCENE = 9   /* 0 triangles, NaN vertices, bad indices      */
};

/* ------------------------------------------------------------- materials */
/* MaterialType, src/Core/Material.cuh:13-27 (same numeric values). */
enum {
    CT_MAT_NONE = -1,
    CT_MAT_LIGHT = 0,
    CT_MAT_SPECULAR = 1,
    CT_MAT_FRESNEL = 2,     /* smooth dielectric                          */
    CT_MAT_LAMBERTIAN = 3,
    CT_MAT_BECKMANN = 4,
    CT_MAT_GGX = 5,
    CT_MAT_TROWBRIDGE = 6,  /* GGX and Trowbridge-Reitz are one distribution */
    CT_MAT_FRESNEL_BECKMANN = 7,
    CT_MAT_FRESNEL_GGX = 8,
    CT_MAT_FRESNEL_TROWBRIDGE = 9
};

/* Field order of struct Material, src/Core/Material.cuh:29-41
 * (albedo doubles as emission for lights).  diffuse_tex >= 0 replaces the
 * albedo by the texture colour at the hit's texture coordinates
 * (Kernel.cu:146); normal_tex >= 0 perturbs the shading normal
 * (Kernel.cu:109-112).  mask_tex and displace_tex are carried but unused, as
 * in the reference's shading. */
typedef struct ct_material {
    float albedo[3];
    float refract_idx;
    int32_t diffuse_tex, normal_tex, mask_tex, displace_tex;
    int32_t type;
    float roughness;
    float absorption[3];
} ct_material;

/* --------------------------------------------------- acceleration nodes */
/* BVHNode, src/BVH/BVHNode.cuh:29-40 (32 B). count == -1: interior with
 * children at left_first and left_first+1; count >= 0: leaf over
 * prim_indices[left_first, left_first+count).  Node 1 is unused padding. */
typedef struct ct_bvh_node {
    float min[3];
    int32_t left_first;
    float max[3];
    int32_t count;
} ct_bvh_node;

/* MBVHNode, src/BVH/MBVHNode.cuh:28-63 (128 B). counts[i] == -1: child i is
 * MBVH node childs[i]; counts[i] > 0: leaf over prim_indices[childs[i], +counts[i]);
 * counts[i] == 0: empty slot (bounds 1e34 / -1e34). */
typedef struct ct_mbvh_node {
    float bminx[4], bmaxx[4], bminy[4], bmaxy[4], bminz[4], bmaxz[4];
    int32_t childs[4];
    int32_t counts[4];
} ct_mbvh_node;

/* -------------------------------------------------------------- scene */
/* A texture: core::Surface's float buffer (src/Utils/Surface.cpp:33-100),
 * width * height RGBA texels, row 0 at the bottom of the image (FreeImage's
 * scan-line order), texel (x, y) at rgba[4 * (x + y * width)]. */
typedef struct ct_texture {
    uint32_t width, height;
    const float* rgba;
} ct_texture;

/* Borrowed host arrays in the layout of TriangleList (src/Core/TriangleList.h:59-73)
 * plus the two trees (BVHTree::m_BVHPool / m_PrimitiveIndices and
 * MBVHTree::m_Tree[0..m_FinalPtr)).  The library copies everything it needs
 * to the device during ct_scene_upload and keeps no pointer to these. */
typedef struct ct_scene_view {
    uint32_t n_vertices;
    const float* vertices;        /* 3 * n_vertices */
    const float* normals;         /* 3 * n_vertices (normalised)            */
    const float* texcoords;       /* 2 * n_vertices, may be NULL            */
    uint32_t n_triangles;
    const uint32_t* indices;      /* 3 * n_triangles                        */
    const float* center_normals;  /* 3 * n_triangles                        */
    const uint32_t* material_idx; /* n_triangles                            */
    uint32_t n_lights;
    const uint32_t* light_indices;/* n_lights triangle ids                  */
    uint32_t n_materials;
    const ct_material* materials;
    const uint32_t* prim_indices; /* n_triangles                            */
    uint32_t n_bvh_nodes;         /* 0 if the binary tree is not supplied    */
    const ct_bvh_node* bvh_nodes;
    uint32_t n_mbvh_nodes;        /* 0 if the quad tree is not supplied      */
    const ct_mbvh_node* mbvh_nodes;
    int32_t sky_enabled;          /* SceneData::skyboxEnabled               */
    float sky[3];                 /* constant environment radiance (a 1x1 sky texture) */
    /* TriangleList::createTextureBuffer (TriangleList.cpp:148-178) */
    uint32_t n_textures;
    const ct_texture* textures;
    int32_t sky_texture;          /* SceneData::skyboxTexture: -1 = the constant sky;
                                     ignored without textures (a zeroed view keeps
                                     the constant sky)                            */
} ct_scene_view;

/* Camera state, src/Core/Camera.h:81-92; plane_distance = tan(fov/2). */
typedef struct ct_camera {
    float position[3];
    float forward[3];
    float up[3];
    float plane_distance;
} ct_camera;

/* ------------------------------------------------------------ options */
enum { CT_ACCEL_BVH2 = 0, CT_ACCEL_BVH4 = 1 };
/* CT_BSDF_REFERENCE: SampleBSDF of src/Core/BSDF.h:33-63 (roughness < 0.1 is a
 * mirror, everything else Lambert; material type ignored except Light).
 * CT_BSDF_TYPED: dispatch on the material type (Lambert, mirror, smooth
 * dielectric, GGX and Beckmann microfacet reflection). */
enum { CT_BSDF_REFERENCE = 0, CT_BSDF_TYPED = 1 };
/* CT_TRAV_EXACT: reference node order and culling, bit-exact hit ids.
 * CT_TRAV_FAST: wide compressed tree (8 children, 128-byte nodes) for either
 * accel setting; hit distances equal, ids equal except exact t-ties.
 * CT_TRAV_COMPACT: the quad tree itself (MBVHNode.cuh:28-63: same nodes,
 * children and leaves) re-laid as 64-byte nodes with 8-bit child boxes on a
 * per-node grid, rounded outwards; CT_ACCEL_BVH4 only.  Children are visited
 * nearest first as in exact mode.  With primitive boxes padded as the
 * builders pad them (1e-3) hit distances equal, ids equal except exact
 * t-ties; a hand-made tree with unpadded boxes can differ on rays that
 * graze a box within rounding. */
enum { CT_TRAV_EXACT = 0, CT_TRAV_FAST = 1, CT_TRAV_COMPACT = 2 };
/* Light-selection pdf in the MIS weights.  CT_LIGHTPDF_REFERENCE multiplies
 * the per-triangle solid-angle pdf by the light count as the reference does
 * (Kernel.cu:128,184), which biases direct light by 1/count^2 when a scene
 * has more than one light triangle; CT_LIGHTPDF_UNIFORM divides by it (the
 * pdf of uniform light selection, unbiased).  Both agree for one light. */
enum { CT_LIGHTPDF_REFERENCE = 0, CT_LIGHTPDF_UNIFORM = 1 };

typedef struct ct_frame_opts {
    int32_t max_depth;        /* path segments (reference: 5, Kernel.cu:320)  */
    int32_t spp;              /* samples per pixel per call (reference: 1)    */
    int32_t rr_start;         /* Russian roulette from this bounce; -1 = off  */
    int32_t accel;            /* CT_ACCEL_*                                   */
    int32_t bsdf_mode;        /* CT_BSDF_*                                    */
    int32_t traversal;        /* CT_TRAV_*                                    */
    int32_t shadow;           /* SceneData::shadow (NEE on)                   */
    int32_t indirect;         /* SceneData::indirect (emission on light hits) */
    float normal_epsilon;     /* SceneData defaults 1e-5 / 1e-5 / 1e-3        */
    float dist_epsilon;
    float triangle_epsilon;
    int32_t sort_by_material; /* material-key sort before shading             */
    int32_t light_pdf;        /* CT_LIGHTPDF_*                                */
} ct_frame_opts;

typedef struct ct_frame_stats {
    uint64_t primary_rays;    /* camera rays traced                           */
    uint64_t extension_rays;  /* closest-hit rays after the first segment     */
    uint64_t shadow_rays;     /* any-hit rays traced                          */
    uint64_t samples;         /* path samples completed                       */
    double gpu_ms;            /* device time of the frame's chunks (HIP events; chunks on
                                 the two lanes overlap, so this can exceed wall time) */
    double extend_ms, shade_ms, connect_ms, other_ms; /* per-stage device time, if profiled;
                                 stages of different chunks overlap.  extend_ms:
                                 the traversal launches of the bounces (each
                                 traces bounce b's extension rays and bounce
                                 b - 1's shadow rays); connect_ms: the launches
                                 of the last bounce's shadow rays             */
    uint32_t chunks;          /* wavefront chunks the frame was split into    */
    uint32_t extend_launches; /* traversal launches of the bounces (chunks x depth;
                                 plus one per chunk for the last shadow rays
                                 when NEE is on; with roulette, the bounces
                                 up to the tail's first)                      */
    double tail_ms;           /* with roulette (rr_start >= 0): the device time of
                                 the launches that run each chunk's remaining
                                 bounces once roulette has thinned its paths
                                 (a lane per path: shading, the light sample's
                                 shadow ray and the next ray traced in turn) */
} ct_frame_stats;

/* Fill with the reference defaults (depth 5, 1 spp, no RR, BVH4, reference BSDF,
 * exact traversal, NEE and indirect on, epsilons 1e-5/1e-5/1e-3, no sort,
 * reference light pdf). */
void ct_frame_opts_default(ct_frame_opts* opts);

/* ------------------------------------------------------------ context */
typedef struct ct_ctx ct_ctx;

typedef struct ct_device_opts {
    int32_t device;           /* HIP device ordinal                          */
    uint64_t mem_budget;      /* bytes the context may allocate; 0 = 90% of free */
    int32_t block_size;       /* 0 = autotuned/default                       */
    int32_t stack_lds;        /* per-lane LDS stack entries, 0 = default      */
    int32_t chunk_paths;      /* wavefront chunk size in paths, 0 = default   */
} ct_device_opts;

int ct_abi_version(void);
int ct_ctx_create(const ct_device_opts* opts, ct_ctx** out);
void ct_ctx_destroy(ct_ctx* ctx);
const char* ct_last_error(const ct_ctx* ctx);
int ct_device_name(const ct_ctx* ctx, char* buf, size_t len);

/* Application::allocateTriangles (Application.cpp:156-193) + node export. */
int ct_scene_upload(ct_ctx* ctx, const ct_scene_view* scene);
/* BVHTree::constructBVH + MBVHTree::constructBVH (BVHTree.cpp:22-58,
 * MBVHTree.cpp:15-30, run by Application::init, Application.cpp:132-135) on
 * the device, then the upload of ct_scene_upload.  The view's prim_indices
 * and trees are ignored.  The binary tree, quad tree and primitive order are
 * byte-identical to the host build of the same view (the reference's
 * single-threaded node numbering).  CT_BUILD_WIDE also collapses the wide
 * tree of CT_TRAV_FAST (on the host, from the device-built trees); without it
 * CT_TRAV_FAST is refused for this scene.  stats may be NULL.  On failure no
 * scene is loaded. */
enum { CT_BUILD_WIDE = 1 };
typedef struct ct_build_stats {
    uint32_t n_bvh_nodes, n_mbvh_nodes;
    uint32_t levels;      /* tree levels processed                         */
    uint32_t max_depth;   /* BVHTree::m_MaxDepth (root depth 2)            */
    uint32_t long_chains; /* partition slots whose swap chain exceeded 256 steps */
    uint32_t jump_rounds; /* pointer-jumping rounds spent on those chains   */
    float device_ms;      /* first to last build kernel (read-backs included) */
    float total_ms;       /* whole call: checks, staging, build, uploads    */
} ct_build_stats;
int ct_scene_upload_build(ct_ctx* ctx, const ct_scene_view* scene, uint32_t flags, ct_build_stats* stats);
/* Node counts of the current scene's trees (0 for a tree not present). */
int ct_scene_tree_sizes(ct_ctx* ctx, uint32_t* n_bvh, uint32_t* n_mbvh, uint32_t* n_prims);
/* Copy the current scene's trees and primitive order back (BVHTree::m_BVHPool,
 * MBVHTree::m_Tree, m_PrimitiveIndices layouts); NULL outputs are skipped. */
int ct_scene_read_trees(ct_ctx* ctx, ct_bvh_node* bvh, uint32_t cap_bvh, ct_mbvh_node* mbvh, uint32_t cap_mbvh,
                        uint32_t* prim_indices, uint32_t cap_prims);
/* Application::updateMaterials (Application.cpp:222-225). */
int ct_material_update(ct_ctx* ctx, uint32_t index, const ct_material* mat);
/* Resize callback + allocateBuffers (Application.cpp:64-75,141-154); resets. */
int ct_resize(ct_ctx* ctx, int32_t width, int32_t height);
/* Params::reset (Kernel.cuh:66-71): zero the accumulator and sample count. */
int ct_reset(ct_ctx* ctx);

/* launchKernels (Kernel.cu:355-372): trace opts->spp samples per pixel from
 * the camera, add them to the persistent accumulator and advance the sample
 * counter.  opts may be NULL (reference defaults).  stats may be NULL. */
int ct_trace_frame(ct_ctx* ctx, const ct_camera* cam, const ct_frame_opts* opts,
                   ct_frame_stats* stats);

/* Accumulator access (replaces the GL surface, CudaRenderer.cpp:36-42).
 * rgba: 4 floats per pixel; .w counts samples.  samples may be NULL. */
int ct_read_accum(ct_ctx* ctx, float* rgba, size_t n_floats, int32_t* samples);
/* Checkpoint/resume: restore sums and sample count written by ct_read_accum. */
int ct_write_accum(ct_ctx* ctx, const float* rgba, size_t n_floats, int32_t samples);
/* draw_framebuffer (Kernel.cu:258-270): rgb / samples, then x^(1/2.2). */
int ct_read_image(ct_ctx* ctx, float* rgba, size_t n_floats);
/* The same image as 8-bit RGBA for a viewer's texture upload (a quarter of
 * the PCIe bytes of the RGBA32F surface the reference writes): each channel
 * clamp(v, 0, 1) * 255 + 0.5 truncated, alpha 255.  n_bytes >= 4 * pixels. */
int ct_read_image_rgba8(ct_ctx* ctx, uint8_t* rgba, size_t n_bytes);

/* Ray-level entry points (the extend / connect stages on their own).
 * origins, dirs: 3 floats per ray (host).  Closest hit writes t (1e34 on a
 * miss) and the primitive id (-1 on a miss); any-hit writes 1 if something is
 * hit in (triangle_epsilon, tmax).  opts selects accel/traversal/epsilon. */
int ct_trace_rays(ct_ctx* ctx, const float* origins, const float* dirs, uint32_t n,
                  const ct_frame_opts* opts, float* t_out, int32_t* prim_out);
int ct_trace_shadow_rays(ct_ctx* ctx, const float* origins, const float* dirs,
                         const float* tmax, uint32_t n, const ct_frame_opts* opts,
                         uint8_t* occluded_out);
/* Device-pointer form for benchmarking the closest-hit kernel alone:
 * d_rays_o, d_rays_d: float4 per ray (xyz used); d_hits: float2 (t, prim bits).
 * kernel_ms receives the kernel's device time. */
int ct_trace_rays_device(ct_ctx* ctx, const void* d_rays_o, const void* d_rays_d,
                         uint32_t n, const ct_frame_opts* opts, void* d_hits,
                         float* kernel_ms);

/* ct_trace_rays_device without any device allocation (no counterpart in the
 * reference: for callers that cannot allocate, such as code under HIP graph
 * capture, tight serving loops and harnesses that forbid hipMalloc).  The
 * caller passes the rays and hits as for ct_trace_rays_device and must have
 * uploaded the scene; the kernel's counters, error flags and traversal-stack
 * spill live in one 32 MiB + 64 KiB device array per GPU, held until the
 * process ends.  Limits: a traversal stack of 48 entries (deeper trees return
 * CT_ERR_STACK_OVERFLOW, those rays reported as misses), and launch shapes
 * whose spill does not fit the array are refused (CT_ERR_INVALID_ARG).  All
 * calls on one GPU share the array: calls that overlap on one GPU race and
 * must be serialised by the caller.  Same results as ct_trace_rays_device. */
int ct_trace_rays_device_noalloc(ct_ctx* ctx, const void* d_rays_o, const void* d_rays_d,
                                 uint32_t n, const ct_frame_opts* opts, void* d_hits,
                                 float* kernel_ms);

/* --------------------------------------------------- batched multi-view */
/* Many camera views of uneven resolution and spp rendered by one set of
 * wavefront launches (one persistent path queue, hipGraph-captured bounce
 * loop).  Each view's accumulator is written to accum_out (host, 4 floats
 * per pixel, sums over the view's spp samples; frame_base is the first
 * frame index, so a view equals ct_trace_frame of the same spp from a reset
 * context when frame_base == 0). */
typedef struct ct_view {
    ct_camera camera;
    int32_t width, height, spp;
    uint32_t frame_base;
    float* accum_out;
} ct_view;

typedef struct ct_batch_stats {
    uint64_t primary_rays, extension_rays, shadow_rays, samples;
    double gpu_ms;
    int32_t chunks;
} ct_batch_stats;

int ct_render_batch(ct_ctx* ctx, const ct_view* views, int32_t n_views,
                    const ct_frame_opts* opts, ct_batch_stats* stats);

/* ------------------------------------------------------------ autotuner */
typedef struct ct_tune_result {
    int32_t block_size;       /* threads per block of the wavefront kernels  */
    int32_t stack_lds;        /* per-lane LDS stack entries                  */
    int32_t chunk_paths;      /* paths per wavefront chunk                   */
    double best_ms;           /* time of the probe workload with the result  */
    int32_t trials;
} ct_tune_result;
/* Time a probe workload of the uploaded scene over the candidate grid (block
 * size x LDS stack depth, then chunk size) and keep the fastest configuration
 * on the context.  The probe view is sampled often enough that every timed
 * run has at least 2M paths (8M for the chunk size), and the current setting
 * is kept unless a candidate is at least 2% faster. */
int ct_autotune(ct_ctx* ctx, const ct_camera* cam, const ct_frame_opts* opts,
                int32_t probe_width, int32_t probe_height, ct_tune_result* out);

/* -------------------------------------------------------------- serving */
typedef struct ct_server ct_server;

typedef struct ct_server_opts {
    uint64_t mem_budget;      /* device bytes of the server's state, 0 = 8 GiB: the
                                 shared path-state queues (2 x chunk_paths path
                                 slots of 288 B) and, in the rest, the arena that
                                 holds admitted requests' accumulators          */
    int32_t chunk_paths;      /* paths per shared wavefront chunk, 0 = 2^23  */
    int32_t max_active;       /* cap on concurrently admitted requests, 0 = none */
} ct_server_opts;

typedef struct ct_request {
    ct_camera camera;
    int32_t width, height, spp;
    uint32_t frame_base;
    ct_frame_opts opts;       /* every field but spp applies to this request;
                                 requests share a wavefront chunk only with
                                 requests of equal options                    */
} ct_request;

typedef struct ct_result {
    int32_t status;           /* CT_OK, CT_ERR_CANCELLED, ...                */
    int32_t width, height, spp;
    double submit_ms, admit_ms, done_ms;  /* server clock (ms since creation) */
    uint64_t rays;            /* rays traced for this request                */
} ct_result;

/* The server thread owns ctx until ct_server_destroy returns.  Admission:
 * a request is admitted in arrival order once its accumulator (width x height
 * x 16 bytes) fits the arena; its path samples then take slots of the shared
 * queues chunk by chunk.  A request whose accumulator is larger than the
 * whole arena is refused at submit with CT_ERR_ADMISSION; invalid sizes,
 * cameras or options with CT_ERR_INVALID_ARG (ct_server_last_error gives the
 * reason). */
int ct_server_create(ct_ctx* ctx, const ct_server_opts* opts, ct_server** out);
int ct_server_submit(ct_server* srv, const ct_request* req, uint64_t* id);
/* Message of the server's last refused call (valid until the next call). */
const char* ct_server_last_error(ct_server* srv);
/* Non-blocking: CT_OK with done=1 when finished (accumulator copied to
 * accum_out, 4 floats per pixel, if not NULL), done=0 while pending. */
int ct_server_poll(ct_server* srv, uint64_t id, int32_t* done, ct_result* res,
                   float* accum_out, size_t n_floats);
int ct_server_wait(ct_server* srv, uint64_t id, ct_result* res, float* accum_out,
                   size_t n_floats);
int ct_server_cancel(ct_server* srv, uint64_t id);
double ct_server_clock_ms(const ct_server* srv);
void ct_server_destroy(ct_server* srv);

/* ------------------------------------------------- host scene building */
/* The host classes (TriangleList, BVHTree, MBVHTree, Camera, Material) are
 * C++ (cudatracer_amd/csrc/host); this block exposes them for C and ctypes
 * callers.  A ct_host_scene owns a TriangleList and both trees. */
typedef struct ct_host_scene ct_host_scene;

enum {
    CT_SCENE_C1_BOX_SPHERE = 1,   /* closed box + 20,480-tri icosphere + quad light */
    CT_SCENE_C2_ATRIUM = 2,       /* ~260k-tri procedural atrium                     */
    CT_SCENE_C3_DISPLACED = 3,    /* 5.2M displaced icosphere + 2M noisy ground      */
    CT_SCENE_C2_TEXTURED = 4      /* config 2 with diffuse and normal textures and an
                                     equirectangular sky texture                     */
};

int ct_host_scene_create(ct_host_scene** out);
int ct_host_scene_generate(int32_t config, uint64_t seed, ct_host_scene** out);
void ct_host_scene_destroy(ct_host_scene* s);
/* TriangleList::addMaterial (TriangleList.cpp:52-61): forces roughness = 1
 * unless keep_roughness != 0.  Returns the material index (>= 0) or -status. */
int ct_host_scene_add_material(ct_host_scene* s, const ct_material* m, int32_t keep_roughness);
/* TriangleList::addTriangle (TriangleList.cpp:25-50); also registers the
 * triangle as a light when its material is a Light. */
int ct_host_scene_add_triangle(ct_host_scene* s, const float p[9], const float n[9],
                               uint32_t material, const float uv[6]);
/* TriangleList::loadModel (src/Core/TriangleList.h:49, TriangleList.cpp:63-146)
 * for Wavefront OBJ files: positions mat * (p * scale, 1), normals
 * mat * (n, 0); mat is a column-major 4x4 (glm layout), NULL = identity;
 * normalize != 0 first fits the model into [-1, 1] (centred, divided by half
 * its largest extent).  Polygons become triangle fans with reversed winding,
 * identical corners are shared, missing normals are generated smooth.
 * material < 0: the faces take their MTL materials (mtllib / usemtl), made
 * as TriangleList::processMesh makes them (TriangleList.cpp:227-345) with
 * the OBJ importer's defaults (diffuse 0.6) for faces without one; map_Kd,
 * norm, map_d and map_Disp textures load through ct_host_scene_load_texture.
 * Triangles of a Light material become lights.  CT_ERR_INVALID_ARG if the
 * file cannot be opened, CT_ERR_INVALID_SCENE if it does not parse;
 * ct_host_scene_last_error gives the reason. */
int ct_host_scene_load_model(ct_host_scene* s, const char* path, float scale, const float mat[16], int32_t material,
                             int32_t normalize);
const char* ct_host_scene_last_error(const ct_host_scene* s);
/* BVHTree::constructBVH + MBVHTree::constructBVH. */
int ct_host_scene_build(ct_host_scene* s);
int ct_host_scene_view(const ct_host_scene* s, ct_scene_view* out);
int ct_host_scene_set_sky(ct_host_scene* s, int32_t enabled, const float rgb[3]);
/* TriangleList::loadTexture (TriangleList.cpp:377-391) from memory: a copy of
 * width * height RGBA floats, row 0 at the bottom.  Returns the texture index
 * (>= 0) or -status. */
int ct_host_scene_add_texture(ct_host_scene* s, uint32_t width, uint32_t height, const float* rgba);
/* TriangleList::loadTexture from a file: binary PPM (P6, maxval <= 255) as
 * 8-bit colour / 255 with alpha 1 (FreeImage's 32-bit conversion), PFM (PF colour, Pf grey) as floats with
 * alpha 1; rows stored bottom-up as the reference's FreeImage load does.  A
 * path loaded before returns its earlier index.  Returns the index or
 * -status (ct_host_scene_last_error gives the reason).  FreeImage's other
 * formats (PNG, JPEG, Radiance HDR) are not decoded. */
int ct_host_scene_load_texture(ct_host_scene* s, const char* path);
/* Application::loadSkybox (Application.cpp:79-89): texture index (-1 = none;
 * the constant sky colour of ct_host_scene_set_sky applies). */
int ct_host_scene_set_sky_texture(ct_host_scene* s, int32_t texture);
/* The generator's camera for a config scene (Camera ctor, Camera.cpp:9-17). */
int ct_host_scene_camera(const ct_host_scene* s, int32_t width, int32_t height, ct_camera* cam);
/* Make a camera from position/look-at/fov (degrees). */
void ct_camera_look_at(const float pos[3], const float target[3], float fov_deg, ct_camera* cam);
/* The configuration's frame options (depth, spp, accel, BSDF mode, epsilons)
 * and image size for a generated scene. */
int ct_host_scene_frame_opts(const ct_host_scene* s, ct_frame_opts* opts, int32_t* width, int32_t* height);
/* [triangles, binary nodes, quad nodes, max binary depth, leaves, lights,
 *  binary build us, quad build us]. */
int ct_host_scene_stats(const ct_host_scene* s, uint64_t out[8]);

/* ---------------------------------------------------------- measurement */
/* ct_trace_frame with per-stage work counters: work[0..3] = closest-hit node
 * tests, closest-hit triangle tests, shadow node tests, shadow triangle
 * tests.  Used to derive algorithmic bytes per launch; slower than
 * ct_trace_frame (atomics), never used in timed runs. */
int ct_trace_frame_counted(ct_ctx* ctx, const ct_camera* cam, const ct_frame_opts* opts,
                           ct_frame_stats* stats, uint64_t work[4]);
/* ct_trace_frame with the frame's wavefront chunks run one at a time and
 * every kernel timed (stats required): each traversal launch has the device
 * to itself, so stats->extend_ms / connect_ms are the kernels' own times
 * (in ct_trace_frame two chunks and their launches overlap).  Same image;
 * slower, for measurement only. */
int ct_trace_frame_serial(ct_ctx* ctx, const ct_camera* cam, const ct_frame_opts* opts,
                          ct_frame_stats* stats);

/* The wide compressed tree of CT_TRAV_FAST, which ct_scene_upload builds
 * from the view's binary tree (or quad tree): up to 8 children per node,
 * 32 words per node (no counterpart in the reference; exposed so the
 * quantisation can be checked on the host).  Words: p.xyz (float), biased
 * exponents of the per-axis steps | valid-child mask << 24, per axis the low
 * and high 8-bit child offsets (8 bytes each), eight child entries (node
 * index, or 0x80000000 | first primitive-list position), 8 zero words.
 * Call with out_words == NULL for the node count; stack_bound (may be NULL)
 * receives the deepest traversal stack the tree can need. */
int ct_build_wide_nodes(const ct_scene_view* v, uint32_t* out_words, uint32_t capacity_nodes, uint32_t* n_nodes,
                        uint32_t* stack_bound);

#ifdef __cplusplus
}
#endif
#endif /* CUDATRACER_H */
