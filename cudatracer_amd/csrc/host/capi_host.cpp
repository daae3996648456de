// C ABI over the host scene classes (TriangleList, BVHTree, MBVHTree, Camera)
// and the configuration scene generators.
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <string>

#include "../../../include/cudatracer.h"
#include "bvh.h"
#include "cnode.h"
#include "scene.h"
#include "scenegen.h"

// The ABI structs mirror the reference's records byte for byte, so a caller
// can pass its arrays by reinterpret_cast: Material (Material.cuh:29-41,
// 13 four-byte fields), BVHNode (BVHNode.cuh:29-40), MBVHNode
// (MBVHNode.cuh:28-63) and the camera state (Camera.h:81-92).
static_assert(sizeof(ct_material) == 52, "ct_material must be the 52-byte core::Material");
static_assert(sizeof(ct_bvh_node) == 32, "ct_bvh_node must be the 32-byte BVHNode");
static_assert(sizeof(ct_mbvh_node) == 128, "ct_mbvh_node must be the 128-byte MBVHNode");
static_assert(sizeof(ct_camera) == 40, "ct_camera: position, forward, up, plane distance");
static_assert(sizeof(ct_frame_opts) == 52, "ct_frame_opts: 13 four-byte fields");
static_assert(sizeof(ct_frame_stats) == 88, "ct_frame_stats: 4 counts, 6 times, 2 launch counts (ABI 5)");

struct ct_host_scene {
    ct::TriangleList list;
    std::unique_ptr<ct::BVHTree> bvh;
    std::unique_ptr<ct::MBVHTree> mbvh;
    ct::SceneConfig cfg;
    bool generated = false;
    std::vector<ct_material> mats;
    std::vector<ct_texture> textures;
    std::string error;
};

extern "C" {

int ct_build_wide_nodes(const ct_scene_view* v, uint32_t* out_words, uint32_t capacity_nodes, uint32_t* n_nodes,
                        uint32_t* stack_bound) {
    if (!v || !n_nodes) return CT_ERR_INVALID_ARG;
    std::vector<uint32_t> w;
    uint32_t bound = 0;
    if (ct::build_wide_nodes(*v, w, bound)) return CT_ERR_INVALID_SCENE;
    const uint32_t nn = (uint32_t)(w.size() / ct::kWideNodeWords);
    *n_nodes = nn;
    if (stack_bound) *stack_bound = bound;
    if (!out_words) return CT_OK;
    if (capacity_nodes < nn) return CT_ERR_INVALID_ARG;
    std::memcpy(out_words, w.data(), w.size() * 4);
    return CT_OK;
}

int ct_host_scene_create(ct_host_scene** out) {
    if (!out) return CT_ERR_INVALID_ARG;
    *out = new (std::nothrow) ct_host_scene();
    if (!*out) return CT_ERR_OUT_OF_MEMORY;
    ct_frame_opts_default(&(*out)->cfg.opts);
    return CT_OK;
}

int ct_host_scene_generate(int32_t config, uint64_t seed, ct_host_scene** out) {
    if (!out) return CT_ERR_INVALID_ARG;
    *out = nullptr;
    ct_host_scene* s = new (std::nothrow) ct_host_scene();
    if (!s) return CT_ERR_OUT_OF_MEMORY;
    if (!ct::generate_scene(config, seed, s->list, s->cfg)) {
        delete s;
        return CT_ERR_INVALID_ARG;
    }
    s->generated = true;
    *out = s;
    return CT_OK;
}

void ct_host_scene_destroy(ct_host_scene* s) { delete s; }

int ct_host_scene_add_material(ct_host_scene* s, const ct_material* m, int32_t keep_roughness) {
    if (!s || !m) return -CT_ERR_INVALID_ARG;
    s->bvh.reset();
    s->mbvh.reset();
    return (int)s->list.addMaterial(ct::Material::from_abi(*m), keep_roughness != 0);
}

int ct_host_scene_load_model(ct_host_scene* s, const char* path, float scale, const float mat[16], int32_t material,
                             int32_t normalize) {
    if (!s || !path || !std::isfinite(scale)) return CT_ERR_INVALID_ARG;
    if (mat)
        for (int i = 0; i < 16; ++i)
            if (!std::isfinite(mat[i])) return CT_ERR_INVALID_ARG;
    if (material >= (int32_t)s->list.m_Materials.size()) return CT_ERR_INVALID_ARG;
    std::string why;
    if (!s->list.loadModel(path, scale, mat, material, normalize != 0, &why)) {
        s->error = why;
        return why.rfind("cannot open", 0) == 0 ? CT_ERR_INVALID_ARG : CT_ERR_INVALID_SCENE;
    }
    s->bvh.reset();
    s->mbvh.reset();
    return CT_OK;
}

const char* ct_host_scene_last_error(const ct_host_scene* s) { return s ? s->error.c_str() : ""; }

int ct_host_scene_add_triangle(ct_host_scene* s, const float p[9], const float n[9], uint32_t material,
                               const float uv[6]) {
    if (!s || !p || !n) return CT_ERR_INVALID_ARG;
    if (material >= s->list.m_Materials.size()) return CT_ERR_INVALID_ARG;
    for (int i = 0; i < 9; ++i)
        if (!std::isfinite(p[i]) || !std::isfinite(n[i])) return CT_ERR_INVALID_ARG;
    s->bvh.reset();
    s->mbvh.reset();
    const ct::vec2 t0 = uv ? ct::vec2(uv[0], uv[1]) : ct::vec2(), t1 = uv ? ct::vec2(uv[2], uv[3]) : ct::vec2(),
                   t2 = uv ? ct::vec2(uv[4], uv[5]) : ct::vec2();
    s->list.addTriangle(ct::vec3(p[0], p[1], p[2]), ct::vec3(p[3], p[4], p[5]), ct::vec3(p[6], p[7], p[8]),
                        ct::vec3(n[0], n[1], n[2]), ct::vec3(n[3], n[4], n[5]), ct::vec3(n[6], n[7], n[8]), material,
                        t0, t1, t2);
    return CT_OK;
}

int ct_host_scene_build(ct_host_scene* s) {
    if (!s) return CT_ERR_INVALID_ARG;
    if (s->list.getPrimCount() == 0) return CT_ERR_INVALID_SCENE;
    s->bvh.reset(new ct::BVHTree(&s->list));
    s->bvh->constructBVH();
    s->mbvh.reset(new ct::MBVHTree(s->bvh.get()));
    s->mbvh->constructBVH();
    return CT_OK;
}

int ct_host_scene_view(const ct_host_scene* cs, ct_scene_view* out) {
    if (!cs || !out) return CT_ERR_INVALID_ARG;
    ct_host_scene* s = const_cast<ct_host_scene*>(cs);
    s->mats.clear();
    for (const auto& m : s->list.m_Materials) s->mats.push_back(m.to_abi());
    std::memset(out, 0, sizeof(*out));
    const ct::TriangleList& L = s->list;
    out->n_vertices = (uint32_t)L.m_Vertices.size();
    out->vertices = reinterpret_cast<const float*>(L.m_Vertices.data());
    out->normals = reinterpret_cast<const float*>(L.m_Normals.data());
    out->texcoords = reinterpret_cast<const float*>(L.m_TexCoords.data());
    out->n_triangles = (uint32_t)L.m_Indices.size();
    out->indices = reinterpret_cast<const uint32_t*>(L.m_Indices.data());
    out->center_normals = reinterpret_cast<const float*>(L.m_CenterNormals.data());
    out->material_idx = L.m_MaterialIdxs.data();
    out->n_lights = (uint32_t)L.m_LightIndices.size();
    out->light_indices = L.m_LightIndices.data();
    out->n_materials = (uint32_t)s->mats.size();
    out->materials = s->mats.data();
    // trees only after ct_host_scene_build (a view without them suits ct_scene_upload_build)
    if (s->bvh && s->mbvh) {
        out->prim_indices = s->bvh->m_PrimitiveIndices.data();
        out->n_bvh_nodes = (uint32_t)s->bvh->m_BVHPool.size();
        out->bvh_nodes = reinterpret_cast<const ct_bvh_node*>(s->bvh->m_BVHPool.data());
        out->n_mbvh_nodes = (uint32_t)s->mbvh->m_Tree.size();
        out->mbvh_nodes = reinterpret_cast<const ct_mbvh_node*>(s->mbvh->m_Tree.data());
    }
    out->sky_enabled = s->cfg.sky_enabled;
    out->sky[0] = s->cfg.sky.x;
    out->sky[1] = s->cfg.sky.y;
    out->sky[2] = s->cfg.sky.z;
    s->textures.clear();
    for (const auto& t : L.m_Textures) s->textures.push_back(ct_texture{t.width, t.height, t.rgba.data()});
    out->n_textures = (uint32_t)s->textures.size();
    out->textures = s->textures.data();
    out->sky_texture = s->cfg.sky_texture;
    return CT_OK;
}

int ct_host_scene_add_texture(ct_host_scene* s, uint32_t width, uint32_t height, const float* rgba) {
    if (!s || !rgba || width == 0 || height == 0) return -CT_ERR_INVALID_ARG;
    const int idx = s->list.addTexture(width, height, rgba);
    return idx < 0 ? -CT_ERR_INVALID_ARG : idx;
}

int ct_host_scene_load_texture(ct_host_scene* s, const char* path) {
    if (!s || !path) return -CT_ERR_INVALID_ARG;
    std::string why;
    const int idx = s->list.loadTexture(path, &why);
    if (idx < 0) {
        s->error = why;
        return why.rfind("cannot open", 0) == 0 ? -CT_ERR_INVALID_ARG : -CT_ERR_INVALID_SCENE;
    }
    return idx;
}

int ct_host_scene_set_sky_texture(ct_host_scene* s, int32_t texture) {
    if (!s || texture < -1 || texture >= (int32_t)s->list.m_Textures.size()) return CT_ERR_INVALID_ARG;
    s->cfg.sky_texture = texture;
    if (texture >= 0) s->cfg.sky_enabled = 1;  // loadSkybox enables the sky box (Application.cpp:85)
    return CT_OK;
}

int ct_host_scene_set_sky(ct_host_scene* s, int32_t enabled, const float rgb[3]) {
    if (!s) return CT_ERR_INVALID_ARG;
    s->cfg.sky_enabled = enabled ? 1 : 0;
    if (rgb) s->cfg.sky = ct::vec3(rgb[0], rgb[1], rgb[2]);
    return CT_OK;
}

void ct_camera_look_at(const float pos[3], const float target[3], float fov_deg, ct_camera* cam) {
    if (!pos || !target || !cam) return;
    ct::Camera c(1, 1, fov_deg, ct::vec3(pos[0], pos[1], pos[2]));
    c.lookAt(ct::vec3(target[0], target[1], target[2]));
    *cam = c.to_abi();
}

int ct_host_scene_camera(const ct_host_scene* s, int32_t width, int32_t height, ct_camera* cam) {
    if (!s || !cam || width <= 0 || height <= 0) return CT_ERR_INVALID_ARG;
    ct::Camera c(width, height, s->cfg.cam_fov, s->cfg.cam_pos);
    c.lookAt(s->cfg.cam_target);
    *cam = c.to_abi();
    return CT_OK;
}

int ct_host_scene_frame_opts(const ct_host_scene* s, ct_frame_opts* opts, int32_t* width, int32_t* height) {
    if (!s || !opts) return CT_ERR_INVALID_ARG;
    *opts = s->cfg.opts;
    if (width) *width = s->cfg.width;
    if (height) *height = s->cfg.height;
    return CT_OK;
}

int ct_host_scene_stats(const ct_host_scene* s, uint64_t out[8]) {
    if (!s || !out) return CT_ERR_INVALID_ARG;
    std::memset(out, 0, 8 * sizeof(uint64_t));
    out[0] = s->list.getPrimCount();
    out[1] = s->bvh ? s->bvh->m_BVHPool.size() : 0;
    out[2] = s->mbvh ? s->mbvh->m_Tree.size() : 0;
    out[3] = s->bvh ? (uint64_t)s->bvh->m_MaxDepth : 0;
    uint64_t leaves = 0;
    if (s->bvh)
        for (const auto& n : s->bvh->m_BVHPool) leaves += n.count > 0 ? 1 : 0;
    out[4] = leaves;
    out[5] = s->list.m_LightIndices.size();
    out[6] = s->bvh ? (uint64_t)(s->bvh->m_BuildMs * 1000.0) : 0;
    out[7] = s->mbvh ? (uint64_t)(s->mbvh->m_BuildMs * 1000.0) : 0;
    return CT_OK;
}

}  // extern "C"
