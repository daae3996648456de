// C ABI entry points: scene upload (with the node/triangle export layouts),
// resize/reset/material update, per-frame trace, accumulator access,
// ray-level tracing, batched multi-view rendering and the autotuner.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <sstream>

#include "../host/cnode.h"
#include "runtime.h"

using ctd::Segment;
using ctd::ViewDesc;
using namespace ctr;

namespace {

bool finite3(const float* p) { return std::isfinite(p[0]) && std::isfinite(p[1]) && std::isfinite(p[2]); }

// triangle::getArea (Triangle.cuh:89-96) in float, glm order.
float heron_area(const float* p0, const float* p1, const float* p2) {
    auto len = [](const float* a, const float* b) {
        const float x = a[0] - b[0], y = a[1] - b[1], z = a[2] - b[2];
        return std::sqrt(x * x + y * y + z * z);
    };
    const float a = len(p0, p1), b = len(p1, p2), c = len(p2, p0);
    const float s = (a + b + c) / 2.0f;
    return std::sqrt(s * (s - a) * (s - b) * (s - c));
}

float4 f4(const float* p, float w) { return make_float4(p[0], p[1], p[2], w); }
float ubits(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// Collect leaf ranges of a tree while checking that it is a proper tree
// (every interior child in range and reached once).
int collect_bvh2(const ct_scene_view* v, std::vector<std::pair<uint32_t, uint32_t>>& leaves, std::string& why) {
    const uint32_t nn = v->n_bvh_nodes, N = v->n_triangles;
    std::vector<uint8_t> seen(nn, 0);
    std::vector<uint32_t> st{0};
    seen[0] = 1;
    while (!st.empty()) {
        const uint32_t i = st.back();
        st.pop_back();
        const ct_bvh_node& n = v->bvh_nodes[i];
        if (n.count >= 0) {
            if ((uint64_t)(uint32_t)n.left_first + (uint64_t)n.count > N || n.left_first < 0) {
                why = "binary leaf range outside the primitive list";
                return 1;
            }
            if (n.count > 0) leaves.emplace_back((uint32_t)n.left_first, (uint32_t)n.count);
            continue;
        }
        if (n.count != -1 || n.left_first < 1 || (uint64_t)n.left_first + 1 >= nn) {
            why = "binary interior node with bad children";
            return 1;
        }
        for (uint32_t k = 0; k < 2; ++k) {
            const uint32_t ch = (uint32_t)n.left_first + k;
            if (seen[ch]) { why = "binary tree has a shared or cyclic child"; return 1; }
            seen[ch] = 1;
            st.push_back(ch);
        }
    }
    return 0;
}

int collect_bvh4(const ct_scene_view* v, std::vector<std::pair<uint32_t, uint32_t>>& leaves, std::string& why) {
    const uint32_t nn = v->n_mbvh_nodes, N = v->n_triangles;
    std::vector<uint8_t> seen(nn, 0);
    std::vector<uint32_t> st{0};
    seen[0] = 1;
    while (!st.empty()) {
        const uint32_t i = st.back();
        st.pop_back();
        const ct_mbvh_node& n = v->mbvh_nodes[i];
        for (int s = 0; s < 4; ++s) {
            const int cnt = n.counts[s], ch = n.childs[s];
            if (cnt == 0) continue;
            if (cnt == -1) {
                if (ch < 1 || (uint32_t)ch >= nn) { why = "quad interior child out of range"; return 1; }
                if (seen[ch]) { why = "quad tree has a shared or cyclic child"; return 1; }
                seen[ch] = 1;
                st.push_back((uint32_t)ch);
            } else if (cnt > 0) {
                if (ch < 0 || (uint64_t)(uint32_t)ch + (uint64_t)cnt > N) {
                    why = "quad leaf range outside the primitive list";
                    return 1;
                }
                leaves.emplace_back((uint32_t)ch, (uint32_t)cnt);
            } else {
                why = "quad slot with a negative count other than -1";
                return 1;
            }
        }
    }
    return 0;
}

// Mark the last triangle of each leaf; leaves must not overlap.
int leaf_flags(const std::vector<std::pair<uint32_t, uint32_t>>& leaves, uint32_t N, std::vector<uint8_t>& last,
               std::string& why) {
    std::vector<uint8_t> used(N, 0);
    last.assign(N, 0);
    for (auto& l : leaves) {
        for (uint32_t k = l.first; k < l.first + l.second; ++k) {
            if (used[k]) { why = "leaves overlap"; return 1; }
            used[k] = 1;
        }
        last[l.first + l.second - 1] = 1;
    }
    return 0;
}

void free_scene(ct_ctx* c) {
    for (void* p : c->scene_bufs) dfree(c, p);
    c->scene_bufs.clear();
    c->sc = ctd::DevScene{};
    c->has_scene = c->has_bvh2 = c->has_bvh4 = false;
}

template <class T>
int upload_vec(ct_ctx* c, const std::vector<T>& h, const float4** dst, const char* what) {
    void* p = nullptr;
    const int r = dalloc(c, &p, h.size() * sizeof(T), what);
    if (r != CT_OK) return r;
    c->scene_bufs.push_back(p);
    if (!h.empty()) HIPCHK(hipMemcpyAsync(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, c->ln[0].stream), what);
    *dst = reinterpret_cast<const float4*>(p);
    return CT_OK;
}

// Row 1 .w: texture indices + 1 (0 = none), diffuse in bits 0-15, normal in 16-31.
void mat_rows(const ct_material& m, float4* rows) {
    const uint32_t tex = ((uint32_t)(m.diffuse_tex + 1) & 0xFFFFu) | (((uint32_t)(m.normal_tex + 1) & 0xFFFFu) << 16);
    rows[0] = make_float4(m.albedo[0], m.albedo[1], m.albedo[2], m.roughness);
    rows[1] = make_float4(ubits((uint32_t)m.type), m.refract_idx, std::max(m.roughness, 1e-3f), ubits(tex));
}

// Texture indices of a material: -1 or an existing texture.
bool mat_textures_ok(const ct_material& m, uint32_t n_tex) {
    return m.diffuse_tex >= -1 && m.normal_tex >= -1 && (m.diffuse_tex < 0 || (uint32_t)m.diffuse_tex < n_tex) &&
           (m.normal_tex < 0 || (uint32_t)m.normal_tex < n_tex);
}

void light_rows(const ct_scene_view* v, uint32_t tri, const ct_material* mats, float4* rows) {
    const float* p0 = v->vertices + 3 * (size_t)v->indices[3 * (size_t)tri];
    const float* p1 = v->vertices + 3 * (size_t)v->indices[3 * (size_t)tri + 1];
    const float* p2 = v->vertices + 3 * (size_t)v->indices[3 * (size_t)tri + 2];
    rows[0] = f4(p0, 0.0f);
    rows[1] = f4(p1, 0.0f);
    rows[2] = f4(p2, 0.0f);
    rows[3] = f4(v->center_normals + 3 * (size_t)tri, heron_area(p0, p1, p2));
    rows[4] = f4(mats[v->material_idx[tri]].albedo, 0.0f);
}

// Checks shared by both upload paths (everything but the trees).
int check_geometry(ct_ctx* c, const ct_scene_view* v) {
    const uint32_t N = v->n_triangles, V = v->n_vertices;
    if (N == 0) return fail(c, CT_ERR_INVALID_SCENE, "scene has no triangles");
    if (N >= (1u << 30)) return fail(c, CT_ERR_INVALID_SCENE, "too many triangles");
    if (!v->vertices || !v->normals || !v->indices || !v->center_normals || !v->material_idx || !v->materials ||
        v->n_materials == 0 || (v->n_lights && !v->light_indices))
        return fail(c, CT_ERR_INVALID_ARG, "scene view is missing an array");
    for (uint64_t i = 0; i < 3ull * N; ++i)
        if (v->indices[i] >= V) return fail(c, CT_ERR_INVALID_SCENE, "vertex index out of range");
    for (uint32_t i = 0; i < V; ++i)
        if (!finite3(v->vertices + 3 * (size_t)i)) return fail(c, CT_ERR_INVALID_SCENE, "non-finite vertex");
    for (uint32_t i = 0; i < N; ++i)
        if (v->material_idx[i] >= v->n_materials) return fail(c, CT_ERR_INVALID_SCENE, "material index out of range");
    for (uint32_t i = 0; i < v->n_lights; ++i)
        if (v->light_indices[i] >= N) return fail(c, CT_ERR_INVALID_SCENE, "light index out of range");
    if (v->n_textures > 0xFFFEu) return fail(c, CT_ERR_INVALID_SCENE, "too many textures");
    if (v->n_textures && !v->textures) return fail(c, CT_ERR_INVALID_ARG, "scene view is missing an array");
    uint64_t texels = 0;
    for (uint32_t t = 0; t < v->n_textures; ++t) {
        const ct_texture& x = v->textures[t];
        if (x.width == 0 || x.height == 0 || !x.rgba) return fail(c, CT_ERR_INVALID_SCENE, "empty texture");
        texels += (uint64_t)x.width * x.height;
    }
    if (texels >= (1ull << 32)) return fail(c, CT_ERR_INVALID_SCENE, "textures too large");
    for (uint32_t m = 0; m < v->n_materials; ++m)
        if (!mat_textures_ok(v->materials[m], v->n_textures))
            return fail(c, CT_ERR_INVALID_SCENE, "material texture index out of range");
    if (v->n_textures && (v->sky_texture < -1 || (v->sky_texture >= 0 && (uint32_t)v->sky_texture >= v->n_textures)))
        return fail(c, CT_ERR_INVALID_SCENE, "sky texture index out of range");
    return CT_OK;
}

// Host staging of the tree-independent device arrays.
struct Shading {
    std::vector<float4> prim, mats, lights;  // prim: ctd::kPrimRows rows per primitive
    std::vector<float4> texels;
    std::vector<uint32_t> texels8;
    std::vector<uint4> tex_info;
};

// An 8-bit texture: alpha 1 and every colour value c / 255.0f for an integer
// c in [0, 255] (what Surface holds for 8-bit images); stored as bytes.
bool is_8bit(const ct_texture& x) {
    const size_t n = (size_t)x.width * x.height;
    for (size_t i = 0; i < n; ++i) {
        if (x.rgba[4 * i + 3] != 1.0f) return false;
        for (int k = 0; k < 3; ++k) {
            const float v = x.rgba[4 * i + k];
            if (!(v >= 0.0f && v <= 1.0f)) return false;
            const float c = std::nearbyint(v * 255.0f);
            if ((float)(uint32_t)c / 255.0f != v) return false;
        }
    }
    return true;
}

void stage_shading(const ct_scene_view* v, Shading& sh) {
    const uint32_t N = v->n_triangles;
    constexpr size_t R = ctd::kPrimRows;
    sh.prim.assign((size_t)N * R, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
    for (uint32_t prim = 0; prim < N; ++prim) {
        float4* row = &sh.prim[R * prim];
        // texture coordinates u0, v0, u1, v1, u2, v2 (TriangleList.cpp:148-178)
        float t[6] = {0, 0, 0, 0, 0, 0};
        if (v->n_textures && v->texcoords)
            for (int j = 0; j < 3; ++j) {
                const uint32_t vi = v->indices[3 * (size_t)prim + j];
                t[2 * j] = v->texcoords[2 * (size_t)vi];
                t[2 * j + 1] = v->texcoords[2 * (size_t)vi + 1];
            }
        for (int j = 0; j < 3; ++j) {
            const uint32_t vi = v->indices[3 * (size_t)prim + j];
            row[j] = f4(v->vertices + 3 * (size_t)vi, j == 0 ? ubits(v->material_idx[prim]) : t[j - 1]);
            row[3 + j] = f4(v->normals + 3 * (size_t)vi, t[2 + j]);
        }
        row[6] = f4(v->center_normals + 3 * (size_t)prim, t[5]);
    }
    sh.mats.resize((size_t)v->n_materials * 2);
    for (uint32_t m = 0; m < v->n_materials; ++m) mat_rows(v->materials[m], &sh.mats[2 * (size_t)m]);
    sh.lights.resize((size_t)std::max<uint32_t>(v->n_lights, 1u) * 5);
    for (uint32_t l = 0; l < v->n_lights; ++l) light_rows(v, v->light_indices[l], v->materials, &sh.lights[5 * (size_t)l]);
    // TriangleList::createTextureBuffer (TriangleList.cpp:148-178)
    for (uint32_t t = 0; t < v->n_textures; ++t) {
        const ct_texture& x = v->textures[t];
        const size_t n = (size_t)x.width * x.height;
        if (is_8bit(x)) {
            sh.tex_info.push_back(make_uint4((uint32_t)sh.texels8.size(), x.width, x.height, 1u));
            for (size_t i = 0; i < n; ++i) {
                uint32_t p = 0;
                for (int k = 0; k < 3; ++k) p |= (uint32_t)std::nearbyint(x.rgba[4 * i + k] * 255.0f) << (8 * k);
                sh.texels8.push_back(p);
            }
        } else {
            sh.tex_info.push_back(make_uint4((uint32_t)sh.texels.size(), x.width, x.height, 0u));
            for (size_t i = 0; i < n; ++i)
                sh.texels.push_back(make_float4(x.rgba[4 * i], x.rgba[4 * i + 1], x.rgba[4 * i + 2], x.rgba[4 * i + 3]));
        }
    }
}

int upload_shading(ct_ctx* c, const Shading& sh, ctd::DevScene& sc) {
    int r = upload_vec(c, sh.prim, &sc.sh_prim, "shading records");
    if (r == CT_OK) r = upload_vec(c, sh.mats, &sc.mats, "materials");
    if (r == CT_OK) r = upload_vec(c, sh.lights, &sc.lights, "lights");
    if (r == CT_OK && !sh.tex_info.empty()) {
        r = upload_vec(c, sh.texels, &sc.texels, "textures");
        const float4* t8 = nullptr;
        if (r == CT_OK) r = upload_vec(c, sh.texels8, &t8, "8-bit textures");
        sc.texels8 = reinterpret_cast<const uint32_t*>(t8);
        const float4* ti = nullptr;
        if (r == CT_OK) r = upload_vec(c, sh.tex_info, &ti, "texture table");
        sc.tex_info = reinterpret_cast<const uint4*>(ti);
    }
    return r;
}

// The fast node test's copy of the uploaded quad tree (sc.nodes4f), its
// compact layout (sc.cnodes, 64 bytes per node) and the tree's flags
// (sc.plane_flags): the fast test needs planes that are not NaN, empty slots
// whose child is 0xFFFFFFFF, and origins without +0 components only where
// some plane is -0; the compact traversal needs every node encoded.
int prepare_quad(ct_ctx* c, ctd::DevScene& sc) {
    sc.plane_flags = 0;
    sc.nodes4f = nullptr;
    sc.cnodes = nullptr;
    if (!sc.nodes4 || sc.n_nodes4 == 0) return CT_OK;
    void* fast = nullptr;
    int r = dalloc(c, &fast, (size_t)sc.n_nodes4 * sizeof(ct_mbvh_node), "fast quad nodes");
    if (r != CT_OK) return r;
    c->scene_bufs.push_back(fast);
    void* compact = nullptr;
    r = dalloc(c, &compact, (size_t)sc.n_nodes4 * 64, "compact quad nodes");
    if (r != CT_OK) return r;
    c->scene_bufs.push_back(compact);
    uint32_t* d = nullptr;
    HIPCHK(hipMalloc(&d, sizeof(uint32_t)), "quad tree flags");
    const hipStream_t s = c->ln[0].stream;
    uint32_t h = 0;
    hipError_t e = hipMemsetAsync(d, 0, sizeof(uint32_t), s);
    if (e == hipSuccess) e = ctk::launch_prepare_quad(sc.nodes4, sc.n_nodes4, reinterpret_cast<float4*>(fast),
                                                           static_cast<uint4*>(compact), d, s);
    if (e == hipSuccess) e = hipMemcpyAsync(&h, d, sizeof(uint32_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(d);
    if (e != hipSuccess) return ctr::hip_fail(c, e, "quad tree flags");
    sc.nodes4f = reinterpret_cast<const float4*>(fast);
    sc.cnodes = static_cast<const uint4*>(compact);
    sc.plane_flags = h;
    return CT_OK;
}

// Make an uploaded scene current (tree fields of sc already set).
void install_scene(ct_ctx* c, const ct_scene_view* v, ctd::DevScene sc, const Shading& sh, uint32_t wide_bound) {
    const uint32_t N = v->n_triangles;
    sc.n_tris = N;
    sc.n_mats = v->n_materials;
    sc.n_lights = v->n_lights;
    sc.sky_enabled = v->sky_enabled ? 1 : 0;
    sc.n_tex = v->n_textures;
    sc.sky_tex = v->n_textures ? v->sky_texture : -1;
    sc.sky_r = v->sky[0];
    sc.sky_g = v->sky[1];
    sc.sky_b = v->sky[2];
    c->sc = sc;
    c->has_bvh4 = sc.n_nodes4 > 0;
    c->has_bvh2 = sc.n_nodes2 > 0;
    c->wide_stack_bound = wide_bound;
    c->has_scene = true;
    c->h_mats.assign(v->materials, v->materials + v->n_materials);
    c->h_mat_idx.assign(v->material_idx, v->material_idx + N);
    c->h_light_idx.assign(v->light_indices, v->light_indices + v->n_lights);
    c->h_light_rows.assign(reinterpret_cast<const float*>(sh.lights.data()),
                           reinterpret_cast<const float*>(sh.lights.data()) + sh.lights.size() * 4);
    drop_graphs(c);
}

}  // namespace

extern "C" {

int ct_scene_upload(ct_ctx* c, const ct_scene_view* v) {
    if (!c || !v) return CT_ERR_INVALID_ARG;
    (void)hipSetDevice(c->device);
    int r = check_geometry(c, v);
    if (r != CT_OK) return r;
    const uint32_t N = v->n_triangles;
    if (!v->prim_indices) return fail(c, CT_ERR_INVALID_ARG, "scene view is missing an array");
    if (v->n_bvh_nodes == 0 && v->n_mbvh_nodes == 0)
        return fail(c, CT_ERR_INVALID_ARG, "scene view has no acceleration tree");
    {
        std::vector<uint8_t> seen(N, 0);
        for (uint32_t i = 0; i < N; ++i) {
            const uint32_t p = v->prim_indices[i];
            if (p >= N || seen[p]) return fail(c, CT_ERR_INVALID_SCENE, "prim_indices is not a permutation");
            seen[p] = 1;
        }
    }
    std::string why;
    std::vector<uint8_t> last, last2;
    bool have_flags = false;
    if (v->n_mbvh_nodes) {
        std::vector<std::pair<uint32_t, uint32_t>> lv;
        if (collect_bvh4(v, lv, why) || leaf_flags(lv, N, last, why)) return fail(c, CT_ERR_INVALID_SCENE, why);
        have_flags = true;
    }
    if (v->n_bvh_nodes) {
        std::vector<std::pair<uint32_t, uint32_t>> lv;
        if (collect_bvh2(v, lv, why) || leaf_flags(lv, N, last2, why)) return fail(c, CT_ERR_INVALID_SCENE, why);
        if (have_flags && last2 != last)
            return fail(c, CT_ERR_INVALID_SCENE, "binary and quad trees have different leaves");
        if (!have_flags) last.swap(last2);
    }

    // host staging of the device layouts
    std::vector<float4> tris((size_t)N * 3);
    for (uint32_t k = 0; k < N; ++k) {
        const uint32_t prim = v->prim_indices[k];
        const float* p0 = v->vertices + 3 * (size_t)v->indices[3 * (size_t)prim];
        const float* p1 = v->vertices + 3 * (size_t)v->indices[3 * (size_t)prim + 1];
        const float* p2 = v->vertices + 3 * (size_t)v->indices[3 * (size_t)prim + 2];
        tris[3 * (size_t)k] = f4(p0, ubits(prim | (last[k] ? 0x80000000u : 0u)));
        tris[3 * (size_t)k + 1] = make_float4(p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2], 0.0f);
        tris[3 * (size_t)k + 2] = make_float4(p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2], 0.0f);
    }
    Shading sh;
    stage_shading(v, sh);
    std::vector<ct_mbvh_node> n4(v->mbvh_nodes, v->mbvh_nodes + v->n_mbvh_nodes);
    std::vector<ct_bvh_node> n2(v->bvh_nodes, v->bvh_nodes + v->n_bvh_nodes);
    std::vector<uint32_t> fn;
    uint32_t wide_bound = 0;
    if (ct::build_wide_nodes(*v, fn, wide_bound)) return fail(c, CT_ERR_INVALID_SCENE, "tree bounds are not finite");

    sync_lanes(c);
    free_scene(c);
    ctd::DevScene sc{};
    r = upload_vec(c, tris, &sc.tris, "triangles");
    if (r == CT_OK) r = upload_shading(c, sh, sc);
    if (r == CT_OK && !n4.empty()) r = upload_vec(c, n4, &sc.nodes4, "quad nodes");
    if (r == CT_OK) {
        const float4* fp = nullptr;
        r = upload_vec(c, fn, &fp, "wide nodes");
        sc.fnodes = reinterpret_cast<const uint4*>(fp);
    }
    if (r == CT_OK && !n2.empty()) r = upload_vec(c, n2, &sc.nodes2, "binary nodes");
    if (r != CT_OK) { free_scene(c); return r; }
    HIPCHK(hipStreamSynchronize(c->ln[0].stream), "scene upload");
    sc.n_nodes4 = v->n_mbvh_nodes;
    sc.n_nodes2 = v->n_bvh_nodes;
    sc.n_fnodes = (uint32_t)(fn.size() / ct::kWideNodeWords);
    if ((r = prepare_quad(c, sc)) != CT_OK) { free_scene(c); return r; }
    install_scene(c, v, sc, sh, wide_bound);
    return CT_OK;
}

int ct_scene_upload_build(ct_ctx* c, const ct_scene_view* v, uint32_t flags, ct_build_stats* st) {
    if (!c || !v) return CT_ERR_INVALID_ARG;
    if (flags & ~(uint32_t)CT_BUILD_WIDE) return fail(c, CT_ERR_INVALID_ARG, "unknown build flags");
    (void)hipSetDevice(c->device);
    const auto t0 = std::chrono::steady_clock::now();
    int r = check_geometry(c, v);
    if (r != CT_OK) return r;
    Shading sh;
    stage_shading(v, sh);
    sync_lanes(c);
    free_scene(c);
    ctd::DevScene sc{};
    ct_build_stats bs{};
    DeviceTrees t;
    r = device_build(c, v, &t, &bs);
    for (void* p : {(void*)t.tris, (void*)t.nodes2, (void*)t.nodes4})
        if (p) c->scene_bufs.push_back(p);
    if (r != CT_OK) { free_scene(c); return r; }
    sc.tris = t.tris;
    sc.nodes2 = reinterpret_cast<const float4*>(t.nodes2);
    sc.nodes4 = reinterpret_cast<const float4*>(t.nodes4);
    sc.n_nodes2 = t.n2;
    sc.n_nodes4 = t.n4;
    r = upload_shading(c, sh, sc);
    uint32_t wide_bound = 0;
    if (r == CT_OK && (flags & CT_BUILD_WIDE)) {
        // wide nodes are collapsed on the host from the device-built trees
        std::vector<ct_bvh_node> n2(t.n2);
        std::vector<ct_mbvh_node> n4(t.n4);
        HIPCHK(hipMemcpy(n2.data(), t.nodes2, n2.size() * sizeof(ct_bvh_node), hipMemcpyDeviceToHost), "read trees");
        HIPCHK(hipMemcpy(n4.data(), t.nodes4, n4.size() * sizeof(ct_mbvh_node), hipMemcpyDeviceToHost), "read trees");
        ct_scene_view tv = *v;
        tv.n_bvh_nodes = t.n2;
        tv.bvh_nodes = n2.data();
        tv.n_mbvh_nodes = t.n4;
        tv.mbvh_nodes = n4.data();
        std::vector<uint32_t> fn;
        if (ct::build_wide_nodes(tv, fn, wide_bound)) r = fail(c, CT_ERR_INVALID_SCENE, "tree bounds are not finite");
        const float4* fp = nullptr;
        if (r == CT_OK) r = upload_vec(c, fn, &fp, "wide nodes");
        sc.fnodes = reinterpret_cast<const uint4*>(fp);
        sc.n_fnodes = (uint32_t)(fn.size() / ct::kWideNodeWords);
    }
    if (r != CT_OK) { free_scene(c); return r; }
    HIPCHK(hipStreamSynchronize(c->ln[0].stream), "scene upload");
    if ((r = prepare_quad(c, sc)) != CT_OK) { free_scene(c); return r; }
    install_scene(c, v, sc, sh, wide_bound);
    bs.total_ms = (float)std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (st) *st = bs;
    return CT_OK;
}

int ct_scene_tree_sizes(ct_ctx* c, uint32_t* n_bvh, uint32_t* n_mbvh, uint32_t* n_prims) {
    if (!c) return CT_ERR_INVALID_ARG;
    if (!c->has_scene) return fail(c, CT_ERR_NO_SCENE, "no scene uploaded");
    if (n_bvh) *n_bvh = c->has_bvh2 ? c->sc.n_nodes2 : 0;
    if (n_mbvh) *n_mbvh = c->has_bvh4 ? c->sc.n_nodes4 : 0;
    if (n_prims) *n_prims = c->sc.n_tris;
    return CT_OK;
}

int ct_scene_read_trees(ct_ctx* c, ct_bvh_node* bvh, uint32_t cap_bvh, ct_mbvh_node* mbvh, uint32_t cap_mbvh,
                        uint32_t* prim_indices, uint32_t cap_prims) {
    if (!c) return CT_ERR_INVALID_ARG;
    if (!c->has_scene) return fail(c, CT_ERR_NO_SCENE, "no scene uploaded");
    (void)hipSetDevice(c->device);
    sync_lanes(c);
    const uint32_t n2 = c->has_bvh2 ? c->sc.n_nodes2 : 0, n4 = c->has_bvh4 ? c->sc.n_nodes4 : 0, N = c->sc.n_tris;
    if ((bvh && cap_bvh < n2) || (mbvh && cap_mbvh < n4) || (prim_indices && cap_prims < N))
        return fail(c, CT_ERR_INVALID_ARG, "output capacity too small");
    if (bvh && n2) HIPCHK(hipMemcpy(bvh, c->sc.nodes2, n2 * sizeof(ct_bvh_node), hipMemcpyDeviceToHost), "read trees");
    if (mbvh && n4)
        HIPCHK(hipMemcpy(mbvh, c->sc.nodes4, n4 * sizeof(ct_mbvh_node), hipMemcpyDeviceToHost), "read trees");
    if (prim_indices) {
        // primitive id of slot k: the w word of its first record, without the last-in-leaf bit
        HIPCHK(hipMemcpy2D(prim_indices, 4, reinterpret_cast<const char*>(c->sc.tris) + 12, 3 * sizeof(float4), 4, N,
                           hipMemcpyDeviceToHost),
               "read trees");
        for (uint32_t k = 0; k < N; ++k) prim_indices[k] &= 0x7FFFFFFFu;
    }
    return CT_OK;
}

int ct_material_update(ct_ctx* c, uint32_t index, const ct_material* mat) {
    if (!c || !mat) return CT_ERR_INVALID_ARG;
    if (!c->has_scene) return fail(c, CT_ERR_NO_SCENE, "no scene uploaded");
    if (index >= c->h_mats.size()) return fail(c, CT_ERR_INVALID_ARG, "material index out of range");
    if (!mat_textures_ok(*mat, c->sc.n_tex)) return fail(c, CT_ERR_INVALID_ARG, "material texture index out of range");
    (void)hipSetDevice(c->device);
    c->h_mats[index] = *mat;
    float4 rows[2];
    mat_rows(*mat, rows);
    HIPCHK(hipMemcpyAsync((void*)(c->sc.mats + 2 * (size_t)index), rows, sizeof(rows), hipMemcpyHostToDevice,
                          c->ln[0].stream),
           "material update");
    // lights whose triangle uses this material carry its emission
    const float4 e = make_float4(mat->albedo[0], mat->albedo[1], mat->albedo[2], 0.0f);
    for (size_t l = 0; l < c->h_light_idx.size(); ++l)
        if (c->h_mat_idx[c->h_light_idx[l]] == index)
            HIPCHK(hipMemcpyAsync((void*)(c->sc.lights + 5 * l + 4), &e, sizeof(e), hipMemcpyHostToDevice, c->ln[0].stream),
                   "light emission update");
    HIPCHK(hipStreamSynchronize(c->ln[0].stream), "material update");
    return CT_OK;
}

int ct_resize(ct_ctx* c, int32_t width, int32_t height) {
    if (!c) return CT_ERR_INVALID_ARG;
    if (width <= 0 || height <= 0 || (int64_t)width * height > (1ll << 28))
        return fail(c, CT_ERR_INVALID_ARG, "bad frame size");
    (void)hipSetDevice(c->device);
    sync_lanes(c);
    dfree(c, c->accum);
    c->accum = nullptr;
    drop_graphs(c);
    const int r = dalloc(c, (void**)&c->accum, (size_t)width * height * sizeof(float4), "accumulator");
    if (r != CT_OK) { c->width = c->height = 0; return r; }
    c->width = width;
    c->height = height;
    return ct_reset(c);
}

int ct_reset(ct_ctx* c) {
    if (!c) return CT_ERR_INVALID_ARG;
    if (!c->accum) return fail(c, CT_ERR_INVALID_ARG, "no frame buffer (call ct_resize)");
    (void)hipSetDevice(c->device);
    HIPCHK(hipMemsetAsync(c->accum, 0, (size_t)c->width * c->height * sizeof(float4), c->ln[0].stream), "reset");
    HIPCHK(hipStreamSynchronize(c->ln[0].stream), "reset");
    c->samples = 0;
    return CT_OK;
}

static void fill_stats(const ChunkStats& s, ct_frame_stats* out) {
    if (!out) return;
    out->primary_rays = s.primary;
    out->extension_rays = s.extension;
    out->shadow_rays = s.shadow;
    out->samples = s.samples;
    out->gpu_ms = s.total_ms;
    out->extend_ms = s.extend_ms;
    out->chunks = (uint32_t)s.chunks;
    out->extend_launches = (uint32_t)s.extend_launches;
    out->shade_ms = s.shade_ms;
    out->connect_ms = s.connect_ms;
    out->other_ms = s.other_ms;
    out->tail_ms = s.tail_ms;
}

int ct_trace_frame(ct_ctx* c, const ct_camera* cam, const ct_frame_opts* opts, ct_frame_stats* stats) {
    if (!c) return CT_ERR_INVALID_ARG;
    ct_frame_opts o;
    if (opts) o = *opts;
    else ct_frame_opts_default(&o);
    int r = validate_opts(c, o);
    if (r != CT_OK) return r;
    if (!camera_ok(cam)) return fail(c, CT_ERR_INVALID_ARG, "camera is not finite or degenerate");
    if (!c->accum) return fail(c, CT_ERR_INVALID_ARG, "no frame buffer (call ct_resize)");
    (void)hipSetDevice(c->device);
    const uint64_t total = (uint64_t)c->width * c->height * (uint64_t)o.spp;
    if (total >= (1ull << 32)) return fail(c, CT_ERR_INVALID_ARG, "width*height*spp must be < 2^32");
    std::vector<ViewDesc> views{make_view(*cam, c->width, c->height, o.spp, (uint32_t)c->samples, 0)};
    std::vector<Run> runs{{0u, 0u, (uint32_t)total}};
    const uint32_t cap = (uint32_t)std::min<uint64_t>(total, c->chunk_default);
    ChunkStats st;
    r = run_paths(c, views, runs, cap, c->accum, o, stats ? kRunTimed : 0, &st);
    if (r != CT_OK) return r;
    c->samples += o.spp;
    fill_stats(st, stats);
    return CT_OK;
}

// Frame with work counters (node / triangle tests per stage), for the
// benchmark's algorithmic-byte accounting (declared with the measurement
// entry points of the header).
int ct_trace_frame_counted(ct_ctx* c, const ct_camera* cam, const ct_frame_opts* opts, ct_frame_stats* stats,
                           uint64_t work[4]) {
    if (!c || !work) return CT_ERR_INVALID_ARG;
    ct_frame_opts o;
    if (opts) o = *opts;
    else ct_frame_opts_default(&o);
    int r = validate_opts(c, o);
    if (r != CT_OK) return r;
    if (!camera_ok(cam)) return fail(c, CT_ERR_INVALID_ARG, "camera is not finite or degenerate");
    if (!c->accum) return fail(c, CT_ERR_INVALID_ARG, "no frame buffer (call ct_resize)");
    (void)hipSetDevice(c->device);
    const uint64_t total = (uint64_t)c->width * c->height * (uint64_t)o.spp;
    if (total >= (1ull << 32)) return fail(c, CT_ERR_INVALID_ARG, "width*height*spp must be < 2^32");
    std::vector<ViewDesc> views{make_view(*cam, c->width, c->height, o.spp, (uint32_t)c->samples, 0)};
    std::vector<Run> runs{{0u, 0u, (uint32_t)total}};
    const uint32_t cap = (uint32_t)std::min<uint64_t>(total, c->chunk_default);
    ChunkStats st;
    r = run_paths(c, views, runs, cap, c->accum, o, kRunCount, &st);
    if (r != CT_OK) return r;
    c->samples += o.spp;
    fill_stats(st, stats);
    for (int k = 0; k < 4; ++k) work[k] = st.work[k];
    return CT_OK;
}

// Frame with its chunks run one at a time and every kernel timed: each
// traversal launch has the device to itself, so its time is its own (for
// the benchmark's isolated roofline figure; slower than ct_trace_frame).
int ct_trace_frame_serial(ct_ctx* c, const ct_camera* cam, const ct_frame_opts* opts, ct_frame_stats* stats) {
    if (!c || !stats) return CT_ERR_INVALID_ARG;
    ct_frame_opts o;
    if (opts) o = *opts;
    else ct_frame_opts_default(&o);
    int r = validate_opts(c, o);
    if (r != CT_OK) return r;
    if (!camera_ok(cam)) return fail(c, CT_ERR_INVALID_ARG, "camera is not finite or degenerate");
    if (!c->accum) return fail(c, CT_ERR_INVALID_ARG, "no frame buffer (call ct_resize)");
    (void)hipSetDevice(c->device);
    const uint64_t total = (uint64_t)c->width * c->height * (uint64_t)o.spp;
    if (total >= (1ull << 32)) return fail(c, CT_ERR_INVALID_ARG, "width*height*spp must be < 2^32");
    std::vector<ViewDesc> views{make_view(*cam, c->width, c->height, o.spp, (uint32_t)c->samples, 0)};
    std::vector<Run> runs{{0u, 0u, (uint32_t)total}};
    const uint32_t cap = (uint32_t)std::min<uint64_t>(total, c->chunk_default);
    ChunkStats st;
    r = run_paths(c, views, runs, cap, c->accum, o, kRunTimed | kRunSerial, &st);
    if (r != CT_OK) return r;
    c->samples += o.spp;
    fill_stats(st, stats);
    return CT_OK;
}

int ct_read_accum(ct_ctx* c, float* rgba, size_t n_floats, int32_t* samples) {
    if (!c || !rgba) return CT_ERR_INVALID_ARG;
    if (!c->accum) return fail(c, CT_ERR_INVALID_ARG, "no frame buffer");
    const size_t need = (size_t)c->width * c->height * 4;
    if (n_floats < need) return fail(c, CT_ERR_INVALID_ARG, "output buffer too small");
    (void)hipSetDevice(c->device);
    HIPCHK(hipMemcpyAsync(rgba, c->accum, need * 4, hipMemcpyDeviceToHost, c->ln[0].stream), "read accumulator");
    HIPCHK(hipStreamSynchronize(c->ln[0].stream), "read accumulator");
    if (samples) *samples = c->samples;
    return CT_OK;
}

int ct_write_accum(ct_ctx* c, const float* rgba, size_t n_floats, int32_t samples) {
    if (!c || !rgba || samples < 0) return CT_ERR_INVALID_ARG;
    if (!c->accum) return fail(c, CT_ERR_INVALID_ARG, "no frame buffer");
    const size_t need = (size_t)c->width * c->height * 4;
    if (n_floats < need) return fail(c, CT_ERR_INVALID_ARG, "input buffer too small");
    (void)hipSetDevice(c->device);
    HIPCHK(hipMemcpyAsync(c->accum, rgba, need * 4, hipMemcpyHostToDevice, c->ln[0].stream), "write accumulator");
    HIPCHK(hipStreamSynchronize(c->ln[0].stream), "write accumulator");
    c->samples = samples;
    return CT_OK;
}

// Device buffer for the tonemapped image, grown on demand and kept.
static int image_buffer(ct_ctx* c, size_t bytes) {
    if (c->image_bytes >= bytes) return CT_OK;
    dfree(c, c->image);
    c->image = nullptr;
    c->image_bytes = 0;
    const int r = dalloc(c, &c->image, bytes, "image");
    if (r == CT_OK) c->image_bytes = bytes;
    return r;
}

int ct_read_image(ct_ctx* c, float* rgba, size_t n_floats) {
    if (!c || !rgba) return CT_ERR_INVALID_ARG;
    if (!c->accum) return fail(c, CT_ERR_INVALID_ARG, "no frame buffer");
    const size_t npix = (size_t)c->width * c->height;
    if (n_floats < npix * 4) return fail(c, CT_ERR_INVALID_ARG, "output buffer too small");
    (void)hipSetDevice(c->device);
    sync_lanes(c);
    int r = image_buffer(c, npix * sizeof(float4));
    if (r != CT_OK) return r;
    hipError_t e = ctk::launch_tonemap(c->accum, (uint32_t)npix, (float)std::max(c->samples, 1), (float4*)c->image,
                                       c->ln[0].stream);
    if (e == hipSuccess) e = hipMemcpyAsync(rgba, c->image, npix * sizeof(float4), hipMemcpyDeviceToHost, c->ln[0].stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->ln[0].stream);
    if (e != hipSuccess) return hip_fail(c, e, "tonemap");
    return CT_OK;
}

int ct_read_image_rgba8(ct_ctx* c, uint8_t* rgba, size_t n_bytes) {
    if (!c || !rgba) return CT_ERR_INVALID_ARG;
    if (!c->accum) return fail(c, CT_ERR_INVALID_ARG, "no frame buffer");
    const size_t npix = (size_t)c->width * c->height;
    if (n_bytes < npix * 4) return fail(c, CT_ERR_INVALID_ARG, "output buffer too small");
    (void)hipSetDevice(c->device);
    sync_lanes(c);
    int r = image_buffer(c, npix * 4);
    if (r != CT_OK) return r;
    hipError_t e = ctk::launch_tonemap8(c->accum, (uint32_t)npix, (float)std::max(c->samples, 1), (uint32_t*)c->image,
                                        c->ln[0].stream);
    if (e == hipSuccess) e = hipMemcpyAsync(rgba, c->image, npix * 4, hipMemcpyDeviceToHost, c->ln[0].stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->ln[0].stream);
    if (e != hipSuccess) return hip_fail(c, e, "tonemap");
    return CT_OK;
}

static int ray_prologue(ct_ctx* c, const ct_frame_opts* opts, ct_frame_opts& o) {
    if (opts) o = *opts;
    else ct_frame_opts_default(&o);
    if (o.spp < 1) o.spp = 1;
    if (o.max_depth < 1) o.max_depth = 1;
    return validate_opts(c, o);
}

int ct_trace_rays(ct_ctx* c, const float* origins, const float* dirs, uint32_t n, const ct_frame_opts* opts,
                  float* t_out, int32_t* prim_out) {
    if (!c || (n && (!origins || !dirs || !t_out || !prim_out))) return CT_ERR_INVALID_ARG;
    ct_frame_opts o;
    int r = ray_prologue(c, opts, o);
    if (r != CT_OK) return r;
    if (n == 0) return CT_OK;
    (void)hipSetDevice(c->device);
    const uint32_t cap = std::min<uint32_t>(n, 1u << 22);
    r = ensure_paths(c, cap);
    if (r != CT_OK) return r;
    ctd::TraceOpts to{o.accel, o.traversal, o.triangle_epsilon};
    std::vector<float4> ro(cap), rd(cap);
    std::vector<float2> hits(cap);
    for (uint32_t base = 0; base < n; base += cap) {
        const uint32_t m = std::min(cap, n - base);
        for (uint32_t i = 0; i < m; ++i) {
            ro[i] = f4(origins + 3 * (size_t)(base + i), 0.0f);
            rd[i] = f4(dirs + 3 * (size_t)(base + i), 0.0f);
        }
        const ctd::ChunkHdr h{m, 0u, 0u, 0u};
        HIPCHK(hipMemcpyAsync(c->ln[0].q.ray_o[0], ro.data(), m * sizeof(float4), hipMemcpyHostToDevice, c->ln[0].stream), "rays");
        HIPCHK(hipMemcpyAsync(c->ln[0].q.ray_d[0], rd.data(), m * sizeof(float4), hipMemcpyHostToDevice, c->ln[0].stream), "rays");
        HIPCHK(hipMemcpyAsync(c->ln[0].d_hdr, &h, sizeof(h), hipMemcpyHostToDevice, c->ln[0].stream), "header");
        HIPCHK(ctk::launch_setup(c->ln[0].q.counters, c->ln[0].d_hdr, c->ln[0].stream), "setup");
        HIPCHK(ctk::launch_extend(c->shape, c->sc, to, c->ln[0].q, 0, 0, c->ln[0].stream), "extend");
        uint32_t errflag = 0;
        HIPCHK(hipMemcpyAsync(hits.data(), c->ln[0].q.hit, m * sizeof(float2), hipMemcpyDeviceToHost, c->ln[0].stream), "hits");
        HIPCHK(hipMemcpyAsync(&errflag, c->ln[0].q.err, 4, hipMemcpyDeviceToHost, c->ln[0].stream), "flags");
        HIPCHK(hipStreamSynchronize(c->ln[0].stream), "trace rays");
        if (errflag) {
            (void)hipMemsetAsync(c->ln[0].q.err, 0, 4, c->ln[0].stream);
            return fail(c, CT_ERR_STACK_OVERFLOW, "traversal stack bound exceeded");
        }
        for (uint32_t i = 0; i < m; ++i) {
            t_out[base + i] = hits[i].x;
            int32_t p;
            std::memcpy(&p, &hits[i].y, 4);
            prim_out[base + i] = p;
        }
    }
    return CT_OK;
}

int ct_trace_shadow_rays(ct_ctx* c, const float* origins, const float* dirs, const float* tmax, uint32_t n,
                         const ct_frame_opts* opts, uint8_t* occluded_out) {
    if (!c || (n && (!origins || !dirs || !tmax || !occluded_out))) return CT_ERR_INVALID_ARG;
    ct_frame_opts o;
    int r = ray_prologue(c, opts, o);
    if (r != CT_OK) return r;
    if (n == 0) return CT_OK;
    (void)hipSetDevice(c->device);
    const uint32_t cap = std::min<uint32_t>(n, 1u << 22);
    r = ensure_paths(c, cap);
    if (r != CT_OK) return r;
    void* dout = nullptr;
    r = dalloc(c, &dout, cap, "occlusion flags");
    if (r != CT_OK) return r;
    ctd::TraceOpts to{o.accel, o.traversal, o.triangle_epsilon};
    std::vector<float4> so(cap), sd(cap);
    hipError_t e = hipSuccess;
    uint32_t errflag = 0;
    for (uint32_t base = 0; base < n && e == hipSuccess; base += cap) {
        const uint32_t m = std::min(cap, n - base);
        for (uint32_t i = 0; i < m; ++i) {
            so[i] = f4(origins + 3 * (size_t)(base + i), tmax[base + i]);
            sd[i] = f4(dirs + 3 * (size_t)(base + i), 0.0f);
        }
        const ctd::ChunkHdr h{m, 0u, 0u, 0u};
        e = hipMemcpyAsync(c->ln[0].q.sh_o[0], so.data(), m * sizeof(float4), hipMemcpyHostToDevice, c->ln[0].stream);
        if (e == hipSuccess) e = hipMemcpyAsync(c->ln[0].q.sh_d[0], sd.data(), m * sizeof(float4), hipMemcpyHostToDevice, c->ln[0].stream);
        if (e == hipSuccess) e = hipMemcpyAsync(c->ln[0].d_hdr, &h, sizeof(h), hipMemcpyHostToDevice, c->ln[0].stream);
        if (e == hipSuccess) e = ctk::launch_setup(c->ln[0].q.counters, c->ln[0].d_hdr, c->ln[0].stream);
        if (e == hipSuccess) e = ctk::launch_shadow_flags(c->shape, c->sc, to, c->ln[0].q, m, (uint8_t*)dout, c->ln[0].stream);
        if (e == hipSuccess) e = hipMemcpyAsync(occluded_out + base, dout, m, hipMemcpyDeviceToHost, c->ln[0].stream);
        if (e == hipSuccess) e = hipMemcpyAsync(&errflag, c->ln[0].q.err, 4, hipMemcpyDeviceToHost, c->ln[0].stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->ln[0].stream);
        if (errflag) break;
    }
    dfree(c, dout);
    if (e != hipSuccess) return hip_fail(c, e, "trace shadow rays");
    if (errflag) {
        (void)hipMemsetAsync(c->ln[0].q.err, 0, 4, c->ln[0].stream);
        return fail(c, CT_ERR_STACK_OVERFLOW, "traversal stack bound exceeded");
    }
    return CT_OK;
}

int ct_trace_rays_device(ct_ctx* c, const void* d_rays_o, const void* d_rays_d, uint32_t n,
                         const ct_frame_opts* opts, void* d_hits, float* kernel_ms) {
    if (!c || (n && (!d_rays_o || !d_rays_d || !d_hits))) return CT_ERR_INVALID_ARG;
    ct_frame_opts o;
    int r = ray_prologue(c, opts, o);
    if (r != CT_OK) return r;
    (void)hipSetDevice(c->device);
    r = ensure_paths(c, 64);
    if (r != CT_OK) return r;
    ctd::Queues q = c->ln[0].q;
    q.ray_o[0] = (float4*)d_rays_o;
    q.ray_d[0] = (float4*)d_rays_d;
    q.hit = (float2*)d_hits;
    ctd::TraceOpts to{o.accel, o.traversal, o.triangle_epsilon};
    const ctd::ChunkHdr h{n, 0u, 0u, 0u};
    HIPCHK(hipMemcpyAsync(c->ln[0].d_hdr, &h, sizeof(h), hipMemcpyHostToDevice, c->ln[0].stream), "header");
    HIPCHK(ctk::launch_setup(c->ln[0].q.counters, c->ln[0].d_hdr, c->ln[0].stream), "setup");
    HIPCHK(hipEventRecord(c->ln[0].ev[0], c->ln[0].stream), "event");
    HIPCHK(ctk::launch_extend(c->shape, c->sc, to, q, 0, 0, c->ln[0].stream), "extend");
    HIPCHK(hipEventRecord(c->ln[0].ev[1], c->ln[0].stream), "event");
    uint32_t errflag = 0;
    HIPCHK(hipMemcpyAsync(&errflag, c->ln[0].q.err, 4, hipMemcpyDeviceToHost, c->ln[0].stream), "flags");
    HIPCHK(hipStreamSynchronize(c->ln[0].stream), "trace rays");
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, c->ln[0].ev[0], c->ln[0].ev[1]), "elapsed");
    if (kernel_ms) *kernel_ms = ms;
    if (errflag) {
        // cleared here, so the next call on the context does not report it
        HIPCHK(hipMemsetAsync(c->ln[0].q.err, 0, 4, c->ln[0].stream), "clear error flags");
        HIPCHK(hipStreamSynchronize(c->ln[0].stream), "clear error flags");
        return fail(c, CT_ERR_STACK_OVERFLOW, "traversal stack bound exceeded (rays reported as misses)");
    }
    return CT_OK;
}

int ct_render_batch(ct_ctx* c, const ct_view* views, int32_t n_views, const ct_frame_opts* opts,
                    ct_batch_stats* stats) {
    if (!c || !views || n_views <= 0) return CT_ERR_INVALID_ARG;
    ct_frame_opts o;
    if (opts) o = *opts;
    else ct_frame_opts_default(&o);
    o.spp = 1;
    int r = validate_opts(c, o);
    if (r != CT_OK) return r;
    (void)hipSetDevice(c->device);
    std::vector<ViewDesc> vd;
    std::vector<Run> runs;
    uint64_t pix = 0, total = 0;
    for (int32_t i = 0; i < n_views; ++i) {
        const ct_view& v = views[i];
        if (!camera_ok(&v.camera) || v.width <= 0 || v.height <= 0 || v.spp <= 0 || !v.accum_out)
            return fail(c, CT_ERR_INVALID_ARG, "bad view");
        const uint64_t np = (uint64_t)v.width * v.height * (uint64_t)v.spp;
        if (np >= (1ull << 32)) return fail(c, CT_ERR_INVALID_ARG, "view too large");
        vd.push_back(make_view(v.camera, v.width, v.height, v.spp, v.frame_base, pix));
        runs.push_back({(uint32_t)i, 0u, (uint32_t)np});
        pix += (uint64_t)v.width * v.height;
        total += np;
    }
    float4* arena = nullptr;
    r = dalloc(c, (void**)&arena, pix * sizeof(float4), "batch accumulators");
    if (r != CT_OK) return r;
    hipError_t e = hipMemsetAsync(arena, 0, pix * sizeof(float4), c->ln[0].stream);
    if (e != hipSuccess) { dfree(c, arena); return hip_fail(c, e, "clear batch accumulators"); }
    const uint32_t cap = (uint32_t)std::min<uint64_t>(total, c->chunk_default);
    ChunkStats st;
    int chunks_before = 0;
    (void)chunks_before;
    r = run_paths(c, vd, runs, cap, arena, o, kRunGraph, &st);
    if (r == CT_OK) {
        for (int32_t i = 0; i < n_views && e == hipSuccess; ++i)
            e = hipMemcpyAsync(views[i].accum_out, arena + vd[i].accum_offset,
                               (size_t)views[i].width * views[i].height * sizeof(float4), hipMemcpyDeviceToHost,
                               c->ln[0].stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->ln[0].stream);
        if (e != hipSuccess) r = hip_fail(c, e, "read batch accumulators");
    }
    dfree(c, arena);
    if (r != CT_OK) return r;
    if (stats) {
        stats->primary_rays = st.primary;
        stats->extension_rays = st.extension;
        stats->shadow_rays = st.shadow;
        stats->samples = st.samples;
        stats->gpu_ms = st.total_ms;
        stats->chunks = st.chunks;
    }
    return CT_OK;
}

int ct_autotune(ct_ctx* c, const ct_camera* cam, const ct_frame_opts* opts, int32_t probe_width,
                int32_t probe_height, ct_tune_result* out) {
    if (!c || !camera_ok(cam) || probe_width <= 0 || probe_height <= 0) return CT_ERR_INVALID_ARG;
    ct_frame_opts o;
    if (opts) o = *opts;
    else ct_frame_opts_default(&o);
    o.spp = 1;
    int r = validate_opts(c, o);
    if (r != CT_OK) return r;
    (void)hipSetDevice(c->device);
    const uint64_t npix = (uint64_t)probe_width * probe_height;
    // the probe view is sampled several times so that every timed run has at
    // least 2M paths (shape) or 8M (chunk size): on a few hundred thousand
    // paths launch latency and tails decide, and the choice was a slow shape
    const uint64_t spp_shape = std::max<uint64_t>(1, ((1u << 21) + npix - 1) / npix);
    const uint64_t spp_chunk = std::max<uint64_t>(1, ((1u << 23) + npix - 1) / npix);
    if (npix * spp_chunk >= (1ull << 32)) return fail(c, CT_ERR_INVALID_ARG, "autotune probe too large");
    float4* acc = nullptr;
    r = dalloc(c, (void**)&acc, npix * sizeof(float4), "autotune accumulator");
    if (r != CT_OK) return r;
    const ctk::LaunchShape saved = c->shape;
    const uint32_t saved_chunk = c->chunk_default;
    std::vector<ViewDesc> views{make_view(*cam, probe_width, probe_height, (int)spp_shape, 0u, 0)};
    std::vector<Run> runs{{0u, 0u, (uint32_t)(npix * spp_shape)}};
    int trials = 0;
    // time of one shape: two runs after a warm-up run
    auto time_shape = [&](int b, int l, double* t) -> int {
        ctk::LaunchShape trial = saved;
        trial.block = b;
        trial.lds_depth = l;
        if (!ctk::set_trace_grids(trial, c->num_cus)) return CT_ERR_INVALID_ARG;
        c->shape = trial;
        *t = 0;
        for (int rep = 0; rep < 3; ++rep) {
            ChunkStats st;
            const int rr = run_paths(c, views, runs, (uint32_t)(npix * spp_shape), acc, o, 0, &st);
            if (rr != CT_OK) return rr;
            if (rep > 0) *t += st.total_ms;
        }
        ++trials;
        return CT_OK;
    };
    double t_saved = 0;
    if ((r = time_shape(saved.block, saved.lds_depth, &t_saved)) != CT_OK) {
        c->shape = saved;
        dfree(c, acc);
        return r;
    }
    double best = t_saved;
    int best_block = saved.block, best_lds = saved.lds_depth;
    const int blocks[] = {64, 128, 256};
    const int ldss[] = {8, 16, 24, 32};
    for (int b : blocks)
        for (int l : ldss) {
            if (b == saved.block && l == saved.lds_depth) continue;
            double t = 0;
            r = time_shape(b, l, &t);
            if (r == CT_ERR_INVALID_ARG) continue;  // does not fit a CU
            if (r != CT_OK) { c->shape = saved; dfree(c, acc); return r; }
            if (t < best) { best = t; best_block = b; best_lds = l; }
        }
    // a new shape only when it is at least 2% faster than the current one
    if (best > 0.98 * t_saved) { best = t_saved; best_block = saved.block; best_lds = saved.lds_depth; }
    ctr_apply_shape(c, best_block, best_lds);
    // chunk size: paths per wavefront chunk for multi-sample work; again the
    // current one unless another is at least 2% faster
    std::vector<ViewDesc> v2{make_view(*cam, probe_width, probe_height, (int)spp_chunk, 0u, 0)};
    std::vector<Run> r2{{0u, 0u, (uint32_t)(npix * spp_chunk)}};
    auto time_chunk = [&](uint32_t ch, double* t) -> int {
        ChunkStats st;
        const int rr = run_paths(c, v2, r2, ch, acc, o, 0, &st);
        if (rr == CT_OK) { *t = st.total_ms; ++trials; }
        return rr;
    };
    double best_c = 0;
    uint32_t best_chunk = saved_chunk;
    r = time_chunk(saved_chunk, &best_c);
    const double t_chunk_saved = best_c;
    const uint32_t chunks[] = {1u << 21, 1u << 22, 1u << 23};
    for (uint32_t ch : chunks) {
        if (r != CT_OK) break;
        if (ch == saved_chunk) continue;
        double t = 0;
        const int rr = time_chunk(ch, &t);
        if (rr == CT_ERR_OUT_OF_MEMORY) continue;  // path buffers beyond the memory budget
        if ((r = rr) != CT_OK) break;
        if (t < best_c) { best_c = t; best_chunk = ch; }
    }
    if (best_c > 0.98 * t_chunk_saved) best_chunk = saved_chunk;
    dfree(c, acc);
    if (r != CT_OK) { c->shape = saved; c->chunk_default = saved_chunk; return r; }
    c->chunk_default = best_chunk;
    if (out) {
        out->block_size = best_block;
        out->stack_lds = best_lds;
        out->chunk_paths = (int32_t)best_chunk;
        out->best_ms = best / 2.0;
        out->trials = trials;
    }
    return CT_OK;
}

}  // extern "C"
