// TriangleList::loadModel for Wavefront OBJ files (src/Core/TriangleList.cpp:63-146).
//
// The reference imports through Assimp with the post-process flags
// Triangulate, FlipWindingOrder, JoinIdenticalVertices, GenSmoothNormals and
// PreTransformVertices (normalised when `normalize`), among others
// (TriangleList.cpp:67-78).  Assimp is not available here, so this reader
// applies the same steps itself:
//   - polygons are split as fans (0, i, i + 1), then each triangle's index
//     order is reversed (Assimp's FlipWindingOrder reverses every face);
//   - corners with the same (position, texcoord, normal) indices become one
//     vertex, in order of first use;
//   - a file without normals gets per-position smooth normals (sum of the
//     unnormalised face normals, in the file's winding, of the faces sharing
//     the position);
//   - with `normalize`, the model is centred on its bounding box and divided
//     by half its largest extent, so it spans [-1, 1] (Assimp's
//     PretransformVertices normalisation);
// then, as in the reference: normals are normalised, positions become
// mat * (p * scale, 1) and normals mat * (n, 0) (not renormalised), centre
// normals are the mean of the three transformed vertex normals, and
// triangles whose material is a Light are registered as lights.
// With `material` < 0 the faces take their MTL materials (`mtllib` /
// `usemtl`), made as TriangleList::processMesh makes them
// (TriangleList.cpp:240-345); see mtl_material.  Not applied: Assimp's
// FixInfacingNormals heuristic.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <cctype>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "scene.h"

namespace ct {
namespace {

struct Corner {
    long v, t, n;  // 0-based, -1 = absent
    bool operator==(const Corner& o) const { return v == o.v && t == o.t && n == o.n; }
};
struct CornerHash {
    size_t operator()(const Corner& c) const {
        return (size_t)c.v * 0x9E3779B97F4A7C15ull ^ ((size_t)c.t + 0x1234567ull) * 0xC2B2AE3D27D4EB4Full ^
               ((size_t)c.n + 0x7654321ull) * 0x165667B19E3779F9ull;
    }
};

// OBJ index (1-based, negative = relative to the end) to 0-based; -1 if absent/bad.
long resolve(const char* s, size_t count) {
    if (!*s) return -1;
    char* end = nullptr;
    const long i = std::strtol(s, &end, 10);
    if (end == s || i == 0) return -2;
    const long r = i > 0 ? i - 1 : (long)count + i;
    return (r < 0 || (size_t)r >= count) ? -2 : r;
}

bool parse_corner(const std::string& tok, size_t nv, size_t nt, size_t nn, Corner& c) {
    std::string parts[3];
    int k = 0;
    for (char ch : tok) {
        if (ch == '/') {
            if (++k > 2) return false;
        } else {
            parts[k] += ch;
        }
    }
    c.v = resolve(parts[0].c_str(), nv);
    c.t = parts[1].empty() ? -1 : resolve(parts[1].c_str(), nt);
    c.n = parts[2].empty() ? -1 : resolve(parts[2].c_str(), nn);
    return c.v >= 0 && c.t != -2 && c.n != -2;
}

// What an MTL material holds, with the defaults of Assimp's OBJ importer for
// keys a file leaves out (diffuse 0.6, transmission filter 1, shininess 0,
// refraction index 1).
struct MtlDesc {
    vec3 kd = vec3(0.6f), ks = vec3(0.0f), ke = vec3(0.0f), tf = vec3(1.0f);
    float ns = 0.0f, ni = 1.0f;
    std::string map_kd, map_norm, map_d, map_disp;
};

std::string lower(std::string s) {
    for (char& ch : s) ch = (char)std::tolower((unsigned char)ch);
    return s;
}

// Read an MTL file into `out` (name -> material).  Texture maps keep the
// last token of the line (the file name after any options).
void read_mtl(const std::string& path, std::map<std::string, MtlDesc>& out) {
    std::ifstream in(path);
    if (!in) return;  // like Assimp, a missing library leaves the default material
    std::string line, tok;
    MtlDesc* cur = nullptr;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        if (!(ls >> tok) || tok[0] == '#') continue;
        const std::string key = lower(tok);
        if (key == "newmtl") {
            std::string name;
            std::getline(ls >> std::ws, name);
            while (!name.empty() && std::isspace((unsigned char)name.back())) name.pop_back();
            cur = &out[name];
            *cur = MtlDesc();
            continue;
        }
        if (!cur) continue;
        auto rgb = [&](vec3& v) {
            float r, g, b;
            if (ls >> r) {
                g = r;
                b = r;
                if (ls >> g) ls >> b;
                v = vec3(r, g, b);
            }
        };
        auto last = [&](std::string& dst) {
            std::string t, l;
            while (ls >> t) l = t;
            if (!l.empty()) dst = l;
        };
        if (key == "kd") rgb(cur->kd);
        else if (key == "ks") rgb(cur->ks);
        else if (key == "ke") rgb(cur->ke);
        else if (key == "tf") rgb(cur->tf);
        else if (key == "ns") ls >> cur->ns;
        else if (key == "ni") ls >> cur->ni;
        else if (key == "map_kd") last(cur->map_kd);
        else if (key == "norm" || key == "map_kn") last(cur->map_norm);
        else if (key == "map_d") last(cur->map_d);
        else if (key == "map_disp" || key == "disp") last(cur->map_disp);
    }
}

bool nonzero(vec3 c) { return c.x > 0.0001f && c.y > 0.0001f && c.z > 0.0001f; }  // isNotZero, TriangleList.cpp:231

// glm::mat4 is column-major: m[col * 4 + row].
vec3 xform(const float* m, vec3 v, float w) {
    return vec3(m[0] * v.x + m[4] * v.y + m[8] * v.z + m[12] * w, m[1] * v.x + m[5] * v.y + m[9] * v.z + m[13] * w,
                m[2] * v.x + m[6] * v.y + m[10] * v.z + m[14] * w);
}

}  // namespace

// TriangleList::processMesh's material (TriangleList.cpp:227-345): textures
// from the library's directory (a map that does not load gives -1), then the
// first of diffuse-textured / emissive / diffuse / specular / transmissive,
// added with addMaterial (which sets roughness 1).  The reference passes the
// texture indices of a transmissive material into the roughness and diffuse
// slots of Material::fresnel (TriangleList.cpp:333); here they go to the
// diffuse and normal slots.
static unsigned int mtl_material(TriangleList& L, const MtlDesc& m, const std::string& dir) {
    auto tex = [&](const std::string& f) { return f.empty() ? -1 : L.loadTexture(dir + "/" + f); };
    const int d = tex(m.map_kd), n = tex(m.map_norm), mk = tex(m.map_d), disp = tex(m.map_disp);
    Material mat;
    if (d > -1) {
        mat = Material::lambertian(vec3(1.0f), d, n, mk, disp);
        if (!(m.ns < 2.0f)) {
            mat.roughness = m.ns;
            mat.type = GGX;
        }
    } else if (nonzero(m.ke)) {
        mat = Material::light(m.ke);
    } else if (nonzero(m.kd)) {
        vec3 c = m.kd;
        if (dot(m.ks, m.ks) > dot(m.kd, m.kd)) c = m.ks;
        mat = Material::lambertian(c, d, n, mk, disp);
        if (!(m.ns < 1.0f)) {
            mat.roughness = m.ns;
            mat.type = GGX;
        }
    } else if (nonzero(m.ks)) {
        mat = Material::specular(m.ks, 1.0f);
    } else if (nonzero(m.tf)) {
        mat = Material::fresnel(m.tf, m.ni, m.tf, 0.0f, d, n);
        if (!(m.ns < 1.0f)) {
            mat.roughness = m.ns;
            mat.type = FresnelGGX;
        }
    } else {
        mat = Material::lambertian(vec3(0.0f));
    }
    return L.addMaterial(mat);
}

bool TriangleList::loadModel(const std::string& path, float scale, const float* mat, int material, bool normalize,
                             std::string* error) {
    auto fail = [&](const std::string& why) {
        if (error) *error = why;
        return false;
    };
    std::ifstream in(path);
    if (!in) return fail("cannot open " + path);
    if (material >= (int)m_Materials.size()) return fail("material index out of range");
    static const float identity[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    if (!mat) mat = identity;

    std::vector<vec3> pos, nrm;
    std::vector<vec2> tex;
    std::vector<Corner> corners;       // unique corners, in order of first use
    std::unordered_map<Corner, uint32_t, CornerHash> index;
    std::vector<uint32_t> tris;        // 3 corner ids per triangle, winding flipped
    std::vector<int> tri_mtl;          // per triangle: index into used_mtl, -1 = default material
    std::vector<std::string> used_mtl;
    std::map<std::string, MtlDesc> library;
    const size_t slash = path.find_last_of('/');
    const std::string dir = slash == std::string::npos ? std::string(".") : path.substr(0, slash);
    int cur_mtl = -1;
    std::string line, tok;
    size_t lineno = 0;
    std::vector<uint32_t> face;
    while (std::getline(in, line)) {
        ++lineno;
        std::istringstream ls(line);
        if (!(ls >> tok) || tok[0] == '#') continue;
        if (tok == "v" || tok == "vn") {
            vec3 v;
            if (!(ls >> v.x >> v.y >> v.z) || !std::isfinite(v.x) || !std::isfinite(v.y) || !std::isfinite(v.z))
                return fail("bad vertex at line " + std::to_string(lineno));
            (tok == "v" ? pos : nrm).push_back(v);
        } else if (tok == "vt") {
            vec2 t;
            if (!(ls >> t.x)) return fail("bad texcoord at line " + std::to_string(lineno));
            if (!(ls >> t.y)) t.y = 0.0f;
            tex.push_back(t);
        } else if (tok == "f") {
            face.clear();
            while (ls >> tok) {
                Corner c;
                if (!parse_corner(tok, pos.size(), tex.size(), nrm.size(), c))
                    return fail("bad face index at line " + std::to_string(lineno));
                auto it = index.find(c);
                if (it == index.end()) {
                    it = index.emplace(c, (uint32_t)corners.size()).first;
                    corners.push_back(c);
                }
                face.push_back(it->second);
            }
            if (face.size() < 3) return fail("face with fewer than 3 vertices at line " + std::to_string(lineno));
            for (size_t i = 1; i + 1 < face.size(); ++i) {
                tris.push_back(face[i + 1]);
                tris.push_back(face[i]);
                tris.push_back(face[0]);
                tri_mtl.push_back(cur_mtl);
            }
        } else if (tok == "mtllib") {
            while (ls >> tok) read_mtl(dir + "/" + tok, library);
        } else if (tok == "usemtl") {
            std::string name;
            std::getline(ls >> std::ws, name);
            while (!name.empty() && std::isspace((unsigned char)name.back())) name.pop_back();
            const auto it = std::find(used_mtl.begin(), used_mtl.end(), name);
            cur_mtl = (int)(it - used_mtl.begin());
            if (it == used_mtl.end()) used_mtl.push_back(name);
        }
        // g, o, s, l, p: no geometry of their own
    }
    if (tris.empty()) return fail("no faces in " + path);

    // vertex attributes of the unique corners
    const size_t nc = corners.size();
    std::vector<vec3> vp(nc), vn(nc);
    std::vector<vec2> vt(nc, vec2(0.5f, 0.5f));  // the reference's default uv
    bool have_normals = true;
    for (size_t i = 0; i < nc; ++i) {
        vp[i] = pos[(size_t)corners[i].v];
        if (corners[i].t >= 0) vt[i] = tex[(size_t)corners[i].t];
        if (corners[i].n >= 0) vn[i] = nrm[(size_t)corners[i].n];
        else have_normals = false;
    }
    if (normalize) {
        vec3 mn(1e30f), mx(-1e30f);
        for (const vec3& p : vp) { mn = vmin(mn, p); mx = vmax(mx, p); }
        const vec3 delta = mx - mn;
        const vec3 half = mn + delta * 0.5f;
        const float div = gmax(delta.x, gmax(delta.y, delta.z)) * 0.5f;
        if (div > 0.0f)
            for (vec3& p : vp) p = (p - half) / div;
    }
    if (!have_normals) {
        // smooth normals per position index (corners sharing a position share it)
        std::vector<vec3> acc(pos.size(), vec3(0.0f));
        for (size_t t = 0; t < tris.size(); t += 3) {
            const vec3 a = vp[tris[t]], b = vp[tris[t + 1]], c = vp[tris[t + 2]];
            // the file's own winding (normals are generated before the flip)
            const vec3 fn = cross(c - a, b - a);
            for (int k = 0; k < 3; ++k) acc[(size_t)corners[tris[t + k]].v] = acc[(size_t)corners[tris[t + k]].v] + fn;
        }
        for (size_t i = 0; i < nc; ++i)
            if (corners[i].n < 0) vn[i] = acc[(size_t)corners[i].v];
    }

    // materials of the faces: the given one, or the MTL materials in order of
    // first use (faces before any usemtl, or naming an unknown material, get
    // the importer's default material)
    std::vector<unsigned int> mat_of(used_mtl.size() + 1, material < 0 ? 0u : (unsigned int)material);
    if (material < 0) {
        bool need_default = std::find(tri_mtl.begin(), tri_mtl.end(), -1) != tri_mtl.end();
        for (const std::string& nm : used_mtl) need_default |= library.find(nm) == library.end();
        if (need_default) mat_of[0] = mtl_material(*this, MtlDesc(), dir);
        for (size_t k = 0; k < used_mtl.size(); ++k) {
            const auto it = library.find(used_mtl[k]);
            mat_of[k + 1] = it == library.end() ? mat_of[0] : mtl_material(*this, it->second, dir);
        }
    }

    // append like TriangleList.cpp:93-145
    const uint32_t offset = (uint32_t)m_Vertices.size();
    for (size_t i = 0; i < nc; ++i) {
        const float len = length(vn[i]);
        const vec3 n = len > 0.0f ? ct::normalize(vn[i]) : vec3(0.0f, 1.0f, 0.0f);
        m_Vertices.push_back(xform(mat, vp[i] * scale, 1.0f));
        m_Normals.push_back(xform(mat, n, 0.0f));
        m_TexCoords.push_back(vt[i]);
    }
    for (size_t t = 0; t < tris.size(); t += 3) {
        const unsigned int matIdx = mat_of[(size_t)(tri_mtl[t / 3] + 1)];
        const uint32_t i0 = tris[t] + offset, i1 = tris[t + 1] + offset, i2 = tris[t + 2] + offset;
        m_Indices.push_back(uvec3(i0, i1, i2));
        if (m_Materials[matIdx].type == Light) m_LightIndices.push_back((unsigned int)(m_Indices.size() - 1));
        m_MaterialIdxs.push_back(matIdx);
        m_CenterNormals.push_back((m_Normals[i0] + m_Normals[i1] + m_Normals[i2]) / 3.0f);
        m_AABBs.push_back(triangle_bounds(m_Vertices[i0], m_Vertices[i1], m_Vertices[i2]));
    }
    return true;
}

}  // namespace ct
