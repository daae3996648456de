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
// Not applied: MTL materials and textures (`material` < 0 gives material 0)
// and Assimp's FixInfacingNormals heuristic.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
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

// glm::mat4 is column-major: m[col * 4 + row].
vec3 xform(const float* m, vec3 v, float w) {
    return vec3(m[0] * v.x + m[4] * v.y + m[8] * v.z + m[12] * w, m[1] * v.x + m[5] * v.y + m[9] * v.z + m[13] * w,
                m[2] * v.x + m[6] * v.y + m[10] * v.z + m[14] * w);
}

}  // namespace

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
            }
        }
        // g, o, s, usemtl, mtllib, l, p: no geometry of their own
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

    // append like TriangleList.cpp:93-145
    const uint32_t offset = (uint32_t)m_Vertices.size();
    for (size_t i = 0; i < nc; ++i) {
        const float len = length(vn[i]);
        const vec3 n = len > 0.0f ? ct::normalize(vn[i]) : vec3(0.0f, 1.0f, 0.0f);
        m_Vertices.push_back(xform(mat, vp[i] * scale, 1.0f));
        m_Normals.push_back(xform(mat, n, 0.0f));
        m_TexCoords.push_back(vt[i]);
    }
    const unsigned int matIdx = material < 0 ? 0u : (unsigned int)material;
    for (size_t t = 0; t < tris.size(); t += 3) {
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
