#include "scenegen.h"

#include <cmath>
#include <functional>
#include <vector>

namespace ct {

namespace {

vec3 face_normal(vec3 a, vec3 b, vec3 c) { return normalize(cross(b - a, c - a)); }

// Flat-shaded triangle whose normal is oriented along `towards`.
void add_flat(TriangleList& L, vec3 a, vec3 b, vec3 c, vec3 towards, unsigned mat, vec2 ta = vec2(),
              vec2 tb = vec2(), vec2 tc = vec2()) {
    vec3 n = face_normal(a, b, c);
    if (dot(n, towards) < 0.0f) n = -n;
    L.addTriangle(a, b, c, n, n, n, mat, ta, tb, tc);
}

void add_quad(TriangleList& L, vec3 p0, vec3 p1, vec3 p2, vec3 p3, vec3 normal, unsigned mat,
              const vec2* uv = nullptr) {
    const vec2 z;
    add_flat(L, p0, p1, p2, normal, mat, uv ? uv[0] : z, uv ? uv[1] : z, uv ? uv[2] : z);
    add_flat(L, p0, p2, p3, normal, mat, uv ? uv[0] : z, uv ? uv[2] : z, uv ? uv[3] : z);
}

// Axis-aligned-in-local-frame box rotated by `yaw` about +y; outward normals.
void add_box(TriangleList& L, vec3 c, vec3 h, float yaw, unsigned mat) {
    const float cs = std::cos(yaw), sn = std::sin(yaw);
    auto P = [&](float x, float y, float z) {
        const float rx = cs * x + sn * z, rz = -sn * x + cs * z;
        return vec3(c.x + rx, c.y + y, c.z + rz);
    };
    vec3 v[8];
    for (int i = 0; i < 8; ++i) v[i] = P((i & 1) ? h.x : -h.x, (i & 2) ? h.y : -h.y, (i & 4) ? h.z : -h.z);
    const int f[6][4] = {{0, 1, 3, 2}, {4, 6, 7, 5}, {0, 4, 5, 1}, {2, 3, 7, 6}, {0, 2, 6, 4}, {1, 5, 7, 3}};
    for (int k = 0; k < 6; ++k) {
        const vec3 a = v[f[k][0]], b = v[f[k][1]], cc = v[f[k][2]], d = v[f[k][3]];
        const vec3 centre = (a + b + cc + d) * 0.25f;
        add_quad(L, a, b, cc, d, centre - c, mat);
    }
}

// Open cylinder along +y: `seg` segments around, `rings` along the height;
// texture coordinates u = 2 * angle / 2pi, v = height fraction.
void add_cylinder(TriangleList& L, vec3 base, float r, float height, int seg, int rings, unsigned mat) {
    const float twopi = 6.28318530717958647692f;
    for (int j = 0; j < rings; ++j) {
        const float y0 = base.y + height * (float)j / (float)rings;
        const float y1 = base.y + height * (float)(j + 1) / (float)rings;
        for (int i = 0; i < seg; ++i) {
            const float a0 = twopi * (float)i / (float)seg, a1 = twopi * (float)(i + 1) / (float)seg;
            const vec3 p0(base.x + r * std::cos(a0), y0, base.z + r * std::sin(a0));
            const vec3 p1(base.x + r * std::cos(a1), y0, base.z + r * std::sin(a1));
            const vec3 p2(base.x + r * std::cos(a1), y1, base.z + r * std::sin(a1));
            const vec3 p3(base.x + r * std::cos(a0), y1, base.z + r * std::sin(a0));
            const float am = 0.5f * (a0 + a1);
            const float u0 = 2.0f * (float)i / (float)seg, u1 = 2.0f * (float)(i + 1) / (float)seg;
            const float v0 = (float)j / (float)rings, v1 = (float)(j + 1) / (float)rings;
            const vec2 uv[4] = {vec2(u0, v0), vec2(u1, v0), vec2(u1, v1), vec2(u0, v1)};
            add_quad(L, p0, p1, p2, p3, vec3(std::cos(am), 0.0f, std::sin(am)), mat, uv);
        }
    }
}

// Icosphere: icosahedron split `subdiv` times (20 * 4^subdiv faces).  The
// midpoint of an edge is normalize(a + b), symmetric in (a, b), so
// neighbouring faces share exact vertex positions without an index map.
void icosphere_faces(int subdiv, const std::function<void(vec3, vec3, vec3)>& emit) {
    const float t = (1.0f + std::sqrt(5.0f)) * 0.5f;
    vec3 v[12] = {vec3(-1, t, 0), vec3(1, t, 0), vec3(-1, -t, 0), vec3(1, -t, 0),
                  vec3(0, -1, t), vec3(0, 1, t), vec3(0, -1, -t), vec3(0, 1, -t),
                  vec3(t, 0, -1), vec3(t, 0, 1), vec3(-t, 0, -1), vec3(-t, 0, 1)};
    for (auto& p : v) p = normalize(p);
    const int f[20][3] = {{0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
                          {11, 10, 2}, {10, 7, 6}, {7, 1, 8},  {3, 9, 4},  {3, 4, 2},   {3, 2, 6}, {3, 6, 8},
                          {3, 8, 9},  {4, 9, 5},  {2, 4, 11}, {6, 2, 10}, {8, 6, 7},   {9, 8, 1}};
    std::function<void(vec3, vec3, vec3, int)> rec = [&](vec3 a, vec3 b, vec3 c, int d) {
        if (d == 0) { emit(a, b, c); return; }
        const vec3 ab = normalize(a + b), bc = normalize(b + c), ca = normalize(c + a);
        rec(a, ab, ca, d - 1);
        rec(b, bc, ab, d - 1);
        rec(c, ca, bc, d - 1);
        rec(ab, bc, ca, d - 1);
    };
    for (auto& fi : f) rec(v[fi[0]], v[fi[1]], v[fi[2]], subdiv);
}

// 3-D value noise on the integer lattice with smoothstep interpolation.
struct ValueNoise {
    uint64_t seed;
    float lattice(int x, int y, int z) const {
        uint64_t h = seed ^ ((uint64_t)(uint32_t)x * 0x9e3779b97f4a7c15ull) ^
                     ((uint64_t)(uint32_t)y * 0xc2b2ae3d27d4eb4full) ^ ((uint64_t)(uint32_t)z * 0x165667b19e3779f9ull);
        h = (h ^ (h >> 30)) * 0xbf58476d1ce4e5b9ull;
        h = (h ^ (h >> 27)) * 0x94d049bb133111ebull;
        h ^= h >> 31;
        return (float)(h >> 40) * (1.0f / 16777216.0f) * 2.0f - 1.0f;
    }
    static float smooth(float t) { return t * t * (3.0f - 2.0f * t); }
    float eval(float x, float y, float z) const {
        const float fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
        const int ix = (int)fx, iy = (int)fy, iz = (int)fz;
        const float tx = smooth(x - fx), ty = smooth(y - fy), tz = smooth(z - fz);
        float c[2][2][2];
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                for (int d = 0; d < 2; ++d) c[a][b][d] = lattice(ix + a, iy + b, iz + d);
        auto lerp = [](float p, float q, float t) { return p + (q - p) * t; };
        const float x00 = lerp(c[0][0][0], c[1][0][0], tx), x10 = lerp(c[0][1][0], c[1][1][0], tx);
        const float x01 = lerp(c[0][0][1], c[1][0][1], tx), x11 = lerp(c[0][1][1], c[1][1][1], tx);
        return lerp(lerp(x00, x10, ty), lerp(x01, x11, ty), tz);
    }
    float fbm(vec3 p, int octaves) const {
        float sum = 0.0f, amp = 0.5f, freq = 1.0f, norm = 0.0f;
        for (int o = 0; o < octaves; ++o) {
            sum += amp * eval(p.x * freq, p.y * freq, p.z * freq);
            norm += amp;
            amp *= 0.5f;
            freq *= 2.0f;
        }
        return sum / norm;
    }
};

ct_frame_opts base_opts() {
    ct_frame_opts o;
    ct_frame_opts_default(&o);
    return o;
}

vec3 rand_albedo(SplitMix64& rng, float lo, float hi) {
    const float r = rng.uniform(lo, hi), g = rng.uniform(lo, hi), b = rng.uniform(lo, hi);
    return vec3(r, g, b);
}

// C1: closed box (12 tris) + icosphere subdivision 5 (20,480 tris) + 1 quad
// light (2 tris, added first), radiance 15, Lambert albedo U[0.2, 0.8].
void gen_c1(uint64_t seed, TriangleList& L, SceneConfig& cfg) {
    SplitMix64 rng(seed);
    const unsigned lightMat = L.addMaterial(Material::light(vec3(15.0f)));
    add_quad(L, vec3(-15, 49.5f, -15), vec3(15, 49.5f, -15), vec3(15, 49.5f, 15), vec3(-15, 49.5f, 15),
             vec3(0, -1, 0), lightMat);
    const float s = 50.0f;
    unsigned wall[6];
    for (int i = 0; i < 6; ++i) wall[i] = L.addMaterial(Material::lambertian(rand_albedo(rng, 0.2f, 0.8f)));
    add_quad(L, vec3(-s, -s, -s), vec3(s, -s, -s), vec3(s, -s, s), vec3(-s, -s, s), vec3(0, 1, 0), wall[0]);
    add_quad(L, vec3(-s, s, -s), vec3(-s, s, s), vec3(s, s, s), vec3(s, s, -s), vec3(0, -1, 0), wall[1]);
    add_quad(L, vec3(-s, -s, -s), vec3(-s, -s, s), vec3(-s, s, s), vec3(-s, s, -s), vec3(1, 0, 0), wall[2]);
    add_quad(L, vec3(s, -s, -s), vec3(s, s, -s), vec3(s, s, s), vec3(s, -s, s), vec3(-1, 0, 0), wall[3]);
    add_quad(L, vec3(-s, -s, -s), vec3(-s, s, -s), vec3(s, s, -s), vec3(s, -s, -s), vec3(0, 0, 1), wall[4]);
    add_quad(L, vec3(-s, -s, s), vec3(s, -s, s), vec3(s, s, s), vec3(-s, s, s), vec3(0, 0, -1), wall[5]);
    const unsigned sphereMat = L.addMaterial(Material::lambertian(rand_albedo(rng, 0.2f, 0.8f)));
    const float R = 25.0f;
    icosphere_faces(5, [&](vec3 a, vec3 b, vec3 c) {
        const vec3 pa = a * R, pb = b * R, pc = c * R;
        add_flat(L, pa, pb, pc, (pa + pb + pc), sphereMat);
    });
    cfg.cam_pos = vec3(0.0f, 5.0f, 48.0f);
    cfg.cam_target = vec3(0.0f, 0.0f, 0.0f);
    cfg.cam_fov = 70.0f;
    cfg.sky_enabled = 0;
    cfg.opts = base_opts();
    cfg.opts.max_depth = 4;
    cfg.opts.spp = 1;
    cfg.opts.accel = CT_ACCEL_BVH2;
    cfg.opts.bsdf_mode = CT_BSDF_REFERENCE;
    // distance epsilon scaled to the world (distances ~1e2, float ulp ~1e-5)
    cfg.opts.dist_epsilon = 1e-3f;
    cfg.width = 256;
    cfg.height = 256;
}

// Random material of the C2 mix: 60% Lambert, 25% GGX r~U[0.05,0.8],
// 5% specular, 10% dielectric (IOR 1.5).
unsigned c2_material(TriangleList& L, SplitMix64& rng) {
    const float u = rng.uniform();
    if (u < 0.60f) return L.addMaterial(Material::lambertian(rand_albedo(rng, 0.2f, 0.8f)), true);
    if (u < 0.85f) {
        const vec3 a = rand_albedo(rng, 0.5f, 0.95f);
        return L.addMaterial(Material::microfacet(a, GGX, rng.uniform(0.05f, 0.8f)), true);
    }
    if (u < 0.90f) return L.addMaterial(Material::specular(rand_albedo(rng, 0.8f, 0.95f), 1.0f), true);
    return L.addMaterial(Material::fresnel(vec3(1.0f), 1.5f), true);
}

// Tessellated plane at height y over [-ext, ext]^2 with n x n cells,
// skipping cells inside |x|,|z| < hole; texture coordinates (x, z) * uv_scale.
void add_grid(TriangleList& L, float y, float ext, int n, float hole, vec3 normal, unsigned mat,
              const std::function<float(float, float)>& height = nullptr, float uv_scale = 0.0f) {
    const float step = 2.0f * ext / (float)n;
    auto P = [&](int i, int j) {
        const float x = -ext + step * (float)i, z = -ext + step * (float)j;
        return vec3(x, y + (height ? height(x, z) : 0.0f), z);
    };
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            const float cx = -ext + step * ((float)i + 0.5f), cz = -ext + step * ((float)j + 0.5f);
            if (hole > 0.0f && std::fabs(cx) < hole && std::fabs(cz) < hole) continue;
            const vec3 q[4] = {P(i, j), P(i + 1, j), P(i + 1, j + 1), P(i, j + 1)};
            vec2 uv[4];
            for (int k = 0; k < 4; ++k) uv[k] = vec2(q[k].x * uv_scale, q[k].z * uv_scale);
            add_quad(L, q[0], q[1], q[2], q[3], normal, mat, uv);
        }
}

// Procedural textures of the textured atrium (RGBA, row 0 at the bottom).
// Colour textures are quantised to 8 bits (c / 255), like image files.
float q8(float v) { return std::nearbyint(std::fmin(std::fmax(v, 0.0f), 1.0f) * 255.0f) / 255.0f; }
// Floor tiles: 8 x 8 tiles of two tones with dark grout and value-noise grain.
int tiles_texture(TriangleList& L, uint64_t seed) {
    const uint32_t n = 1024;
    const ValueNoise vn{seed ^ 0x7431ull};
    std::vector<float> t((size_t)n * n * 4);
    for (uint32_t y = 0; y < n; ++y)
        for (uint32_t x = 0; x < n; ++x) {
            const bool grout = (x % 128) < 4 || (y % 128) < 4;
            const bool dark = ((x / 128) + (y / 128)) & 1;
            const float g = 0.06f * vn.eval((float)x / 24.0f, (float)y / 24.0f, 0.5f);
            const vec3 c = grout ? vec3(0.12f) : dark ? vec3(0.32f, 0.30f, 0.28f) : vec3(0.72f, 0.68f, 0.62f);
            float* p = &t[4 * ((size_t)x + (size_t)y * n)];
            p[0] = q8(c.x + g);
            p[1] = q8(c.y + g);
            p[2] = q8(c.z + g);
            p[3] = 1.0f;
        }
    return L.addTexture(n, n, t.data());
}

// Tangent-space normals of a bump field h = 0.5 sin(2pi x/64) sin(2pi y/64),
// stored as floats (the reference normalises the raw texel colour).
int bumps_texture(TriangleList& L) {
    const uint32_t n = 512;
    const float k = 6.28318530717958647692f / 64.0f;
    std::vector<float> t((size_t)n * n * 4);
    for (uint32_t y = 0; y < n; ++y)
        for (uint32_t x = 0; x < n; ++x) {
            const float dx = 0.5f * k * std::cos(k * (float)x) * std::sin(k * (float)y);
            const float dy = 0.5f * k * std::sin(k * (float)x) * std::cos(k * (float)y);
            const vec3 nn = normalize(vec3(-4.0f * dx, -4.0f * dy, 1.0f));
            float* p = &t[4 * ((size_t)x + (size_t)y * n)];
            p[0] = nn.x;
            p[1] = nn.y;
            p[2] = nn.z;
            p[3] = 1.0f;
        }
    return L.addTexture(n, n, t.data());
}

// Column stone: banded fbm noise.
int marble_texture(TriangleList& L, uint64_t seed) {
    const uint32_t n = 512;
    const ValueNoise vn{seed ^ 0x3a7bull};
    std::vector<float> t((size_t)n * n * 4);
    for (uint32_t y = 0; y < n; ++y)
        for (uint32_t x = 0; x < n; ++x) {
            const float f = vn.fbm(vec3((float)x / 64.0f, (float)y / 64.0f, 0.25f), 4);
            const float b = 0.5f + 0.5f * std::sin(0.05f * (float)y + 6.0f * f);
            float* p = &t[4 * ((size_t)x + (size_t)y * n)];
            p[0] = q8(0.55f + 0.35f * b);
            p[1] = q8(0.52f + 0.33f * b);
            p[2] = q8(0.50f + 0.30f * b);
            p[3] = 1.0f;
        }
    return L.addTexture(n, n, t.data());
}

// Equirectangular sky (u = azimuth, v = 1 - polar angle / pi): zenith-to-
// horizon gradient over a dim ground, with faint azimuthal banding, 8-bit.
int sky_texture(TriangleList& L) {
    const uint32_t w = 1024, h = 512;
    const float pi = 3.14159265358979323846f;
    std::vector<float> t((size_t)w * h * 4);
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x) {
            const float v = ((float)y + 0.5f) / (float)h, u = ((float)x + 0.5f) / (float)w;
            const float elev = (v - 0.5f) * pi;  // -pi/2 (nadir) .. pi/2 (zenith)
            vec3 c;
            if (elev < 0.0f) {
                c = vec3(0.22f, 0.2f, 0.18f);
            } else {
                const float s = std::sin(elev);
                c = vec3(0.75f, 0.8f, 0.88f) * (1.0f - s) + vec3(0.22f, 0.36f, 0.66f) * s;
            }
            c = c * (0.92f + 0.08f * std::sin(6.28318530717958647692f * 4.0f * u));
            float* p = &t[4 * ((size_t)x + (size_t)y * w)];
            p[0] = q8(c.x);
            p[1] = q8(c.y);
            p[2] = q8(c.z);
            p[3] = 1.0f;
        }
    return L.addTexture(w, h, t.data());
}

// C2: atrium of ~260k triangles: 2 tessellated floors, 24 columns of 2,048
// triangles, 400 random boxes (log-uniform size), 4 quad lights + constant sky.
// textured: the same geometry with tiled, bump-mapped floors, stone columns
// (the Lambert and GGX ones) and a sky texture instead of the constant sky.
void gen_c2(uint64_t seed, TriangleList& L, SceneConfig& cfg, bool textured = false) {
    SplitMix64 rng(seed);
    // lights first (light ordinal == triangle id)
    const unsigned lightMat = L.addMaterial(Material::light(vec3(40.0f, 38.0f, 34.0f)));
    const float ly = 29.0f, lh = 6.0f;
    const float lxz[4][2] = {{-65, -65}, {65, -65}, {-65, 65}, {65, 65}};
    for (auto& c : lxz)
        add_quad(L, vec3(c[0] - lh, ly, c[1] - lh), vec3(c[0] + lh, ly, c[1] - lh), vec3(c[0] + lh, ly, c[1] + lh),
                 vec3(c[0] - lh, ly, c[1] + lh), vec3(0, -1, 0), lightMat);
    int tiles = -1, bumps = -1, marble = -1;
    if (textured) {
        tiles = tiles_texture(L, seed);
        bumps = bumps_texture(L);
        marble = marble_texture(L, seed);
        cfg.sky_texture = sky_texture(L);
    }
    const unsigned floorMat = L.addMaterial(
        textured ? Material::lambertian(vec3(1.0f), tiles, bumps) : Material::lambertian(vec3(0.6f, 0.58f, 0.55f)),
        true);
    const unsigned upperMat = L.addMaterial(
        textured ? Material::lambertian(vec3(1.0f), tiles, bumps) : Material::lambertian(vec3(0.5f, 0.52f, 0.6f)),
        true);
    add_grid(L, 0.0f, 100.0f, 236, 0.0f, vec3(0, 1, 0), floorMat, nullptr, 0.125f);
    add_grid(L, 30.0f, 100.0f, 236, 40.0f, vec3(0, -1, 0), upperMat, nullptr, 0.125f);
    // 24 columns: 64 segments x 16 rings x 2 = 2048 triangles each
    for (int iz = 0; iz < 4; ++iz)
        for (int ix = 0; ix < 6; ++ix) {
            const float x = -75.0f + 30.0f * (float)ix, z = -67.5f + 45.0f * (float)iz;
            const unsigned m = c2_material(L, rng);
            if (textured && (L.m_Materials[m].type == Lambertian || L.m_Materials[m].type == GGX))
                L.m_Materials[m].diffuseTex = marble;
            add_cylinder(L, vec3(x, 0.0f, z), 2.0f, 30.0f, 64, 16, m);
        }
    // 400 boxes, edge lengths log-uniform in [0.1, 2] * 5 world units
    for (int b = 0; b < 400; ++b) {
        const float lo = std::log(0.5f), hi = std::log(10.0f);
        const vec3 size(std::exp(rng.uniform(lo, hi)), std::exp(rng.uniform(lo, hi)), std::exp(rng.uniform(lo, hi)));
        const float x = rng.uniform(-95.0f, 95.0f), z = rng.uniform(-95.0f, 95.0f);
        const float yaw = rng.uniform(0.0f, 6.2831853f);
        const bool upper = rng.uniform() < 0.25f && (std::fabs(x) > 45.0f || std::fabs(z) > 45.0f);
        const float y0 = upper ? 30.0f + 0.05f : 0.05f;
        add_box(L, vec3(x, y0 + size.y * 0.5f, z), size * 0.5f, yaw, c2_material(L, rng));
    }
    cfg.cam_pos = vec3(-85.0f, 14.0f, -80.0f);
    cfg.cam_target = vec3(10.0f, 9.0f, 15.0f);
    cfg.cam_fov = 70.0f;
    cfg.sky_enabled = 1;
    cfg.sky = vec3(0.35f, 0.4f, 0.5f);
    cfg.opts = base_opts();
    cfg.opts.max_depth = 8;
    cfg.opts.spp = 1;
    cfg.opts.accel = CT_ACCEL_BVH4;
    cfg.opts.bsdf_mode = CT_BSDF_TYPED;
    cfg.opts.sort_by_material = 1;  // mixed BSDFs: shade in material order
    cfg.opts.normal_epsilon = 1e-3f;
    cfg.opts.dist_epsilon = 1e-2f;
    cfg.width = 1920;
    cfg.height = 1080;
}

// C3: icosphere subdivision 9 (5,242,880 tris) displaced by 4-octave value
// noise (5% amplitude) on a 2M-triangle noisy ground; 50/50 Beckmann/GGX with
// roughness U[0.02, 0.5]; one small area light.
void gen_c3(uint64_t seed, TriangleList& L, SceneConfig& cfg) {
    SplitMix64 rng(seed);
    const unsigned lightMat = L.addMaterial(Material::light(vec3(800.0f)));
    add_quad(L, vec3(-15, 260, -15), vec3(15, 260, -15), vec3(15, 260, 15), vec3(-15, 260, 15), vec3(0, -1, 0),
             lightMat);
    unsigned mats[32];
    for (int i = 0; i < 32; ++i) {
        const MaterialType t = (i & 1) ? GGX : Beckmann;
        mats[i] = L.addMaterial(Material::microfacet(rand_albedo(rng, 0.3f, 0.9f), t, rng.uniform(0.02f, 0.5f)), true);
    }
    const ValueNoise noise{seed ^ 0x5eedull};
    const float R = 100.0f;
    size_t face = 0;
    auto disp = [&](vec3 u) { return u * (R * (1.0f + 0.05f * noise.fbm(u * 4.0f, 4))); };
    icosphere_faces(9, [&](vec3 a, vec3 b, vec3 c) {
        const vec3 pa = disp(a), pb = disp(b), pc = disp(c);
        add_flat(L, pa, pb, pc, a + b + c, mats[(face >> 14) & 31]);
        ++face;
    });
    const ValueNoise gnoise{seed ^ 0x9001ull};
    add_grid(L, -115.0f, 400.0f, 1000, 0.0f, vec3(0, 1, 0), mats[(seed >> 3) & 31],
             [&](float x, float z) { return 3.0f * gnoise.fbm(vec3(x * 0.02f, 0.0f, z * 0.02f), 4); });
    cfg.cam_pos = vec3(0.0f, 40.0f, 260.0f);
    cfg.cam_target = vec3(0.0f, -10.0f, 0.0f);
    cfg.cam_fov = 60.0f;
    cfg.sky_enabled = 1;
    cfg.sky = vec3(0.05f, 0.06f, 0.08f);
    cfg.opts = base_opts();
    cfg.opts.max_depth = 16;
    cfg.opts.spp = 1;
    cfg.opts.rr_start = 3;
    cfg.opts.accel = CT_ACCEL_BVH4;
    cfg.opts.bsdf_mode = CT_BSDF_TYPED;
    // two microfacet types that shade along nearly the same code: grouping
    // them saves less than the sort's key lookups cost (shading 8% faster
    // without the sort on this scene)
    cfg.opts.sort_by_material = 0;
    cfg.opts.normal_epsilon = 1e-3f;
    cfg.opts.dist_epsilon = 1e-2f;
    cfg.width = 3840;
    cfg.height = 2160;
}

}  // namespace

bool generate_scene(int config, uint64_t seed, TriangleList& list, SceneConfig& cfg) {
    cfg.config = config;
    cfg.seed = seed;
    switch (config) {
        case CT_SCENE_C1_BOX_SPHERE: gen_c1(seed, list, cfg); return true;
        case CT_SCENE_C2_ATRIUM: gen_c2(seed, list, cfg); return true;
        case CT_SCENE_C2_TEXTURED: gen_c2(seed, list, cfg, true); return true;
        case CT_SCENE_C3_DISPLACED: gen_c3(seed, list, cfg); return true;
        default: return false;
    }
}

}  // namespace ct
