// CPU oracle of the wavefront path-tracing loop — TEST INFRASTRUCTURE ONLY
// (see oracle.h).  Scalar restatement of the reference, templated on the
// arithmetic type T (float: bit-comparable with the HIP kernels; double:
// accuracy reference).  Every routine cites the reference file:line it
// restates.  Decisions on the reference's quirks (the "Q" list) are noted
// where they apply.
#include "oracle.h"

#include <atomic>
#include <cmath>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/ct_fp.h"

namespace {

// ------------------------------------------------------------- numerics --
inline float m_sin(float x) { return ct_sinf(x); }
inline double m_sin(double x) { return std::sin(x); }
inline float m_cos(float x) { return ct_cosf(x); }
inline double m_cos(double x) { return std::cos(x); }
inline float m_exp(float x) { return ct_expf(x); }
inline double m_exp(double x) { return std::exp(x); }
inline float m_log(float x) { return ct_logf(x); }
inline double m_log(double x) { return std::log(x); }
inline float m_pow(float x, float y) { return ct_powf(x, y); }
inline double m_pow(double x, double y) { return std::pow(x, y); }
inline float m_acos(float x) { return ct_acosf(x); }
inline double m_acos(double x) { return std::acos(x); }
inline float m_sqrt(float x) { return ct_sqrtf(x); }
inline double m_sqrt(double x) { return std::sqrt(x); }
inline float m_abs(float x) { return ct_fabsf(x); }
inline double m_abs(double x) { return std::fabs(x); }

template <class T> struct K;
template <> struct K<float> {
    static constexpr float pi = CT_PI_F, two_pi = CT_TWO_PI_F, inv_pi = CT_INV_PI_F, root_pi = CT_ROOT_PI_F;
};
template <> struct K<double> {
    static constexpr double pi = 3.14159265358979323846, two_pi = 6.28318530717958647692,
                            inv_pi = 0.318309886183790671538, root_pi = 1.77245385090551602730;
};

// Microfacet.cuh:69-82 (A&S 7.1.26) and :36-67 (Giles); float uses ct_fp.h.
inline float m_erf(float x) { return ct_erff(x); }
inline double m_erf(double x) {
    double sign = 1.0;
    if (x < 0.0) sign = -1.0;
    x = std::fabs(x);
    const double t = 1.0 / (1.0 + 0.3275911 * x);
    const double y = 1.0 - (((((1.061405429 * t + -1.453152027) * t) + 1.421413741) * t + -0.284496736) * t +
                            0.254829592) * t * std::exp(-x * x);
    return sign * y;
}
inline float m_erfinv(float x) { return ct_erfinvf(x); }
inline double m_erfinv(double x) {
    double w, p;
    x = x < -0.999 ? -0.999 : (x > 0.999 ? 0.999 : x);
    w = -std::log((1.0 - x) * (1.0 + x));
    if (w < 5.0) {
        w = w - 2.5;
        p = 2.81022636e-08;
        p = 3.43273939e-07 + p * w;
        p = -3.5233877e-06 + p * w;
        p = -4.39150654e-06 + p * w;
        p = 0.00021858087 + p * w;
        p = -0.00125372503 + p * w;
        p = -0.00417768164 + p * w;
        p = 0.246640727 + p * w;
        p = 1.50140941 + p * w;
    } else {
        w = std::sqrt(w) - 3.0;
        p = -0.000200214257;
        p = 0.000100950558 + p * w;
        p = 0.00134934322 + p * w;
        p = -0.00367342844 + p * w;
        p = 0.00573950773 + p * w;
        p = -0.0076224613 + p * w;
        p = 0.00943887047 + p * w;
        p = 1.00167406 + p * w;
        p = 2.83297682 + p * w;
    }
    return p * x;
}

// ------------------------------------------------------ glm-order vec3 --
// include/glm/detail/func_geometric.inl:48-54 (dot), :82-90 (cross),
// :104-110 (normalize = v * inversesqrt(dot)), func_common.inl:16-28 (min/max).
template <class T> struct V3 {
    T x, y, z;
};
template <class T> inline V3<T> v3(T a, T b, T c) { return V3<T>{a, b, c}; }
template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> operator*(V3<T> a, V3<T> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class T> inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T> inline V3<T> operator/(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> inline V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline V3<T> cross(V3<T> x, V3<T> y) {
    return {x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y};
}
template <class T> inline V3<T> normalize(V3<T> v) { return v * (T(1) / m_sqrt(dot(v, v))); }
template <class T> inline T length(V3<T> v) { return m_sqrt(dot(v, v)); }
template <class T> inline T gmin(T x, T y) { return (y < x) ? y : x; }
template <class T> inline T gmax(T x, T y) { return (x < y) ? y : x; }
template <class T> inline V3<T> vmax0(V3<T> v) { return {gmax(T(0), v.x), gmax(T(0), v.y), gmax(T(0), v.z)}; }
template <class T> inline V3<T> ld(const float* p) { return {T(p[0]), T(p[1]), T(p[2])}; }
template <class T> inline bool isnan_(T x) { return x != x; }

// Tagging the two low mantissa bits of a box tmin (MBVHNode.cuh:91-94).
inline float tag_bits(float v, int slot) {
    uint32_t u;
    std::memcpy(&u, &v, 4);
    u = (u & 0xFFFFFFFCu) | (uint32_t)slot;
    std::memcpy(&v, &u, 4);
    return v;
}
inline double tag_bits(double v, int slot) {
    uint64_t u;
    std::memcpy(&u, &v, 8);
    u = (u & ~3ull) | (uint64_t)slot;
    std::memcpy(&v, &u, 8);
    return v;
}
inline int tag_of(float v) {
    uint32_t u;
    std::memcpy(&u, &v, 4);
    return (int)(u & 3u);
}
inline int tag_of(double v) {
    uint64_t u;
    std::memcpy(&u, &v, 8);
    return (int)(u & 3ull);
}

// ---------------------------------------------------------------- scene --
struct Box {
    float mn[3] = {1e34f, 1e34f, 1e34f};
    float mx[3] = {-1e34f, -1e34f, -1e34f};
};

}  // namespace

struct or_texture {
    uint32_t w = 0, h = 0;
    std::vector<float> rgba;  // row 0 at the bottom (Surface.cpp:33-100 via FreeImage)
};

struct or_scene {
    std::vector<float> vertices, normals, center_normals;
    std::vector<float> texcoords;  // 2 per vertex (TriangleList::m_TexCoords)
    std::vector<or_texture> textures;
    int sky_tex = -1;              // SceneData::skyboxTexture
    std::vector<uint32_t> indices, mat_idx, light_idx;
    std::vector<ct_material> materials;
    int sky_enabled = 0;
    float sky[3] = {0, 0, 0};
    // trees
    std::vector<ct_bvh_node> bvh;
    std::vector<ct_mbvh_node> mbvh;
    std::vector<uint32_t> prims;
    std::vector<Box> aabbs;
    bool built = false;
    uint32_t ntris() const { return (uint32_t)indices.size() / 3; }
    const float* vtx(uint32_t tri, int k) const { return &vertices[3 * (size_t)indices[3 * (size_t)tri + k]]; }
    const float* nrm(uint32_t tri, int k) const { return &normals[3 * (size_t)indices[3 * (size_t)tri + k]]; }
    const float* uv(uint32_t tri, int k) const { return &texcoords[2 * (size_t)indices[3 * (size_t)tri + k]]; }
};

namespace {

// ------------------------------------------------------------- builder --
// _mm_min_ps / _mm_max_ps lanes (AABB.h:311-314)
inline float smin(float a, float b) { return a < b ? a : b; }
inline float smax(float a, float b) { return a > b ? a : b; }
inline void grow(Box& b, const Box& a) {
    for (int k = 0; k < 3; ++k) { b.mn[k] = smin(b.mn[k], a.mn[k]); b.mx[k] = smax(b.mx[k], a.mx[k]); }
}
// AABB::Area (AABB.h:113-121)
inline float area(const Box& b) {
    const float e0 = b.mx[0] - b.mn[0], e1 = b.mx[1] - b.mn[1], e2 = b.mx[2] - b.mn[2];
    const float a = e0 * e1 + e0 * e2 + e1 * e2;
    return a > 0.0f ? a : 0.0f;
}
inline float centroid(const Box& b, int axis) { return (b.mn[axis] + b.mx[axis]) * 0.5f; }

struct Builder {
    or_scene& s;
    std::vector<ct_bvh_node>& pool;
    int poolPtr = 2;
    explicit Builder(or_scene& sc) : s(sc), pool(sc.bvh) {}

    // BVHNode::CalculateBounds, BVHNode.cu:25-38
    void calculateBounds(int n) {
        Box b;
        for (int i = 0; i < pool[n].count; ++i) grow(b, s.aabbs[s.prims[pool[n].left_first + i]]);
        for (int k = 0; k < 3; ++k) { pool[n].min[k] = b.mn[k]; pool[n].max[k] = b.mx[k]; }
    }
    // BVHNode::Partition (MT overload), BVHNode.cu:121-191: axis-outer,
    // planes i = 1..10 inner, full rescan per plane.
    bool partition(int n, int& left, int& right) {
        const ct_bvh_node node = pool[n];
        const int lFirst = node.left_first, count = node.count;
        Box nb;
        for (int k = 0; k < 3; ++k) { nb.mn[k] = node.min[k]; nb.mx[k] = node.max[k]; }
        const float parentNodeCost = area(nb) * (float)count;
        float lengths[3];
        for (int k = 0; k < 3; ++k) lengths[k] = node.max[k] - node.min[k];
        float lowest = 1e34f, bestCoord = 0.0f;
        int bestAxis = 0;
        for (int axis = 0; axis < 3; ++axis)
            for (int i = 1; i < 11; ++i) {
                const float binOffset = (float)i / (float)11;
                const float splitCoord = node.min[axis] + lengths[axis] * binOffset;
                int lc = 0, rc = 0;
                Box lb, rb;
                for (int idx = 0; idx < count; ++idx) {
                    const Box& a = s.aabbs[s.prims[lFirst + idx]];
                    if (centroid(a, axis) <= splitCoord) { grow(lb, a); lc++; }
                    else { grow(rb, a); rc++; }
                }
                const float cost = area(lb) * (float)lc + area(rb) * (float)rc;
                if (cost < lowest) { lowest = cost; bestCoord = splitCoord; bestAxis = axis; }
            }
        if (parentNodeCost < lowest) return false;
        int lCount = 0;
        for (int idx = 0; idx < count; ++idx) {
            const Box& a = s.aabbs[s.prims[lFirst + idx]];
            if (centroid(a, bestAxis) <= bestCoord) {
                std::swap(s.prims[lFirst + idx], s.prims[lFirst + lCount]);
                lCount++;
            }
        }
        left = poolPtr++;
        right = poolPtr++;
        pool[left].left_first = lFirst;
        pool[left].count = lCount;
        pool[right].left_first = lFirst + lCount;
        pool[right].count = count - lCount;
        return true;
    }
    // BVHNode::Subdivide(MT), BVHNode.cu:40-68/70-119: leaf if count < 4 or depth >= 64;
    // children allocated as a pair, left subtree first.
    void subdivide(int n, unsigned depth) {
        depth++;
        if (pool[n].count < 4 || depth >= 64u) return;
        int l, r;
        if (!partition(n, l, r)) return;
        pool[n].left_first = l;
        pool[n].count = -1;
        if (pool[l].count > 0) { calculateBounds(l); subdivide(l, depth); }
        if (pool[r].count > 0) { calculateBounds(r); subdivide(r, depth); }
    }
    // BVHTree::buildBVH, BVHTree.cpp:22-58
    void build() {
        const uint32_t N = s.ntris();
        pool.assign((size_t)N * 2 + 2, ct_bvh_node{{0, 0, 0}, -1, {0, 0, 0}, -1});
        poolPtr = 2;
        pool[0].left_first = 0;
        pool[0].count = (int)N;
        calculateBounds(0);
        subdivide(0, 1);
        if (poolPtr > 2) { pool[0].count = -1; pool[0].left_first = 2; }
        else { pool[0].count = (int)N; pool[0].left_first = 0; }
        pool.resize(poolPtr);
    }
};

struct Collapser {
    const std::vector<ct_bvh_node>& pool;
    std::vector<ct_mbvh_node>& tree;
    uint32_t finalPtr = 1;
    static void setBounds(ct_mbvh_node& m, int i, const float* mn, const float* mx) {
        m.bminx[i] = mn[0]; m.bminy[i] = mn[1]; m.bminz[i] = mn[2];
        m.bmaxx[i] = mx[0]; m.bmaxy[i] = mx[1]; m.bmaxz[i] = mx[2];
    }
    // MBVHNode::GetBVHNodeInfo, MBVHNode.cu:226-277
    int info(uint32_t mi, const ct_bvh_node& node) {
        ct_mbvh_node& m = tree[mi];
        for (int i = 0; i < 4; ++i) { m.childs[i] = -1; m.counts[i] = -1; }
        if (node.count > -1) return 0;
        int nc = 0;
        const ct_bvh_node& L = pool[node.left_first];
        const ct_bvh_node& R = pool[node.left_first + 1];
        if (L.count > -1) { const int i = nc++; setBounds(m, i, L.min, L.max); m.childs[i] = L.left_first; m.counts[i] = L.count; }
        else { const int a = nc++, b = nc++; m.childs[a] = L.left_first; m.childs[b] = L.left_first + 1; }
        if (R.count > -1) { const int i = nc++; setBounds(m, i, R.min, R.max); m.childs[i] = R.left_first; m.counts[i] = R.count; }
        else {
            const int a = nc++, b = nc++;
            setBounds(m, a, pool[R.left_first].min, pool[R.left_first].max);
            setBounds(m, b, pool[R.left_first + 1].min, pool[R.left_first + 1].max);
            m.childs[a] = R.left_first; m.childs[b] = R.left_first + 1;
        }
        return nc;
    }
    // MBVHNode::MergeNodes (vector overload), MBVHNode.cu:126-160
    void merge(uint32_t mi, const ct_bvh_node& node) {
        const int nc = info(mi, node);
        for (int i = 0; i < nc; ++i) {
            if (tree[mi].counts[i] != -1) continue;
            const ct_bvh_node& cur = pool[tree[mi].childs[i]];
            if (cur.count > -1) {
                tree[mi].counts[i] = cur.count;
                tree[mi].childs[i] = cur.left_first;
                setBounds(tree[mi], i, cur.min, cur.max);
            } else {
                const uint32_t ni = finalPtr++;
                tree[mi].childs[i] = (int)ni;
                tree[mi].counts[i] = -1;
                setBounds(tree[mi], i, cur.min, cur.max);
                merge(ni, cur);
            }
        }
        const float imn[3] = {1e34f, 1e34f, 1e34f}, imx[3] = {-1e34f, -1e34f, -1e34f};
        for (int i = nc; i < 4; ++i) { setBounds(tree[mi], i, imn, imx); tree[mi].counts[i] = 0; }
    }
    void build(uint32_t nprims) {
        tree.assign((size_t)nprims * 2 + 1, ct_mbvh_node());
        finalPtr = 1;
        const ct_bvh_node& root = pool[0];
        if (root.count > -1) {
            // Q17: keep a leaf root visible (the reference empties it).
            const float imn[3] = {1e34f, 1e34f, 1e34f}, imx[3] = {-1e34f, -1e34f, -1e34f};
            for (int i = 0; i < 4; ++i) { setBounds(tree[0], i, imn, imx); tree[0].childs[i] = -1; tree[0].counts[i] = 0; }
            setBounds(tree[0], 0, root.min, root.max);
            tree[0].childs[0] = root.left_first;
            tree[0].counts[0] = root.count;
        } else {
            merge(0, root);
        }
        tree.resize(finalPtr);
    }
};

// ------------------------------------------------------------ geometry --
// triangle::intersect, Triangle.cuh:13-44 (Moller-Trumbore, strict t < *rayt)
template <class T>
inline bool tri_intersect(V3<T> org, V3<T> dir, T* rayt, V3<T> p0, V3<T> p1, V3<T> p2, T eps) {
    const V3<T> edge1 = p1 - p0;
    const V3<T> edge2 = p2 - p0;
    const V3<T> h = cross(dir, edge2);
    const T a = dot(edge1, h);
    if (a > -eps && a < eps) return false;
    const T f = T(1) / a;
    const V3<T> s = org - p0;
    const T u = f * dot(s, h);
    if (u < T(0) || u > T(1)) return false;
    const V3<T> q = cross(s, edge1);
    const T v = f * dot(dir, q);
    if (v < T(0) || u + v > T(1)) return false;
    const T t = f * dot(edge2, q);
    if (t > eps && *rayt > t) { *rayt = t; return true; }
    return false;
}

struct Counters {
    uint32_t nodes = 0, tris = 0;
};

template <class T>
inline bool test_prim(const or_scene& s, uint32_t prim, V3<T> o, V3<T> d, T* t, T eps, Counters& c) {
    c.tris++;
    return tri_intersect(o, d, t, ld<T>(s.vtx(prim, 0)), ld<T>(s.vtx(prim, 1)), ld<T>(s.vtx(prim, 2)), eps);
}

// BVHNode::Intersect, BVHNode.cuh:50-62 (no t-culling)
template <class T>
inline bool bvh_box(const ct_bvh_node& n, V3<T> o, V3<T> inv, T* tmin, T* tmax) {
    const V3<T> t1 = (ld<T>(n.min) - o) * inv;
    const V3<T> t2 = (ld<T>(n.max) - o) * inv;
    const V3<T> mn{gmin(t1.x, t2.x), gmin(t1.y, t2.y), gmin(t1.z, t2.z)};
    const V3<T> mx{gmax(t1.x, t2.x), gmax(t1.y, t2.y), gmax(t1.z, t2.z)};
    *tmin = gmax(mn.x, gmax(mn.y, mn.z));
    *tmax = gmin(mx.x, gmin(mx.y, mx.z));
    return *tmax >= T(0) && *tmin < *tmax;
}

// BVHNode::traverseBVH / traverseBVHShadow, BVHNode.cuh:82-204.  The pushes
// put the strictly nearer child first, so the farther child is popped first
// (kept as written).  Stack unbounded here (the reference's 32 is not checked).
template <class T, bool AnyHit>
bool traverse_bvh2(const or_scene& s, V3<T> o, V3<T> d, T* t, int* hit, T eps, Counters& c) {
    const V3<T> inv{T(1) / d.x, T(1) / d.y, T(1) / d.z};
    int stack[256];
    int sp = 0;
    stack[0] = 0;
    while (sp >= 0) {
        const ct_bvh_node& node = s.bvh[stack[sp]];
        sp--;
        if (node.count > -1) {
            for (int i = 0; i < node.count; ++i) {
                const uint32_t prim = s.prims[node.left_first + i];
                if (test_prim(s, prim, o, d, t, eps, c)) {
                    if (AnyHit) return true;
                    *hit = (int)prim;
                }
            }
        } else {
            T n1, f1, n2, f2;
            c.nodes += 2;
            const bool hl = bvh_box(s.bvh[node.left_first], o, inv, &n1, &f1);
            const bool hr = bvh_box(s.bvh[node.left_first + 1], o, inv, &n2, &f2);
            if (hl && hr) {
                if (n1 < n2) { stack[++sp] = node.left_first; stack[++sp] = node.left_first + 1; }
                else { stack[++sp] = node.left_first + 1; stack[++sp] = node.left_first; }
            } else if (hl) {
                stack[++sp] = node.left_first;
            } else if (hr) {
                stack[++sp] = node.left_first + 1;
            }
        }
    }
    return false;
}

// MBVHNode::intersect, MBVHNode.cuh:69-110: 4-wide slab test, tmin tagged
// with the slot id, hit on the tagged value, 5-comparator network.
template <class T>
inline void mbvh_intersect(const ct_mbvh_node& n, V3<T> o, V3<T> inv, T t, T tmin[4], bool res[4]) {
    T tmax[4];
    for (int i = 0; i < 4; ++i) {
        T t1 = (T(n.bminx[i]) - o.x) * inv.x, t2 = (T(n.bmaxx[i]) - o.x) * inv.x;
        T mn = gmin(t1, t2), mx = gmax(t1, t2);
        t1 = (T(n.bminy[i]) - o.y) * inv.y;
        t2 = (T(n.bmaxy[i]) - o.y) * inv.y;
        mn = gmax(mn, gmin(t1, t2));
        mx = gmin(mx, gmax(t1, t2));
        t1 = (T(n.bminz[i]) - o.z) * inv.z;
        t2 = (T(n.bmaxz[i]) - o.z) * inv.z;
        mn = gmax(mn, gmin(t1, t2));
        mx = gmin(mx, gmax(t1, t2));
        tmin[i] = tag_bits(mn, i);
        tmax[i] = mx;
    }
    for (int i = 0; i < 4; ++i) res[i] = tmax[i] > T(0) && tmin[i] <= tmax[i] && tmin[i] < t;
    if (tmin[0] > tmin[1]) std::swap(tmin[0], tmin[1]);
    if (tmin[2] > tmin[3]) std::swap(tmin[2], tmin[3]);
    if (tmin[0] > tmin[2]) std::swap(tmin[0], tmin[2]);
    if (tmin[1] > tmin[3]) std::swap(tmin[1], tmin[3]);
    if (tmin[2] > tmin[3]) std::swap(tmin[2], tmin[3]);
}

// MBVHNode::traverseMBVH / traverseMBVHShadow, MBVHNode.cuh:126-205.
template <class T, bool AnyHit>
bool traverse_bvh4(const or_scene& s, V3<T> o, V3<T> d, T* t, int* hit, T eps, Counters& c) {
    struct Entry { int leftFirst, count; };
    Entry stack[512];
    int sp = 0;
    stack[0] = {0, -1};
    const V3<T> inv{T(1) / d.x, T(1) / d.y, T(1) / d.z};
    while (sp >= 0) {
        const Entry e = stack[sp--];
        if (e.count > -1) {
            for (int i = 0; i < e.count; ++i) {
                const uint32_t prim = s.prims[e.leftFirst + i];
                if (test_prim(s, prim, o, d, t, eps, c)) {
                    if (AnyHit) return true;
                    *hit = (int)prim;
                }
            }
            continue;
        }
        const ct_mbvh_node& n = s.mbvh[e.leftFirst];
        T tmin[4];
        bool res[4];
        c.nodes++;
        mbvh_intersect(n, o, inv, *t, tmin, res);
        for (int i = 3; i >= 0; --i) {
            const int idx = tag_of(tmin[i]);
            if (res[idx]) stack[++sp] = {n.childs[idx], n.counts[idx]};
        }
    }
    return false;
}

template <class T, bool AnyHit>
bool traverse_brute(const or_scene& s, V3<T> o, V3<T> d, T* t, int* hit, T eps, Counters& c) {
    const uint32_t n = s.ntris();
    for (uint32_t p = 0; p < n; ++p)
        if (test_prim(s, p, o, d, t, eps, c)) {
            if (AnyHit) return true;
            *hit = (int)p;
        }
    return false;
}

template <class T, bool AnyHit>
bool trace(const or_scene& s, int accel, V3<T> o, V3<T> d, T* t, int* hit, T eps, Counters& c) {
    if (accel == 0) return traverse_bvh2<T, AnyHit>(s, o, d, t, hit, eps, c);
    if (accel == 1) return traverse_bvh4<T, AnyHit>(s, o, d, t, hit, eps, c);
    return traverse_brute<T, AnyHit>(s, o, d, t, hit, eps, c);
}

// ------------------------------------------------------------ textures --
// SceneData::getTextureColor (SceneData.cuh:81-97): nearest texel after
// wrapping each coordinate into [0, 1) (ct_texel_index, shared with the
// kernels).  getTextureNormal (SceneData.cuh:99-114) skips the fmod and
// reads outside the texture above 1; the decision here wraps it the same way.
inline const float* texel(const or_scene& s, int t, float u, float v) {
    const or_texture& x = s.textures[(size_t)t];
    return &x.rgba[4 * (size_t)ct_texel_index(x.w, x.h, u, v)];
}
// convertToLocalSpace + sampleToWorld (Random.cuh:37-53)
template <class T> V3<T> sample_to_world(V3<T> smp, V3<T> N) {
    const V3<T> W = m_abs(N.x) > T(0.99f) ? V3<T>{T(0), T(1), T(0)} : V3<T>{T(1), T(0), T(0)};
    const V3<T> Tv = normalize(cross(N, W));
    const V3<T> Bv = cross(N, Tv);
    return smp.x * Tv + smp.y * Bv + smp.z * N;
}
// Sky texture coordinates (Kernel.cu:96-100): the shared float functions, or
// double libm for the FP64 render.
inline void sky_uv(float dx, float dy, float dz, float* u, float* v) { ct_sky_uv(dx, dy, dz, u, v); }
inline void sky_uv(double dx, double dy, double dz, float* u, float* v) {
    const double pi = 3.14159265358979323846;
    *u = (float)(1.0 + std::atan2(dx, -dz) / pi * 0.5);
    *v = (float)(1.0 - std::acos(dy < -1.0 ? -1.0 : dy > 1.0 ? 1.0 : dy) / pi);
}

// triangle helpers, Triangle.cuh:46-96
template <class T> V3<T> bary_coords(V3<T> p, V3<T> n, V3<T> p0, V3<T> p1, V3<T> p2) {
    const T abc = dot(n, cross(p1 - p0, p2 - p0));
    const T pbc = dot(n, cross(p1 - p, p2 - p));
    const T pca = dot(n, cross(p2 - p, p0 - p));
    const T alpha = pbc / abc;
    const T beta = pca / abc;
    const T gamma = T(1) - alpha - beta;
    return {alpha, beta, gamma};
}
template <class T> T tri_area(V3<T> p0, V3<T> p1, V3<T> p2) {
    const T a = length(p0 - p1), b = length(p1 - p2), c = length(p2 - p0);
    const T s = (a + b + c) / T(2);
    return m_sqrt(s * (s - a) * (s - b) * (s - c));
}
template <class T> V3<T> random_point(V3<T> p0, V3<T> p1, V3<T> p2, T r1, T r2) {
    if (r1 + r2 > T(1)) { r1 = T(1) - r1; r2 = T(1) - r2; }
    const T a = T(1) - r1 - r2;
    return a * p0 + r1 * p1 + r2 * p2;
}

// Tangent2World, Tools.h:110-137 (the N.z < 0 branch form); T and B also
// give the local frame of the microfacet lobes.
template <class T> void tangent_frame(V3<T> N, V3<T>& Tv, V3<T>& Bv) {
    if (N.z < T(0)) {
        const T a = T(1) / (T(1) - N.z);
        const T b = N.x * N.y * a;
        Bv = {T(1) - N.x * N.x * a, -b, N.x};
        Tv = {b, N.y * N.y * a - T(1), -N.y};
    } else {
        const T a = T(1) / (T(1) + N.z);
        const T b = -N.x * N.y * a;
        Bv = {T(1) - N.x * N.x * a, b, -N.x};
        Tv = {b, T(1) - N.y * N.y * a, -N.y};
    }
}
template <class T> V3<T> tangent_to_world(V3<T> V, V3<T> N) {
    V3<T> Tv, Bv;
    tangent_frame(N, Tv, Bv);
    return V.x * Tv + V.y * Bv + V.z * N;
}
// DiffuseReflectionCosWeighted, Tools.h:162-169 (the sqrt(1.0 - r1) is in double)
template <class T> V3<T> cos_weighted(T r0, T r1) {
    const T term1 = K<T>::two_pi * r0;
    const T term2 = (T)std::sqrt(1.0 - (double)r1);
    const T s = m_sin(term1);
    const T c = m_cos(term1);
    return {c * term2, s * term2, m_sqrt(r1)};
}
// glm::reflect(I, N) = I - N * dot(N, I) * 2
template <class T> V3<T> reflect(V3<T> I, V3<T> N) { return I - N * dot(N, I) * T(2); }

// ---------------------------------------------------------- microfacet --
// Microfacet.cuh:84-127 trigonometric helpers (local frame, z = normal)
template <class T> T cos2t(V3<T> w) { return w.z * w.z; }
template <class T> T sin2t(V3<T> w) { return gmax(T(0), T(1) - cos2t(w)); }
template <class T> T sint(V3<T> w) { return m_sqrt(sin2t(w)); }
template <class T> T tant(V3<T> w) { return sint(w) / w.z; }
template <class T> T tan2t(V3<T> w) { return sin2t(w) / cos2t(w); }
template <class T> T clamp11(T v) { return v < T(-1) ? T(-1) : (v > T(1) ? T(1) : v); }
template <class T> T cosphi(V3<T> w) { const T st = sint(w); return (st == T(0)) ? T(1) : clamp11(w.x / st); }
template <class T> T sinphi(V3<T> w) { const T st = sint(w); return (st == T(0)) ? T(0) : clamp11(w.y / st); }
template <class T> bool isinf_(T v) { return v == v && (v - v) != (v - v); }

// Beckmann D, Microfacet.cuh:129-139 (alphax = alphay = alpha)
template <class T> T D_beckmann(V3<T> wh, T alpha) {
    const T tan2 = tan2t(wh);
    if ((T(2) * tan2) == tan2) return T(0);
    const T cos4 = cos2t(wh) * cos2t(wh);
    const T cp = cosphi(wh), sp = sinphi(wh);
    return m_exp(-tan2 * (cp * cp / (alpha * alpha) + sp * sp / (alpha * alpha))) / (K<T>::pi * alpha * alpha * cos4);
}
// GGX (Trowbridge-Reitz) D, PBRT-v3 microfacet.cpp TrowbridgeReitzDistribution::D
// (the reference has no GGX D): 1 / (pi a^2 cos^4 (1 + tan^2 / a^2)^2).
template <class T> T D_ggx(V3<T> wh, T alpha) {
    const T tan2 = tan2t(wh);
    if ((T(2) * tan2) == tan2) return T(0);
    const T cos4 = cos2t(wh) * cos2t(wh);
    const T e = tan2 / (alpha * alpha);
    return T(1) / (K<T>::pi * alpha * alpha * cos4 * (T(1) + e) * (T(1) + e));
}
// lambda_ggx, Microfacet.cuh:417-426
template <class T> T lambda_ggx(V3<T> w, T alpha) {
    const T absTan = m_abs(tant(w));
    if (T(2) * absTan == absTan) return T(0);
    const T cp = cosphi(w), sp = sinphi(w);
    const T a = m_sqrt(cp * cp * alpha * alpha + sp * sp * alpha * alpha);
    const T a2t2 = a * absTan * a * absTan;
    return (T(-1) + m_sqrt(T(1) + a2t2)) / T(2);
}
// lambda_beckmann, Microfacet.cuh:500-513
template <class T> T lambda_beckmann(V3<T> w, T alpha) {
    const T absTan = m_abs(tant(w));
    if ((T(2) * absTan) == absTan) return T(0);
    const T cp = cosphi(w), sp = sinphi(w);
    const T al = m_sqrt(cp * cp * alpha * alpha + sp * sp * alpha * alpha);
    const T a = T(1) / (al * absTan);
    if (a >= T(1.6)) return T(0);
    return (T(1) - T(1.259) * a + T(0.396) * a * a) / (T(3.535) * a + T(2.181) * a * a);
}
// sampleGGX_P22_11, Microfacet.cuh:312-351
template <class T> void ggx_p22_11(T cosThetaI, T* sx, T* sy, T r1, T r2) {
    if (cosThetaI > T(0.9999)) {
        const T r = m_sqrt(r1 / (T(1) - r1));
        const T sinPhi = m_sin(K<T>::two_pi * r2);
        const T cosPhi = m_cos(K<T>::two_pi * r2);
        *sx = r * cosPhi;
        *sy = r * sinPhi;
        return;
    }
    const T sinThetaI = m_sqrt(gmax(T(0), T(1) - cosThetaI * cosThetaI));
    const T tanThetaI = sinThetaI / cosThetaI;
    const T a = T(1) / tanThetaI;
    const T G1 = T(2) / (T(1) + m_sqrt(T(1) + T(1) / (a * a)));
    const T A = T(2) * r1 / G1 - T(1);
    const T B = tanThetaI;
    const T tmp = gmin(T(1) / (A * A - T(1)), T(1.0e12));
    const T D = m_sqrt(B * B * tmp * tmp - (A * A - B * B) * tmp);
    const T sx1 = B * tmp - D;
    const T sx2 = B * tmp + D;
    *sx = (A < T(0) || sx2 > T(1) / tanThetaI) ? sx1 : sx2;
    T S;
    if (r2 > T(0.5)) { S = T(1); r2 = T(2) * (r2 - T(0.5)); }
    else { S = T(-1); r2 = T(2) * (T(0.5) - r2); }
    const T z = (r2 * (r2 * (r2 * T(0.27385) - T(0.73369)) + T(0.46341))) /
                (r2 * (r2 * (r2 * T(0.093073) + T(0.309420)) - T(1.000000)) + T(0.597999));
    *sy = S * z * m_sqrt(T(1) + (*sx) * (*sx));
}
// BeckmannSample11, Microfacet.cuh:141-209
template <class T> void beckmann_11(T cosThetaI, T r1, T r2, T* sx, T* sy) {
    if (cosThetaI > T(0.9999)) {
        const T r = m_sqrt(-m_log(T(1) - r1));
        const T sinPhi = m_sin(T(2) * K<T>::pi * r2);
        const T cosPhi = m_cos(T(2) * K<T>::pi * r2);
        *sx = r * cosPhi;
        *sy = r * sinPhi;
        return;
    }
    const T sinThetaI = m_sqrt(gmax(T(0), T(1) - cosThetaI * cosThetaI));
    const T tanThetaI = sinThetaI / cosThetaI;
    const T cotThetaI = T(1) / tanThetaI;
    T a = T(-1);
    T c = m_erf(cotThetaI);
    const T sample_x = gmax(r1, T(1e-6));
    const T thetaI = m_acos(cosThetaI);
    const T fit = T(1) + thetaI * (T(-0.876) + thetaI * (T(0.4265) - T(0.0594) * thetaI));
    T b = c - (T(1) + c) * m_pow(T(1) - sample_x, fit);
    const T normalization =
        T(1) / (T(1) + c + T(1) / K<T>::root_pi * tanThetaI * m_exp(-cotThetaI * cotThetaI));
    int it = 0;
    while (++it < 10) {
        if (!(b >= a && b <= c)) b = T(0.5) * (a + c);
        const T invErf = m_erfinv(b);
        const T value =
            normalization * (T(1) + b + T(1) / m_sqrt(K<T>::pi) * tanThetaI * m_exp(-invErf * invErf)) - sample_x;
        const T derivative = normalization * (T(1) - invErf * tanThetaI);
        if (m_abs(value) < T(1e-5)) break;
        if (value > T(0)) c = b;
        else a = b;
        b -= value / derivative;
    }
    *sx = m_erfinv(b);
    *sy = m_erfinv(T(2) * gmax(r2, T(1e-6)) - T(1));
}
// sampleGGX / BeckmannSample (stretch, sample, rotate, unstretch),
// Microfacet.cuh:211-230 and :353-373, with the wo.z < 0 flips of
// sample_ggx (:376-415) and sample_beckmann (:441-486).
template <class T> V3<T> sample_vndf(int type, V3<T> wo, T alpha, T r1, T r2) {
    const bool flip = wo.z < T(0);
    const V3<T> wi = flip ? -wo : wo;
    const V3<T> ws = normalize(V3<T>{alpha * wi.x, alpha * wi.y, wi.z});
    T sx, sy;
    if (type == CT_MAT_BECKMANN || type == CT_MAT_FRESNEL_BECKMANN) beckmann_11(ws.z, r1, r2, &sx, &sy);
    else ggx_p22_11(ws.z, &sx, &sy, r1, r2);
    const T cp = cosphi(ws), sp = sinphi(ws);
    const T tmp = cp * sx - sp * sy;
    sy = sp * sx + cp * sy;
    sx = tmp;
    sx = alpha * sx;
    sy = alpha * sy;
    V3<T> wm = normalize(V3<T>{-sx, -sy, T(1)});
    if (type == CT_MAT_BECKMANN || type == CT_MAT_FRESNEL_BECKMANN) {
        if (flip) wm = -wm;
    } else if (wm.z < T(0)) {
        wm = -wm;
    }
    return wm;
}

inline bool is_beckmann(int type) { return type == CT_MAT_BECKMANN || type == CT_MAT_FRESNEL_BECKMANN; }
inline bool is_microfacet(int type) {
    return type == CT_MAT_BECKMANN || type == CT_MAT_GGX || type == CT_MAT_TROWBRIDGE;
}
inline bool is_dielectric(int type) {
    return type == CT_MAT_FRESNEL || type == CT_MAT_FRESNEL_BECKMANN || type == CT_MAT_FRESNEL_GGX ||
           type == CT_MAT_FRESNEL_TROWBRIDGE;
}

// Microfacet reflection in the local frame (z = shading normal):
// f = F D G2 / (4 cos_o cos_i), pdf = D G1(wo) / (4 cos_o) for visible-normal
// sampling (Heitz 2014 / PBRT-v3 §8.4), G2 height-correlated 1/(1+L(o)+L(i)),
// F = Schlick with F0 = albedo.
template <class T>
void mf_eval(int type, T alpha, V3<T> albedo, V3<T> wo, V3<T> wi, V3<T>* f, T* pdf) {
    if (wo.z <= T(0) || wi.z <= T(0)) { *f = {T(0), T(0), T(0)}; *pdf = T(0); return; }
    V3<T> wm = normalize(wo + wi);
    const T D = is_beckmann(type) ? D_beckmann(wm, alpha) : D_ggx(wm, alpha);
    const T Lo = is_beckmann(type) ? lambda_beckmann(wo, alpha) : lambda_ggx(wo, alpha);
    const T Li = is_beckmann(type) ? lambda_beckmann(wi, alpha) : lambda_ggx(wi, alpha);
    const T G1o = T(1) / (T(1) + Lo);
    const T G2 = T(1) / (T(1) + Lo + Li);
    const T c = gmax(T(0), dot(wo, wm));
    const T m = T(1) - c;
    const T m5 = m * m * m * m * m;
    const V3<T> F = albedo + (V3<T>{T(1), T(1), T(1)} - albedo) * m5;
    *pdf = D * G1o / (T(4) * wo.z);
    *f = F * (D * G2 / (T(4) * wo.z * wi.z));
}

// Dielectric Fresnel, relative IOR eta = n_t / n_i, after
// FresnelDielectricExact (Tools.h:96-108), which returns 0 for eta <= 1;
// here eta < 1 (leaving the medium) is handled, with total internal
// reflection returning 1.
template <class T> T fresnel_dielectric(T cosThetaI, T eta) {
    const T scale = T(1) / eta;
    const T cosThetaTSqr = T(1) - (T(1) - cosThetaI * cosThetaI) * (scale * scale);
    if (cosThetaTSqr <= T(0)) return T(1);
    const T cosThetaT = m_sqrt(cosThetaTSqr);
    const T Rs = (cosThetaI - eta * cosThetaT) / (cosThetaI + eta * cosThetaT);
    const T Rp = (eta * cosThetaI - cosThetaT) / (eta * cosThetaI + cosThetaT);
    return T(0.5) * (Rs * Rs + Rp * Rp);
}

// glm::refract(I, N, eta), include/glm/detail/func_geometric.inl:112-122:
// k = 1 - eta^2 (1 - (N.I)^2); the direction eta I - (eta N.I + sqrt(k)) N,
// valid (no total internal reflection) when k >= 0.
template <class T> bool refract_dir(V3<T> I, V3<T> N, T eta, V3<T>* out) {
    const T dv = dot(N, I);
    const T k = T(1) - eta * eta * (T(1) - dv * dv);
    *out = eta * I - (eta * dv + m_sqrt(k)) * N;
    return k >= T(0);
}

// --------------------------------------------------------------- shading --
enum { SPECULAR = 0, DIFFUSE = 1, GLOSSY = 2 };

// Kernel.cu:128,184 multiply the light pdf by lightCount; uniform light
// selection divides by it (CT_LIGHTPDF_UNIFORM).
template <class T> T light_scale(T pdf, uint32_t n, int mode) {
    return mode == CT_LIGHTPDF_UNIFORM ? pdf / T((float)n) : pdf * T((float)n);
}

template <class T> struct Camera3 {
    V3<T> pos, w, hor, ver;
    T invw, invh;
};
// sample(): camera basis, Kernel.cu:274-290
template <class T> Camera3<T> camera_basis(const ct_camera& c, int width, int height) {
    Camera3<T> cb;
    cb.pos = ld<T>(c.position);
    const V3<T> w = ld<T>(c.forward), up = ld<T>(c.up);
    const V3<T> u = normalize(cross(w, up));
    const V3<T> v = normalize(cross(u, w));
    const T dist = T(c.plane_distance);
    cb.w = w;
    cb.hor = u * dist;
    cb.ver = v * dist;
    const T aspect = T((float)width) / T((float)height);
    if (width > height) cb.hor = cb.hor * aspect;
    else cb.ver = cb.ver * aspect;
    cb.invw = T(1) / T((float)width);
    cb.invh = T(1) / T((float)height);
    return cb;
}

struct Stats {
    uint64_t primary = 0, extension = 0, shadow = 0, samples = 0;
    uint64_t ncl = 0, tcl = 0, nsh = 0, tsh = 0;
};

// One path sample: generatePrimaryRays (Kernel.cu:208-234, Ray.cuh:85-95),
// then per segment extend (Kernel.cu:236-245), shade (Kernel.cu:88-206) and
// connect (Kernel.cu:247-256).  The radiance of the path is summed in path
// order (emission / sky at a segment's hit, then that vertex's NEE).
// Decisions: barycentrics at the hit point (Q8); NEE emission of
// lightIndices[light] (Q1); a rejected BSDF sample ends the path without NEE
// (Q9); start_position = 0 so thread == pixel (Q7); every queued ray is
// processed (Q11); uint32 seed arithmetic (Q16).
template <class T>
V3<T> path_sample(const or_scene& s, const Camera3<T>& cb, const ct_frame_opts& o, int x, int y, int width,
                  uint32_t frame, Stats& st) {
    const uint32_t pixel = (uint32_t)(x + y * width);
    uint32_t seed = ct_primary_seed(pixel, frame);
    const float jx = ct_rand_float(&seed);
    const float jy = ct_rand_float(&seed);
    const T px = T((float)x) + T(jx) - T(0.5f);
    const T py = T((float)y) + T(jy) - T(0.5f);
    const T sx = T(2) * (px * cb.invw) - T(1);
    const T sy = T(1) - T(2) * (py * cb.invh);
    V3<T> dir = normalize(cb.w + cb.hor * sx + cb.ver * sy);
    V3<T> org = cb.pos;
    V3<T> thr{T(1), T(1), T(1)};
    T lastPdf = T(1);
    int lastType = SPECULAR;
    V3<T> L{T(0), T(0), T(0)};
    const T teps = T(o.triangle_epsilon);
    const T neps = T(o.normal_epsilon);
    const T deps = T(o.dist_epsilon);
    const uint32_t lightCount = (uint32_t)s.light_idx.size();
    const int accel = o.accel;
    for (int bounce = 0; bounce < o.max_depth; ++bounce) {
        if (bounce == 0) st.primary++;
        else st.extension++;
        T t = T(1e34f);
        int prim = -1;
        Counters cc;
        trace<T, false>(s, accel, org, dir, &t, &prim, teps, cc);
        st.ncl += cc.nodes;
        st.tcl += cc.tris;
        if (!(t < T(1e34f))) {  // miss (Ray::valid, Ray.cuh:38-41)
            if (s.sky_enabled) {
                V3<T> sky{T(s.sky[0]), T(s.sky[1]), T(s.sky[2])};
                if (s.sky_tex >= 0) {  // Kernel.cu:96-100
                    float u, v;
                    sky_uv(dir.x, dir.y, dir.z, &u, &v);
                    sky = ld<T>(texel(s, s.sky_tex, u, v));
                }
                L = L + (thr / lastPdf) * sky;
            }
            break;
        }
        const ct_material& mat = s.materials[s.mat_idx[prim]];
        const V3<T> p0 = ld<T>(s.vtx(prim, 0)), p1 = ld<T>(s.vtx(prim, 1)), p2 = ld<T>(s.vtx(prim, 2));
        const V3<T> hitp = org + t * dir;
        V3<T> N = ld<T>(&s.center_normals[3 * (size_t)prim]);
        const V3<T> bary = bary_coords(hitp, N, p0, p1, p2);
        V3<T> iN = bary.x * ld<T>(s.nrm(prim, 0)) + bary.y * ld<T>(s.nrm(prim, 1)) + bary.z * ld<T>(s.nrm(prim, 2));
        // textures (Kernel.cu:106-112,146): coordinates at the hit point (Q8)
        float tu = 0.0f, tv = 0.0f;
        if (mat.diffuse_tex >= 0 || mat.normal_tex >= 0) {
            const float* t0 = s.uv(prim, 0);
            const float* t1 = s.uv(prim, 1);
            const float* t2 = s.uv(prim, 2);
            tu = (float)((bary.x * T(t0[0]) + bary.y * T(t1[0])) + bary.z * T(t2[0]));
            tv = (float)((bary.x * T(t0[1]) + bary.y * T(t1[1])) + bary.z * T(t2[1]));
        }
        if (mat.normal_tex >= 0) iN = sample_to_world(normalize(ld<T>(texel(s, mat.normal_tex, tu, tv))), iN);
        if (mat.type == CT_MAT_LIGHT) {
            if (!o.indirect) break;
            const T DdotNL = -dot(dir, iN);
            const V3<T> Le = ld<T>(mat.albedo);
            V3<T> c;
            if (lastType == SPECULAR) {
                c = thr * Le;
            } else {
                const T lightPdf = light_scale(t * t / (DdotNL * tri_area(p0, p1, p2)), lightCount, o.light_pdf);
                c = thr * Le * (T(1) / (lastPdf + lightPdf));
            }
            L = L + vmax0(c);
            break;
        }
        org = hitp;
        thr = vmax0(thr / lastPdf);
        const bool flipped = dot(dir, N) > T(0);
        if (flipped) { N = -N; iN = -iN; }
        seed = ct_shade_seed(pixel, frame, ct_path_length_key(bounce));
        const T r3 = T(ct_rand_float(&seed));
        const T r4 = T(ct_rand_float(&seed));
        const V3<T> color = mat.diffuse_tex >= 0 ? ld<T>(texel(s, mat.diffuse_tex, tu, tv)) : ld<T>(mat.albedo);
        const V3<T> wiw = -dir;  // direction towards the previous vertex
        // --- BSDF sample (BSDF.h:33-63, extended for the typed materials)
        V3<T> wo{T(0), T(0), T(0)}, f{T(0), T(0), T(0)};
        T pdf = T(0);
        int type = DIFFUSE;
        bool valid = false;
        bool transmit = false;
        const bool typed = o.bsdf_mode == CT_BSDF_TYPED;
        const int mtype = mat.type;
        const T alpha = gmax(T(mat.roughness), T(1e-3f));
        // `type` is the material's lobe, set before sampling: next-event
        // estimation depends on the lobe, not on whether the sample is kept
        if ((!typed && mat.roughness < 0.1f) || (typed && mtype == CT_MAT_SPECULAR)) {
            type = SPECULAR;
            wo = reflect(dir, iN);
            if (dot(N, wo) > T(0)) {
                pdf = T(1);
                f = color * (T(1) / m_abs(dot(iN, wo)));
                valid = true;
            }
        } else if (typed && is_dielectric(mtype)) {
            const T etaI = flipped ? T(mat.refract_idx) : T(1);
            const T etaT = flipped ? T(1) : T(mat.refract_idx);
            const T cosI = dot(wiw, iN);
            const T F = fresnel_dielectric(cosI, etaT / etaI);
            type = SPECULAR;
            pdf = T(1);
            if (r3 < F || F >= T(1)) {
                wo = reflect(dir, iN);
                valid = dot(N, wo) > T(0);
            } else {
                valid = refract_dir(dir, iN, etaI / etaT, &wo) && dot(N, wo) < T(0);
                transmit = true;
            }
            if (valid) f = color * (T(1) / m_abs(dot(iN, wo)));
        } else if (typed && is_microfacet(mtype)) {
            type = GLOSSY;
            V3<T> Tv, Bv;
            tangent_frame(iN, Tv, Bv);
            const V3<T> wol{dot(wiw, Tv), dot(wiw, Bv), dot(wiw, iN)};
            if (wol.z > T(0)) {
                const V3<T> wm = sample_vndf(mtype, wol, alpha, r3, r4);
                const V3<T> wil = reflect(-wol, wm);
                if (wil.z > T(0)) {
                    wo = normalize(wil.x * Tv + wil.y * Bv + wil.z * iN);
                    if (dot(N, wo) > T(0)) {
                        mf_eval(mtype, alpha, color, wol, wil, &f, &pdf);
                        valid = pdf > T(0);
                    }
                }
            }
        } else {
            type = DIFFUSE;
            wo = normalize(tangent_to_world(cos_weighted(r3, r4), iN));
            if (dot(N, wo) > T(0)) {
                pdf = gmax(T(0), dot(wo, iN)) * K<T>::inv_pi;
                f = color * K<T>::inv_pi;
                valid = true;
            }
        }
        // --- next-event estimation (Kernel.cu:161-195), for a diffuse or
        // glossy lobe also when the BSDF sample is rejected below (Q9): with
        // NEE skipped for rejected samples the MIS sum loses the light
        // sample's share in proportion to the rejection rate (a bias)
        if (type != SPECULAR && o.shadow && lightCount > 0) {
            int light = ct_rand_int_max(&seed, (int)lightCount - 1);
            // RandomIntMax can reach lightCount once float(max)+0.99999f rounds up
            // (more than 2^22 lights); clamp like the kernels.
            if (light > (int)lightCount - 1) light = (int)lightCount - 1;
            const T r1 = T(ct_rand_float(&seed));
            const T r2 = T(ct_rand_float(&seed));
            const uint32_t lt = s.light_idx[light];
            const V3<T> l0 = ld<T>(s.vtx(lt, 0)), l1 = ld<T>(s.vtx(lt, 1)), l2 = ld<T>(s.vtx(lt, 2));
            const V3<T> lightPos = random_point(l0, l1, l2, r1, r2);
            V3<T> Lv = lightPos - org;
            const T sqDist = dot(Lv, Lv);
            const T distance = m_sqrt(sqDist);
            Lv = Lv / distance;
            const V3<T> lN = ld<T>(&s.center_normals[3 * (size_t)lt]);
            const T NdotL = dot(iN, Lv);
            const T LNdotL = -dot(lN, Lv);
            if (NdotL > T(0) && LNdotL > T(0)) {
                const T ar = tri_area(l0, l1, l2);
                const V3<T> emission = ld<T>(s.materials[s.mat_idx[lt]].albedo);
                V3<T> fe;
                T shadowPdf;
                if (type == DIFFUSE) {
                    // EvaluateBSDF, BSDF.h:26-31
                    shadowPdf = m_abs(dot(Lv, iN)) * K<T>::inv_pi;
                    fe = color * K<T>::inv_pi;
                } else {
                    V3<T> Tv, Bv;
                    tangent_frame(iN, Tv, Bv);
                    const V3<T> wol{dot(wiw, Tv), dot(wiw, Bv), dot(wiw, iN)};
                    const V3<T> wil{dot(Lv, Tv), dot(Lv, Bv), dot(Lv, iN)};
                    mf_eval(mtype, alpha, color, wol, wil, &fe, &shadowPdf);
                }
                if (shadowPdf > T(0)) {
                    const T lightPdf = light_scale(sqDist / (LNdotL * ar), lightCount, o.light_pdf);
                    const V3<T> col = thr * emission * fe * (NdotL / (shadowPdf + lightPdf));
                    const V3<T> so = org + neps * Lv;
                    T tmax = distance - T(2) * deps;
                    int dummy = -1;
                    Counters sc;
                    st.shadow++;
                    const bool occ = trace<T, true>(s, accel, so, Lv, &tmax, &dummy, teps, sc);
                    st.nsh += sc.nodes;
                    st.tsh += sc.tris;
                    if (!occ) L = L + col;
                }
            }
        }
        if (!valid) break;  // Q9: a rejected sample ends the path (after NEE)
        // --- continuation (Kernel.cu:197-205)
        if (pdf < T(1e-6f) || isnan_(pdf) || isnan_(thr.x) || isnan_(thr.y) || isnan_(thr.z) ||
            (thr.x <= T(0) && thr.y <= T(0) && thr.z <= T(0)))
            break;
        thr = vmax0(thr * f * m_abs(dot(iN, wo)));
        org = transmit ? org - N * neps : org + N * neps;
        lastPdf = pdf;
        lastType = type;
        dir = wo;
        // Russian roulette (absent in the reference): survival q = min(1,
        // max(throughput / pdf)), one extra draw, throughput /= q.
        if (o.rr_start >= 0 && bounce >= o.rr_start) {
            const V3<T> tt = thr / lastPdf;
            const T q = gmin(T(1), gmax(tt.x, gmax(tt.y, tt.z)));
            const T rr = T(ct_rand_float(&seed));
            if (!(rr < q)) break;
            thr = thr / q;
        }
    }
    st.samples++;
    return L;
}

template <class T, class A>
void render_rows(const or_scene& s, const ct_camera& cam, const ct_frame_opts& o, int width, int height,
                 uint32_t frame_base, int spp, int r0, int r1, A* accum, int threads, Stats& total) {
    const Camera3<T> cb = camera_basis<T>(cam, width, height);
    const long long npix = (long long)(r1 - r0) * width;
    std::vector<Stats> per(threads > 0 ? threads : 1);
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads > 0 ? threads : 1)
    for (long long k = 0; k < npix; ++k) {
#ifdef _OPENMP
        Stats& st = per[omp_get_thread_num()];
#else
        Stats& st = per[0];
#endif
        const int y = r0 + (int)(k / width), x = (int)(k % width);
        A* a = accum + 4 * ((size_t)y * width + x);
        for (int smp = 0; smp < spp; ++smp) {
            const V3<T> L = path_sample<T>(s, cb, o, x, y, width, frame_base + (uint32_t)smp, st);
            a[0] = a[0] + (A)L.x;
            a[1] = a[1] + (A)L.y;
            a[2] = a[2] + (A)L.z;
            a[3] = a[3] + (A)1;
        }
    }
    for (const Stats& p : per) {
        total.primary += p.primary; total.extension += p.extension; total.shadow += p.shadow;
        total.samples += p.samples; total.ncl += p.ncl; total.tcl += p.tcl; total.nsh += p.nsh; total.tsh += p.tsh;
    }
}

}  // namespace

extern "C" {

or_scene* or_scene_create(uint32_t n_vertices, const float* vertices, const float* normals, uint32_t n_tris,
                          const uint32_t* indices, const float* center_normals, const uint32_t* mat_idx,
                          uint32_t n_lights, const uint32_t* light_idx, uint32_t n_materials,
                          const ct_material* materials, int32_t sky_enabled, const float* sky) {
    or_scene* s = new or_scene();
    s->vertices.assign(vertices, vertices + 3 * (size_t)n_vertices);
    s->normals.assign(normals, normals + 3 * (size_t)n_vertices);
    s->indices.assign(indices, indices + 3 * (size_t)n_tris);
    s->center_normals.assign(center_normals, center_normals + 3 * (size_t)n_tris);
    s->mat_idx.assign(mat_idx, mat_idx + n_tris);
    if (n_lights) s->light_idx.assign(light_idx, light_idx + n_lights);
    s->materials.assign(materials, materials + n_materials);
    s->sky_enabled = sky_enabled;
    if (sky) for (int k = 0; k < 3; ++k) s->sky[k] = sky[k];
    // triangle::getBounds (Triangle.cuh:59-65): [min - 0.001, max + 0.01]
    s->aabbs.resize(n_tris);
    for (uint32_t t = 0; t < n_tris; ++t) {
        const float* a = s->vtx(t, 0);
        const float* b = s->vtx(t, 1);
        const float* c = s->vtx(t, 2);
        for (int k = 0; k < 3; ++k) {
            const float mn = gmin(a[k], gmin(b[k], c[k]));
            const float mx = gmax(a[k], gmax(b[k], c[k]));
            s->aabbs[t].mn[k] = mn - 0.001f;
            s->aabbs[t].mx[k] = mx + 0.01f;
        }
    }
    s->prims.resize(n_tris);
    for (uint32_t t = 0; t < n_tris; ++t) s->prims[t] = t;
    return s;
}

void or_scene_destroy(or_scene* s) { delete s; }

int or_scene_set_textures(or_scene* s, const float* texcoords, uint32_t n_textures, const ct_texture* textures,
                          int32_t sky_texture) {
    if (!s) return 1;
    const size_t nv = s->vertices.size() / 3;
    if (texcoords) s->texcoords.assign(texcoords, texcoords + 2 * nv);
    else s->texcoords.assign(2 * nv, 0.0f);
    s->textures.clear();
    for (uint32_t t = 0; t < n_textures; ++t) {
        or_texture x;
        x.w = textures[t].width;
        x.h = textures[t].height;
        x.rgba.assign(textures[t].rgba, textures[t].rgba + 4 * (size_t)x.w * x.h);
        s->textures.push_back(std::move(x));
    }
    s->sky_tex = n_textures ? sky_texture : -1;
    for (const ct_material& m : s->materials)
        if (m.diffuse_tex >= (int)n_textures || m.normal_tex >= (int)n_textures) return 1;
    return (s->sky_tex >= (int)n_textures) ? 1 : 0;
}

int or_texel(const or_scene* s, int32_t texture, float u, float v, float* rgba) {
    if (!s || texture < 0 || (size_t)texture >= s->textures.size() || !rgba) return 1;
    std::memcpy(rgba, texel(*s, texture, u, v), 16);
    return 0;
}

float or_atan2f(float y, float x) { return ct_atan2f(y, x); }
void or_sky_uv(float dx, float dy, float dz, float* u, float* v) { ct_sky_uv(dx, dy, dz, u, v); }

int or_scene_build(or_scene* s) {
    if (!s || s->ntris() == 0) return 1;
    for (uint32_t t = 0; t < s->ntris(); ++t) s->prims[t] = t;
    Builder b(*s);
    b.build();
    Collapser c{s->bvh, s->mbvh};
    c.build(s->ntris());
    s->built = true;
    return 0;
}

int or_scene_set_trees(or_scene* s, uint32_t n_bvh, const ct_bvh_node* bvh, uint32_t n_mbvh, const ct_mbvh_node* mbvh,
                       const uint32_t* prim_indices) {
    if (!s || !prim_indices || (n_bvh && !bvh) || (n_mbvh && !mbvh)) return 1;
    for (uint32_t i = 0; i < s->ntris(); ++i)
        if (prim_indices[i] >= s->ntris()) return 1;
    s->bvh.assign(bvh, bvh + n_bvh);
    s->mbvh.assign(mbvh, mbvh + n_mbvh);
    s->prims.assign(prim_indices, prim_indices + s->ntris());
    s->built = true;
    return 0;
}

int or_scene_tree_sizes(const or_scene* s, uint32_t* n_bvh, uint32_t* n_mbvh, uint32_t* n_prims) {
    if (!s) return 1;
    if (n_bvh) *n_bvh = (uint32_t)s->bvh.size();
    if (n_mbvh) *n_mbvh = (uint32_t)s->mbvh.size();
    if (n_prims) *n_prims = (uint32_t)s->prims.size();
    return 0;
}

int or_scene_tree_export(const or_scene* s, ct_bvh_node* bvh, ct_mbvh_node* mbvh, uint32_t* prim_indices) {
    if (!s) return 1;
    if (bvh) std::memcpy(bvh, s->bvh.data(), s->bvh.size() * sizeof(ct_bvh_node));
    if (mbvh) std::memcpy(mbvh, s->mbvh.data(), s->mbvh.size() * sizeof(ct_mbvh_node));
    if (prim_indices) std::memcpy(prim_indices, s->prims.data(), s->prims.size() * 4);
    return 0;
}

int or_trace(const or_scene* s, int accel, int precision, uint32_t n, const float* origins, const float* dirs,
             float tri_eps, double* t_out, int32_t* prim_out, uint32_t* counters, int threads) {
    if (!s || (!s->built && accel != 2)) return 1;
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads > 0 ? threads : 1)
    for (long long i = 0; i < (long long)n; ++i) {
        Counters c;
        int hit = -1;
        if (precision == 64) {
            double t = 1e34f;
            trace<double, false>(*s, accel, ld<double>(origins + 3 * i), ld<double>(dirs + 3 * i), &t, &hit,
                                 (double)tri_eps, c);
            t_out[i] = t;
        } else {
            float t = 1e34f;
            trace<float, false>(*s, accel, ld<float>(origins + 3 * i), ld<float>(dirs + 3 * i), &t, &hit, tri_eps,
                                c);
            t_out[i] = t;
        }
        prim_out[i] = hit;
        if (counters) { counters[2 * i] = c.nodes; counters[2 * i + 1] = c.tris; }
    }
    return 0;
}

int or_trace_shadow(const or_scene* s, int accel, int precision, uint32_t n, const float* origins,
                    const float* dirs, const float* tmax, float tri_eps, uint8_t* occluded_out,
                    uint32_t* counters, int threads) {
    if (!s || (!s->built && accel != 2)) return 1;
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads > 0 ? threads : 1)
    for (long long i = 0; i < (long long)n; ++i) {
        Counters c;
        int hit = -1;
        bool occ;
        if (precision == 64) {
            double t = tmax[i];
            occ = trace<double, true>(*s, accel, ld<double>(origins + 3 * i), ld<double>(dirs + 3 * i), &t, &hit,
                                      (double)tri_eps, c);
        } else {
            float t = tmax[i];
            occ = trace<float, true>(*s, accel, ld<float>(origins + 3 * i), ld<float>(dirs + 3 * i), &t, &hit,
                                     tri_eps, c);
        }
        occluded_out[i] = occ ? 1 : 0;
        if (counters) { counters[2 * i] = c.nodes; counters[2 * i + 1] = c.tris; }
    }
    return 0;
}

int or_render(const or_scene* s, int precision, const ct_camera* cam, const ct_frame_opts* opts, int32_t width,
              int32_t height, uint32_t frame_base, int32_t spp, int32_t row_begin, int32_t row_end, float* accum32,
              double* accum64, int threads, or_render_stats* stats) {
    if (!s || !s->built || !cam || !opts || width <= 0 || height <= 0 || row_begin < 0 || row_end > height ||
        row_begin > row_end)
        return 1;
    Stats st;
    if (precision == 64) {
        if (!accum64) return 1;
        render_rows<double, double>(*s, *cam, *opts, width, height, frame_base, spp, row_begin, row_end, accum64,
                                    threads, st);
    } else {
        if (!accum32) return 1;
        render_rows<float, float>(*s, *cam, *opts, width, height, frame_base, spp, row_begin, row_end, accum32,
                                  threads, st);
    }
    if (stats) {
        stats->primary_rays = st.primary;
        stats->extension_rays = st.extension;
        stats->shadow_rays = st.shadow;
        stats->samples = st.samples;
        stats->node_tests_closest = st.ncl;
        stats->tri_tests_closest = st.tcl;
        stats->node_tests_shadow = st.nsh;
        stats->tri_tests_shadow = st.tsh;
    }
    return 0;
}

float or_ct_sinf(float x) { return ct_sinf(x); }
float or_ct_cosf(float x) { return ct_cosf(x); }
float or_ct_expf(float x) { return ct_expf(x); }
float or_ct_logf(float x) { return ct_logf(x); }
float or_ct_acosf(float x) { return ct_acosf(x); }
float or_ct_powf(float x, float y) { return ct_powf(x, y); }
float or_ct_erff(float x) { return ct_erff(x); }
float or_ct_erfinvf(float x) { return ct_erfinvf(x); }
uint32_t or_wang_hash(uint32_t s) { return ct_wang_hash(s); }
uint32_t or_xorshift(uint32_t s) { ct_rand_u32(&s); return s; }

double or_mf_D(int type, double alpha, const double* wh) {
    const V3<double> w{wh[0], wh[1], wh[2]};
    return is_beckmann(type) ? D_beckmann(w, alpha) : D_ggx(w, alpha);
}
double or_mf_lambda(int type, double alpha, const double* w) {
    const V3<double> v{w[0], w[1], w[2]};
    return is_beckmann(type) ? lambda_beckmann(v, alpha) : lambda_ggx(v, alpha);
}
void or_mf_sample_vndf32(int type, float alpha, const float* wo, float r1, float r2, float* wm_out) {
    const V3<float> wm = sample_vndf<float>(type, V3<float>{wo[0], wo[1], wo[2]}, alpha, r1, r2);
    wm_out[0] = wm.x; wm_out[1] = wm.y; wm_out[2] = wm.z;
}
double or_fresnel_dielectric(double cos_i, double eta) { return fresnel_dielectric<double>(cos_i, eta); }

int or_refract(int precision, const float* I, const float* N, float eta, double* out) {
    if (precision == 64) {
        V3<double> w;
        const bool ok = refract_dir(ld<double>(I), ld<double>(N), (double)eta, &w);
        out[0] = w.x; out[1] = w.y; out[2] = w.z;
        return ok ? 1 : 0;
    }
    V3<float> w;
    const bool ok = refract_dir(ld<float>(I), ld<float>(N), eta, &w);
    out[0] = w.x; out[1] = w.y; out[2] = w.z;
    return ok ? 1 : 0;
}

int or_bsdf_local(int type, float roughness, const float* wo, const float* wi_in, float r1, float r2,
                  float* wi_out, float* f_out, float* pdf_out) {
    const V3<float> w{wo[0], wo[1], wo[2]};
    const V3<float> one{1.0f, 1.0f, 1.0f};
    const float alpha = gmax(roughness, 1e-3f);
    V3<float> wi{0, 0, 0}, f{0, 0, 0};
    float pdf = 0;
    if (wi_in) {
        wi = V3<float>{wi_in[0], wi_in[1], wi_in[2]};
    } else {
        const V3<float> wm = sample_vndf<float>(type, w, alpha, r1, r2);
        wi = reflect(-w, wm);
    }
    mf_eval<float>(type, alpha, one, w, wi, &f, &pdf);
    if (wi_out) { wi_out[0] = wi.x; wi_out[1] = wi.y; wi_out[2] = wi.z; }
    f_out[0] = f.x; f_out[1] = f.y; f_out[2] = f.z;
    *pdf_out = pdf;
    return 0;
}

}  // extern "C"
