// Wide compressed nodes (see cnode.h).
#include "cnode.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace ct {
namespace {

uint32_t fbits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

// The decode the kernel performs: fmaf(q, s, p), one rounding.
float decode(uint32_t q, float s, float p) { return std::fmaf((float)q, s, p); }

// A child of a wide node before encoding: a box and either a leaf (range of
// the primitive list) or an interior node of the source tree.
struct Item {
    float mn[3], mx[3];
    bool leaf;
    uint32_t first, count;
    int32_t src, slot;  // binary node, or quad node and slot
};

float half_area(const Item& it) {
    const float dx = it.mx[0] - it.mn[0], dy = it.mx[1] - it.mn[1], dz = it.mx[2] - it.mn[2];
    return dx * dy + dy * dz + dz * dx;
}

struct Source {
    const ct_scene_view& v;
    bool binary;

    Item from_binary(uint32_t b) const {
        const ct_bvh_node& n = v.bvh_nodes[b];
        Item it{};
        std::memcpy(it.mn, n.min, 12);
        std::memcpy(it.mx, n.max, 12);
        it.leaf = n.count > -1;
        it.first = (uint32_t)n.left_first;
        it.count = it.leaf ? (uint32_t)n.count : 0u;
        it.src = (int32_t)b;
        return it;
    }
    Item from_quad(uint32_t q, int s) const {
        const ct_mbvh_node& n = v.mbvh_nodes[q];
        Item it{};
        it.mn[0] = n.bminx[s];
        it.mn[1] = n.bminy[s];
        it.mn[2] = n.bminz[s];
        it.mx[0] = n.bmaxx[s];
        it.mx[1] = n.bmaxy[s];
        it.mx[2] = n.bmaxz[s];
        it.leaf = n.counts[s] > 0;
        it.first = (uint32_t)n.childs[s];
        it.count = it.leaf ? (uint32_t)n.counts[s] : 0u;
        it.src = (int32_t)q;
        it.slot = s;
        return it;
    }
    void quad_slots(uint32_t q, std::vector<Item>& out) const {
        for (int s = 0; s < 4; ++s)
            if (v.mbvh_nodes[q].counts[s] != 0) out.push_back(from_quad(q, s));
    }
    void root(std::vector<Item>& out) const {
        if (binary) {
            const Item r = from_binary(0);
            if (r.leaf) {
                out.push_back(r);
            } else {
                out.push_back(from_binary(r.first));
                out.push_back(from_binary(r.first + 1));
            }
        } else {
            quad_slots(0, out);
        }
    }
    // children of an interior item
    void expand(const Item& it, std::vector<Item>& out) const {
        if (binary) {
            out.push_back(from_binary(it.first));
            out.push_back(from_binary(it.first + 1));
        } else {
            quad_slots((uint32_t)v.mbvh_nodes[it.src].childs[it.slot], out);
        }
    }
};

// Open interior items, largest surface area first, while the node keeps at
// most kWideArity children.
void collapse(const Source& src, std::vector<Item>& items) {
    std::vector<Item> kids;
    while (true) {
        int best = -1;
        float best_area = -1.0f;
        for (size_t k = 0; k < items.size(); ++k) {
            if (items[k].leaf) continue;
            kids.clear();
            src.expand(items[k], kids);
            if (items.size() - 1 + kids.size() > (size_t)kWideArity) continue;
            const float a = half_area(items[k]);
            if (a > best_area) {
                best_area = a;
                best = (int)k;
            }
        }
        if (best < 0) return;
        kids.clear();
        src.expand(items[(size_t)best], kids);
        items.erase(items.begin() + best);
        items.insert(items.end(), kids.begin(), kids.end());
    }
}

struct AxisCode {
    float p;
    int k;  // step 2^k
    uint8_t lo[kWideArity], hi[kWideArity];
};

// Quantise the children's [lo, hi] along one axis (n children, range [mn, mx]):
// the smallest step whose grid holds the padded range, coarser steps until
// every decoded plane contains its box in float arithmetic.
bool encode_axis(const float* lo, const float* hi, int n, float mn, float mx, AxisCode& ac) {
    const double ext = (double)mx - (double)mn;
    int k = -126;
    if (ext > 0.0) {
        k = (int)std::ceil(std::log2(ext / 250.0));
        if (k < -126) k = -126;
    }
    for (; k <= 127; ++k) {
        const float s = std::ldexp(1.0f, k);
        float p = (float)((double)mn - (double)s);
        if ((double)p > (double)mn - (double)s) p = std::nextafter(p, -INFINITY);
        if (!std::isfinite(p)) return false;
        bool ok = true;
        for (int c = 0; c < n && ok; ++c) {
            const double ql = std::floor(((double)lo[c] - (double)p) / (double)s) - 1.0;
            const double qh = std::ceil(((double)hi[c] - (double)p) / (double)s) + 1.0;
            if (ql < 0.0 || qh > 255.0 || ql > qh) { ok = false; break; }
            const uint32_t a = (uint32_t)ql, b = (uint32_t)qh;
            if (!(decode(a, s, p) <= lo[c]) || !(decode(b, s, p) >= hi[c])) { ok = false; break; }
            ac.lo[c] = (uint8_t)a;
            ac.hi[c] = (uint8_t)b;
        }
        if (ok) {
            for (int c = n; c < kWideArity; ++c) ac.lo[c] = ac.hi[c] = 0;
            ac.p = p;
            ac.k = k;
            return true;
        }
    }
    return false;
}

bool encode_node(const std::vector<Item>& items, const std::vector<uint32_t>& entries, uint32_t* w) {
    std::memset(w, 0, kWideNodeWords * 4);
    const int n = (int)items.size();
    uint32_t exps = 0;
    for (int a = 0; a < 3; ++a) {
        float lo[kWideArity], hi[kWideArity];
        float mn = INFINITY, mx = -INFINITY;
        for (int c = 0; c < n; ++c) {
            lo[c] = items[(size_t)c].mn[a];
            hi[c] = items[(size_t)c].mx[a];
            if (!std::isfinite(lo[c]) || !std::isfinite(hi[c])) return false;
            mn = std::fmin(mn, lo[c]);
            mx = std::fmax(mx, hi[c]);
        }
        AxisCode ac{};
        if (!encode_axis(lo, hi, n, mn, mx, ac)) return false;
        w[a] = fbits(ac.p);
        exps |= (uint32_t)(ac.k + 127) << (8 * a);
        for (int c = 0; c < kWideArity; ++c) {
            w[4 + 4 * a + c / 4] |= (uint32_t)ac.lo[c] << (8 * (c % 4));
            w[6 + 4 * a + c / 4] |= (uint32_t)ac.hi[c] << (8 * (c % 4));
        }
    }
    const uint32_t valid = n >= 8 ? 0xFFu : ((1u << n) - 1u);
    w[3] = exps | (valid << 24);
    for (int c = 0; c < n; ++c) w[16 + c] = entries[(size_t)c];
    return true;
}

}  // namespace

int build_wide_nodes(const ct_scene_view& v, std::vector<uint32_t>& words, uint32_t& stack_bound) {
    words.clear();
    stack_bound = 0;
    if (v.n_bvh_nodes == 0 && v.n_mbvh_nodes == 0) return 1;
    const Source src{v, v.n_bvh_nodes > 0};
    // nodes are numbered in creation order; a node's children are created
    // (contiguously) when it is processed, so children follow their parent.
    // A node is encoded as soon as its children have numbers.
    std::vector<std::vector<Item>> items(1);
    src.root(items[0]);
    std::vector<uint32_t> entries;
    for (size_t i = 0; i < items.size(); ++i) {
        collapse(src, items[i]);
        entries.assign(items[i].size(), 0u);
        for (size_t c = 0; c < items[i].size(); ++c) {
            const Item it = items[i][c];
            if (it.leaf) {
                entries[c] = 0x80000000u | it.first;
            } else {
                entries[c] = (uint32_t)items.size();
                items.emplace_back();
                src.expand(it, items.back());
            }
        }
        words.resize((i + 1) * kWideNodeWords);
        if (!encode_node(items[i], entries, words.data() + i * kWideNodeWords)) return 1;
        std::vector<Item>().swap(items[i]);
    }
    // worst-case stack: a node pushes all n children; while one is processed
    // the other n - 1 stay on the stack
    const size_t nn = items.size();
    std::vector<uint32_t> bound(nn, 0u);
    for (size_t i = nn; i-- > 0;) {
        const uint32_t* w = words.data() + i * kWideNodeWords;
        const uint32_t n = (uint32_t)__builtin_popcount(w[3] >> 24);
        uint32_t sub = 0;
        for (uint32_t c = 0; c < n; ++c)
            if (!(w[16 + c] & 0x80000000u)) sub = std::max(sub, bound[w[16 + c]]);
        bound[i] = sub ? std::max(n, n - 1 + sub) : n;
    }
    stack_bound = bound[0];
    return 0;
}

}  // namespace ct
