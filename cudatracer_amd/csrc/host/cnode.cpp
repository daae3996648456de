// Compressed quad nodes (see cnode.h).
#include "cnode.h"

#include <cmath>
#include <cstring>

namespace ct {
namespace {

uint32_t fbits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

// 2^k as a float for k in [-126, 127].
float pow2f(int k) { return std::ldexp(1.0f, k); }

// The decode the kernel performs: fmaf(q, s, p), one rounding.
float decode(uint32_t q, float s, float p) { return std::fmaf((float)q, s, p); }

struct AxisCode {
    float p;
    int k;  // s = 2^k
    uint8_t lo[4], hi[4];
};

// Quantise the valid children's [lo, hi] along one axis.  The node range is
// [mn, mx] (union of the children).  Try the smallest step whose grid holds
// the padded range, then coarser steps until every decoded plane contains
// its box in float arithmetic.
bool encode_axis(const float* lo, const float* hi, unsigned valid, float mn, float mx, AxisCode& ac) {
    const double ext = (double)mx - (double)mn;
    int k = -126;
    if (ext > 0.0) {
        k = (int)std::ceil(std::log2(ext / 250.0));
        if (k < -126) k = -126;
    }
    for (; k <= 127; ++k) {
        const float s = pow2f(k);
        // corner one step below the range, rounded down to a float
        float p = (float)((double)mn - (double)s);
        if ((double)p > (double)mn - (double)s) p = std::nextafter(p, -INFINITY);
        if (!std::isfinite(p)) return false;
        bool ok = true;
        for (int c = 0; c < 4 && ok; ++c) {
            if (!(valid >> c & 1u)) {
                ac.lo[c] = 0;
                ac.hi[c] = 0;
                continue;
            }
            const double ql = std::floor(((double)lo[c] - (double)p) / (double)s) - 1.0;
            const double qh = std::ceil(((double)hi[c] - (double)p) / (double)s) + 1.0;
            if (ql < 0.0 || qh > 255.0 || ql > qh) { ok = false; break; }
            uint32_t a = (uint32_t)ql, b = (uint32_t)qh;
            if (!(decode(a, s, p) <= lo[c]) || !(decode(b, s, p) >= hi[c])) { ok = false; break; }
            ac.lo[c] = (uint8_t)a;
            ac.hi[c] = (uint8_t)b;
        }
        if (ok) {
            ac.p = p;
            ac.k = k;
            return true;
        }
    }
    return false;
}

}  // namespace

int compress_quad_nodes(const ct_mbvh_node* nodes, uint32_t n, uint32_t* out) {
    for (uint32_t i = 0; i < n; ++i) {
        const ct_mbvh_node& nd = nodes[i];
        uint32_t* w = out + (size_t)kCompressedNodeWords * i;
        std::memset(w, 0, kCompressedNodeWords * 4);
        unsigned valid = 0;
        for (int c = 0; c < 4; ++c)
            if (nd.counts[c] != 0) valid |= 1u << c;
        const float* los[3] = {nd.bminx, nd.bminy, nd.bminz};
        const float* his[3] = {nd.bmaxx, nd.bmaxy, nd.bmaxz};
        uint32_t exps = 0;
        for (int a = 0; a < 3; ++a) {
            float mn = INFINITY, mx = -INFINITY;
            for (int c = 0; c < 4; ++c) {
                if (!(valid >> c & 1u)) continue;
                if (!std::isfinite(los[a][c]) || !std::isfinite(his[a][c])) return 1;
                mn = std::fmin(mn, los[a][c]);
                mx = std::fmax(mx, his[a][c]);
            }
            AxisCode ac{};
            if (valid == 0) {
                ac.p = 0.0f;
                ac.k = 0;
            } else if (!encode_axis(los[a], his[a], valid, mn, mx, ac)) {
                return 1;
            }
            w[a] = fbits(ac.p);
            exps |= (uint32_t)(ac.k + 127) << (8 * a);
            uint32_t lw = 0, hw = 0;
            for (int c = 0; c < 4; ++c) {
                lw |= (uint32_t)ac.lo[c] << (8 * c);
                hw |= (uint32_t)ac.hi[c] << (8 * c);
            }
            w[4 + 2 * a] = lw;
            w[5 + 2 * a] = hw;
        }
        w[3] = exps | (valid << 24);
        for (int c = 0; c < 4; ++c) {
            if (!(valid >> c & 1u)) continue;
            const uint32_t ch = (uint32_t)nd.childs[c];
            w[10 + c] = nd.counts[c] == -1 ? ch : (0x80000000u | ch);
        }
    }
    return 0;
}

}  // namespace ct
