#include "bvh.h"

#include <chrono>
#include <cstring>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace ct {

namespace {

constexpr int kMaxPrims = 4;   // BVHNode.cu:6
constexpr int kMaxDepth = 64;  // BVHNode.cu:7
constexpr int kBins = 11;      // BVHNode.cu:8

// _mm_min_ps / _mm_max_ps lane semantics (AABB.h:311-314): (a < b) ? a : b.
inline float sse_min(float a, float b) { return a < b ? a : b; }
inline float sse_max(float a, float b) { return a > b ? a : b; }

struct Box {
    float mn[3] = {1e34f, 1e34f, 1e34f};
    float mx[3] = {-1e34f, -1e34f, -1e34f};
    void grow(const AABB& b) {
        for (int k = 0; k < 3; ++k) {
            mn[k] = sse_min(mn[k], b.bmin[k]);
            mx[k] = sse_max(mx[k], b.bmax[k]);
        }
    }
    void grow(const Box& b) {
        for (int k = 0; k < 3; ++k) {
            mn[k] = sse_min(mn[k], b.mn[k]);
            mx[k] = sse_max(mx[k], b.mx[k]);
        }
    }
    // AABB::Area, AABB.h:113-121: fmax(0, e0*e1 + e0*e2 + e1*e2)
    float area() const {
        const float e0 = mx[0] - mn[0], e1 = mx[1] - mn[1], e2 = mx[2] - mn[2];
        const float a = e0 * e1 + e0 * e2 + e1 * e2;
        return a > 0.0f ? a : 0.0f;  // fmaxf(0, a), with NaN -> 0
    }
};

// AABB::Centroid, AABB.h:104-111 and Center(): (bmin + bmax) * 0.5
inline float centroid(const AABB& b, int axis) { return (b.bmin[axis] + b.bmax[axis]) * 0.5f; }

inline double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

BVHTree::BVHTree(const TriangleList* objectList) : m_ObjectList(objectList) { reset(); }

void BVHTree::reset() {
    CanUseBVH = false;
    m_BVHPool.clear();
    m_PrimitiveCount = m_ObjectList->getPrimCount();
    m_PrimitiveIndices.resize(m_PrimitiveCount);
    for (size_t i = 0; i < m_PrimitiveCount; ++i) m_PrimitiveIndices[i] = (unsigned int)i;
}

void BVHTree::calculateBounds(int nodeIdx) {
    BVHNode& node = m_BVHPool[nodeIdx];
    const AABB* aabbs = m_ObjectList->m_AABBs.data();
    const int first = node.leftFirst, count = node.count;
    Box b;
    if (count > 65536) {
#ifdef _OPENMP
        const int nt = omp_get_max_threads();
        std::vector<Box> part(nt);
#pragma omp parallel
        {
            Box local;
#pragma omp for schedule(static)
            for (int i = 0; i < count; ++i) local.grow(aabbs[m_PrimitiveIndices[first + i]]);
            part[omp_get_thread_num()] = local;
        }
        for (const Box& p : part) b.grow(p);
#else
        for (int i = 0; i < count; ++i) b.grow(aabbs[m_PrimitiveIndices[first + i]]);
#endif
    } else {
        for (int i = 0; i < count; ++i) b.grow(aabbs[m_PrimitiveIndices[first + i]]);
    }
    for (int k = 0; k < 3; ++k) {
        node.min[k] = b.mn[k];
        node.max[k] = b.mx[k];
    }
}

bool BVHTree::partition(int nodeIdx, int& left, int& right) {
    const BVHNode node = m_BVHPool[nodeIdx];
    const AABB* aabbs = m_ObjectList->m_AABBs.data();
    unsigned int* prims = m_PrimitiveIndices.data();
    const int lFirst = node.leftFirst;
    const int count = node.count;

    Box nb;
    for (int k = 0; k < 3; ++k) { nb.mn[k] = node.min[k]; nb.mx[k] = node.max[k]; }
    const float parentNodeCost = nb.area() * (float)count;
    float lengths[3];
    for (int k = 0; k < 3; ++k) lengths[k] = node.max[k] - node.min[k];

    // Split coordinates, axis-outer / plane-inner like the MT Partition.
    float split[3][kBins];
    for (int axis = 0; axis < 3; ++axis)
        for (int i = 1; i < kBins; ++i) {
            const float binOffset = (float)i / (float)kBins;
            split[axis][i] = node.min[axis] + lengths[axis] * binOffset;
        }

    // One pass: bin b(axis) = first plane i with centroid <= split[axis][i]
    // (kBins if none).  Candidate i puts bins 1..i on the left.
    Box bins[3][kBins + 1];
    int counts[3][kBins + 1];
    std::memset(counts, 0, sizeof(counts));
    auto bin_range = [&](int begin, int end, Box (&bb)[3][kBins + 1], int (&cc)[3][kBins + 1]) {
        for (int idx = begin; idx < end; ++idx) {
            const AABB& a = aabbs[prims[lFirst + idx]];
            for (int axis = 0; axis < 3; ++axis) {
                const float c = centroid(a, axis);
                int b = 1;
                while (b < kBins && !(c <= split[axis][b])) ++b;
                bb[axis][b].grow(a);
                cc[axis][b]++;
            }
        }
    };
#ifdef _OPENMP
    if (count > 65536) {
        const int nt = omp_get_max_threads();
        std::vector<Box> pb((size_t)nt * 3 * (kBins + 1));
        std::vector<int> pc((size_t)nt * 3 * (kBins + 1), 0);
#pragma omp parallel
        {
            const int t = omp_get_thread_num();
            const int n = omp_get_num_threads();
            const int begin = (int)((long long)count * t / n), end = (int)((long long)count * (t + 1) / n);
            Box lb[3][kBins + 1];
            int lc[3][kBins + 1];
            std::memset(lc, 0, sizeof(lc));
            bin_range(begin, end, lb, lc);
            for (int axis = 0; axis < 3; ++axis)
                for (int b = 0; b <= kBins; ++b) {
                    pb[((size_t)t * 3 + axis) * (kBins + 1) + b] = lb[axis][b];
                    pc[((size_t)t * 3 + axis) * (kBins + 1) + b] = lc[axis][b];
                }
        }
        for (int t = 0; t < nt; ++t)
            for (int axis = 0; axis < 3; ++axis)
                for (int b = 0; b <= kBins; ++b) {
                    bins[axis][b].grow(pb[((size_t)t * 3 + axis) * (kBins + 1) + b]);
                    counts[axis][b] += pc[((size_t)t * 3 + axis) * (kBins + 1) + b];
                }
    } else
#endif
    {
        bin_range(0, count, bins, counts);
    }

    float lowestNodeCost = 1e34f, bestCoord = 0.0f;
    int bestAxis = 0;
    for (int axis = 0; axis < 3; ++axis) {
        // suffix unions for the right side
        Box suffix[kBins + 2];
        int suffixCount[kBins + 2];
        suffixCount[kBins + 1] = 0;
        for (int b = kBins; b >= 1; --b) {
            suffix[b] = suffix[b + 1];
            suffix[b].grow(bins[axis][b]);
            suffixCount[b] = suffixCount[b + 1] + counts[axis][b];
        }
        Box leftBox;
        int leftCount = 0;
        for (int i = 1; i < kBins; ++i) {
            leftBox.grow(bins[axis][i]);
            leftCount += counts[axis][i];
            const Box& rightBox = suffix[i + 1];
            const int rightCount = suffixCount[i + 1];
            const float leftArea = leftBox.area();
            const float rightArea = rightBox.area();
            const float splitNodeCost = leftArea * (float)leftCount + rightArea * (float)rightCount;
            if (splitNodeCost < lowestNodeCost) {
                lowestNodeCost = splitNodeCost;
                bestCoord = split[axis][i];
                bestAxis = axis;
            }
        }
    }

    if (parentNodeCost < lowestNodeCost) return false;

    // In-place swap scan, BVHNode.cu:168-178.
    int lCount = 0;
    for (int idx = 0; idx < count; ++idx) {
        const AABB& a = aabbs[prims[lFirst + idx]];
        if (centroid(a, bestAxis) <= bestCoord) {
            const unsigned int tmp = prims[lFirst + idx];
            prims[lFirst + idx] = prims[lFirst + lCount];
            prims[lFirst + lCount] = tmp;
            lCount++;
        }
    }
    left = m_PoolPtr++;
    right = m_PoolPtr++;
    m_BVHPool[left].leftFirst = lFirst;
    m_BVHPool[left].count = lCount;
    m_BVHPool[right].leftFirst = lFirst + lCount;
    m_BVHPool[right].count = count - lCount;
    return true;
}

void BVHTree::subdivide(int nodeIdx, unsigned int depth) {
    depth++;
    if ((int)depth > m_MaxDepth) m_MaxDepth = (int)depth;
    if (m_BVHPool[nodeIdx].count < kMaxPrims || depth >= (unsigned)kMaxDepth) return;
    int left = -1, right = -1;
    if (!partition(nodeIdx, left, right)) return;
    m_BVHPool[nodeIdx].leftFirst = left;
    m_BVHPool[nodeIdx].count = -1;
    if (m_BVHPool[left].count > 0) {
        calculateBounds(left);
        subdivide(left, depth);
    }
    if (m_BVHPool[right].count > 0) {
        calculateBounds(right);
        subdivide(right, depth);
    }
}

void BVHTree::constructBVH() {
    if (m_ObjectList == nullptr || m_PrimitiveCount == 0) return;
    const double t0 = now_ms();
    m_BVHPool.assign(m_PrimitiveCount * 2 + 2, BVHNode());
    m_PoolPtr = 2;
    m_MaxDepth = 0;
    m_BVHPool[0].leftFirst = 0;
    m_BVHPool[0].count = (int)m_PrimitiveCount;
    calculateBounds(0);
    subdivide(0, 1);
    if (m_PoolPtr > 2) {
        m_BVHPool[0].count = -1;
        m_BVHPool[0].leftFirst = 2;
    } else {
        m_BVHPool[0].count = (int)m_PrimitiveCount;
        m_BVHPool[0].leftFirst = 0;
    }
    m_BVHPool.resize(m_PoolPtr);
    m_BuildMs = now_ms() - t0;
    CanUseBVH = true;
}

void MBVHNode::SetBounds(int slot, const float mn[3], const float mx[3]) {
    bminx[slot] = mn[0];
    bminy[slot] = mn[1];
    bminz[slot] = mn[2];
    bmaxx[slot] = mx[0];
    bmaxy[slot] = mx[1];
    bmaxz[slot] = mx[2];
}

MBVHTree::MBVHTree(const BVHTree* orgTree) : m_OriginalTree(orgTree) {
    m_PrimitiveIndices = orgTree->m_PrimitiveIndices;
}

int MBVHTree::getBVHNodeInfo(unsigned int mIdx, const BVHNode& node) {
    MBVHNode& m = m_Tree[mIdx];
    for (int i = 0; i < 4; ++i) { m.childs[i] = -1; m.counts[i] = -1; }
    int numChildren = 0;
    const std::vector<BVHNode>& pool = m_OriginalTree->m_BVHPool;
    if (node.IsLeaf()) return 0;
    const BVHNode& leftNode = pool[node.leftFirst];
    const BVHNode& rightNode = pool[node.leftFirst + 1];
    if (leftNode.IsLeaf()) {
        const int idx = numChildren++;
        m.SetBounds(idx, leftNode.min, leftNode.max);
        m.childs[idx] = leftNode.leftFirst;
        m.counts[idx] = leftNode.count;
    } else {
        const int idx1 = numChildren++;
        const int idx2 = numChildren++;
        m.childs[idx1] = leftNode.leftFirst;
        m.childs[idx2] = leftNode.leftFirst + 1;
    }
    if (rightNode.IsLeaf()) {
        const int idx = numChildren++;
        m.SetBounds(idx, rightNode.min, rightNode.max);
        m.childs[idx] = rightNode.leftFirst;
        m.counts[idx] = rightNode.count;
    } else {
        const int idx1 = numChildren++;
        const int idx2 = numChildren++;
        m.SetBounds(idx1, pool[rightNode.leftFirst].min, pool[rightNode.leftFirst].max);
        m.SetBounds(idx2, pool[rightNode.leftFirst + 1].min, pool[rightNode.leftFirst + 1].max);
        m.childs[idx1] = rightNode.leftFirst;
        m.childs[idx2] = rightNode.leftFirst + 1;
    }
    return numChildren;
}

void MBVHTree::mergeNodes(unsigned int mIdx, const BVHNode& node) {
    const std::vector<BVHNode>& pool = m_OriginalTree->m_BVHPool;
    const int numChildren = getBVHNodeInfo(mIdx, node);
    for (int idx = 0; idx < numChildren; ++idx) {
        if (m_Tree[mIdx].counts[idx] == -1) {
            const BVHNode& curNode = pool[m_Tree[mIdx].childs[idx]];
            if (curNode.IsLeaf()) {
                m_Tree[mIdx].counts[idx] = curNode.count;
                m_Tree[mIdx].childs[idx] = curNode.leftFirst;
                m_Tree[mIdx].SetBounds(idx, curNode.min, curNode.max);
            } else {
                const unsigned int newIdx = m_FinalPtr++;
                m_Tree.emplace_back();  // nodes are addressed by index: growth is safe
                m_Tree[mIdx].childs[idx] = (int)newIdx;
                m_Tree[mIdx].counts[idx] = -1;
                m_Tree[mIdx].SetBounds(idx, curNode.min, curNode.max);
                mergeNodes(newIdx, curNode);
            }
        }
    }
    const float inv_mn[3] = {1e34f, 1e34f, 1e34f}, inv_mx[3] = {-1e34f, -1e34f, -1e34f};
    for (int idx = numChildren; idx < 4; ++idx) {
        m_Tree[mIdx].SetBounds(idx, inv_mn, inv_mx);
        m_Tree[mIdx].counts[idx] = 0;
    }
}

void MBVHTree::constructBVH() {
    m_Tree.clear();
    m_CanUseBVH = false;
    const std::vector<BVHNode>& pool = m_OriginalTree->m_BVHPool;
    if (pool.empty()) return;
    const double t0 = now_ms();
    // grown one node at a time (the reference sizes it 2N + 1 up front)
    m_Tree.reserve(pool.size() / 2 + 1);
    m_Tree.emplace_back();
    m_FinalPtr = 1;
    const BVHNode& root = pool[0];
    if (root.IsLeaf()) {
        // The reference collapses a leaf root to four empty slots
        // (MBVHNode.cu:233-236), which hides the whole scene.  Here the root
        // keeps the leaf in slot 0 so small scenes stay visible.
        MBVHNode& m = m_Tree[0];
        const float inv_mn[3] = {1e34f, 1e34f, 1e34f}, inv_mx[3] = {-1e34f, -1e34f, -1e34f};
        for (int i = 0; i < 4; ++i) { m.SetBounds(i, inv_mn, inv_mx); m.childs[i] = -1; m.counts[i] = 0; }
        m.SetBounds(0, root.min, root.max);
        m.childs[0] = root.leftFirst;
        m.counts[0] = root.count;
    } else {
        mergeNodes(0, root);
    }
    m_Tree.resize(m_FinalPtr);
    m_BuildMs = now_ms() - t0;
    m_CanUseBVH = true;
}

}  // namespace ct
