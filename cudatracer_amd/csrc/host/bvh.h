// Binary SAH BVH and its 4-wide collapse, with the reference's class names.
//
//   BVHTree::buildBVH         src/BVH/BVHTree.cpp:22-58
//   BVHNode::SubdivideMT      src/BVH/BVHNode.cu:70-119  (leaf if count < 4 or depth >= 64)
//   BVHNode::Partition (MT)   src/BVH/BVHNode.cu:121-191 (axis-outer, 10 planes per axis)
//   BVHNode::CalculateBounds  src/BVH/BVHNode.cu:25-38
//   MBVHTree::constructBVH    src/BVH/MBVHTree.cpp:15-30
//   MBVHNode::MergeNodes      src/BVH/MBVHNode.cu:126-160
//   MBVHNode::GetBVHNodeInfo  src/BVH/MBVHNode.cu:226-277
//
// The split decision is computed from one binning pass per node instead of
// the reference's 30 full scans.  Every candidate plane sees exactly the
// same left/right sets, the same child boxes (min/max unions are exact and
// order-free) and the same float cost expression, so the chosen plane, the
// partition order of the primitive indices and hence the tree are the
// reference's.  Node indices follow the single-threaded allocation order
// (children allocated as a pair when a node splits, left subtree first); the
// reference's multi-threaded build allocates the same topology in thread order.
#pragma once

#include <cstdint>
#include <vector>

#include "../../../include/cudatracer.h"
#include "scene.h"

namespace ct {

struct BVHNode {
    float min[3] = {0, 0, 0};
    int leftFirst = -1;
    float max[3] = {0, 0, 0};
    int count = -1;
    bool IsLeaf() const { return count > -1; }
};
static_assert(sizeof(BVHNode) == 32, "BVHNode is 32 bytes");

class BVHTree {
public:
    explicit BVHTree(const TriangleList* objectList);
    void constructBVH();
    void reset();

    std::vector<BVHNode> m_BVHPool;
    std::vector<unsigned int> m_PrimitiveIndices;
    size_t m_PrimitiveCount = 0;
    const TriangleList* m_ObjectList;
    bool CanUseBVH = false;
    int m_PoolPtr = 0;
    int m_MaxDepth = 0;
    double m_BuildMs = 0.0;

private:
    void subdivide(int nodeIdx, unsigned int depth);
    bool partition(int nodeIdx, int& left, int& right);
    void calculateBounds(int nodeIdx);
};

struct MBVHNode {
    float bminx[4], bmaxx[4], bminy[4], bmaxy[4], bminz[4], bmaxz[4];
    int childs[4];
    int counts[4];
    void SetBounds(int slot, const float mn[3], const float mx[3]);
};
static_assert(sizeof(MBVHNode) == 128, "MBVHNode is 128 bytes");

class MBVHTree {
public:
    explicit MBVHTree(const BVHTree* orgTree);
    void constructBVH();

    const BVHTree* m_OriginalTree;
    std::vector<MBVHNode> m_Tree;
    std::vector<unsigned int> m_PrimitiveIndices;
    bool m_CanUseBVH = false;
    unsigned int m_FinalPtr = 0;
    double m_BuildMs = 0.0;

private:
    void mergeNodes(unsigned int mIdx, const BVHNode& node);
    int getBVHNodeInfo(unsigned int mIdx, const BVHNode& node);
};

}  // namespace ct
