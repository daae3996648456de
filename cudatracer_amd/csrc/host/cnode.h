// Compressed quad nodes for the fast traversal mode (64 bytes per node,
// half the reference's MBVHNode).
//
// Each child box is stored as 8-bit offsets on a per-node grid: a corner p
// (three floats) and a power-of-two step s per axis, after Ylitie, Karras and
// Laine, "Efficient Incoherent Ray Traversal on GPUs Through Compressed Wide
// BVHs" (HPG 2017).  Quantisation rounds outwards and then pads one more step
// on each side, and every decoded plane fmaf(q, s, p) is checked in float to
// contain the original box, so the compressed tree never culls a box the
// exact tree would enter.
//
// Words of node i (16 x u32):
//   0..2  p.x, p.y, p.z (float bits)
//   3     biased exponents of s.x, s.y, s.z (bytes 0..2) | valid-child mask << 24
//   4, 5  x: low bytes of children 0..3, high bytes of children 0..3
//   6, 7  y: likewise
//   8, 9  z: likewise
//   10..13 child entries: interior node index, or 0x80000000 | first triangle
//          record of a leaf (the same entries the exact traversal pushes)
//   14, 15 zero
#pragma once

#include <stdint.h>

#include "../../../include/cudatracer.h"

namespace ct {

constexpr int kCompressedNodeWords = 16;

// Compress n quad nodes into out[16 * n].  Returns 0, or 1 if a node's box is
// not finite (the caller validates the tree before).
int compress_quad_nodes(const ct_mbvh_node* nodes, uint32_t n, uint32_t* out);

}  // namespace ct
