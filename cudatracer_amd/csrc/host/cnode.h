// Wide compressed nodes for the fast traversal mode.
//
// The binary tree (or, without one, the quad tree) is collapsed to a tree of
// up to 8 children per node, opening the child of largest surface area
// until 8 are reached; its leaves are the original leaves, so the leaf-ordered
// triangle records serve both modes.  Each child box is stored as 8-bit
// offsets on a per-node grid, a corner p and a power-of-two step per axis,
// after Ylitie, Karras and Laine, "Efficient Incoherent Ray Traversal on GPUs
// Through Compressed Wide BVHs" (HPG 2017).  Quantisation rounds outwards and
// pads one more step on each side, and every decoded plane fmaf(q, s, p) is
// checked in float to contain the original box, so the wide tree never culls
// a box the exact trees would enter.
//
// One node is 32 words (128 bytes, one cache line):
//   0..2   p.x, p.y, p.z (float bits)
//   3      biased exponents of the x, y, z steps (bytes 0..2) | valid-child mask << 24
//   4..5   low x offsets of children 0..7 (one byte each), 6..7 high x offsets
//   8..11  y (low, high), 12..15 z (low, high)
//   16..23 child entries: wide node index, or 0x80000000 | first triangle
//          record of a leaf (leaves run to the record with the last-in-leaf bit)
//   24..31 zero
#pragma once

#include <stdint.h>

#include <vector>

#include "../../../include/cudatracer.h"

namespace ct {

constexpr int kWideNodeWords = 32;
constexpr int kWideArity = 8;

// Build the wide nodes of a scene view (binary tree if present, else quad
// tree).  stack_bound receives the largest traversal stack the tree can need
// (entries).  Returns 0, or 1 if the view has no tree or a box is not finite.
int build_wide_nodes(const ct_scene_view& v, std::vector<uint32_t>& words, uint32_t& stack_bound);

}  // namespace ct
