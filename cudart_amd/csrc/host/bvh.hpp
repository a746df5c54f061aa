// Binned-SAH BVH build and the four-wide quantized layout the HIP kernels traverse.
// Replaces constructBVH / recursiveBuild / flattenBVHTree (src/geometry/bvh.cpp:65-323)
// and the texture repack of ParallelRenderer::createTextureObjects
// (src/renderer/parallel_renderer.cu:128-202).
#pragma once

#include <cstdint>
#include <vector>

namespace prt {

// Four-wide node with quantized child boxes, 64 bytes = 16 words (one cache line):
//   [0..2]   origin: the node box's lower corner (float)
//   [3]      per axis the biased float exponent of the step 2^e (byte k = axis k)
//   [4..7]   child words (int32)
//   [8..13]  qlo.x qhi.x qlo.y qhi.y qlo.z qhi.z: 4 bytes each, byte c = child c
//   [14..15] 0
// Child c spans fmaf(q, 2^e, origin) per plane: q rounded outward, so each decoded
// box contains the child's exact box (culling stays conservative).
// A child word >= 0 is an inner node index, kEmptyChild an unused slot. A negative
// word is a leaf: bit 31 set, bits 28..30 = count-1 (1..8 triangles), bits 0..27 =
// first record.
// Collapsed from a binary SAH tree into nodes of up to four children (BvhParams::
// collapse_sah: the collapse of least SAH cost, or greedily opening the inner
// child with the largest box area until there are four).
constexpr int kNodeFloats = 16;
constexpr uint32_t kNodeReadBytes = 64;
constexpr int32_t kEmptyChild = 0x7FFFFFFF;
constexpr int kMaxLeaf = 8;
constexpr uint32_t kMaxRecords = 1u << 28;
constexpr int kLeafCountShift = 28;
inline int32_t leaf_word(uint32_t first, uint32_t count) {
    return int32_t(0x80000000u | ((count - 1u) << kLeafCountShift) | first);
}

struct BvhLayout {
    std::vector<float> nodes;     // kNodeFloats floats per node
    std::vector<uint32_t> order;  // record index -> original triangle id
    uint32_t num_nodes = 0;
    uint32_t num_leaves = 0;
    uint32_t max_depth = 0;       // deepest four-wide node (root = 0)
    uint32_t max_leaf = 0;
    float sah_cost = 0.f;
};

#ifndef PRT_BVH_LEAF_SIZE
#define PRT_BVH_LEAF_SIZE 2  // measured best for the four-wide layout (2: 197 ms, 3-6: 203 ms per 1080p frame)
#endif
#ifndef PRT_BVH_COLLAPSE_SAH
#define PRT_BVH_COLLAPSE_SAH 1  // 1080p spheres frame: 235,118 nodes (greedy 299,869), extend + shadow 1.3% faster
#endif
#ifndef PRT_BVH_TRI_COST
#define PRT_BVH_TRI_COST 1.0f
#endif
#ifndef PRT_BVH_MERGE_MAX
#define PRT_BVH_MERGE_MAX 2
#endif
#ifndef PRT_BVH_TREELET_PASSES
#define PRT_BVH_TREELET_PASSES 0
#endif
struct BvhParams {
    int leaf_size = PRT_BVH_LEAF_SIZE;  // SAH leaves hold at most this many triangles
    // Four-wide collapse: 0 = open the largest-area child until four; 1 = the
    // collapse of least SAH cost (node visit 1, triangle test tri_cost), which may
    // also write a subtree of at most merge_max triangles as one leaf.
    int collapse_sah = PRT_BVH_COLLAPSE_SAH;
    float tri_cost = PRT_BVH_TRI_COST;
    int merge_max = PRT_BVH_MERGE_MAX;
    int treelet_passes = PRT_BVH_TREELET_PASSES;  // treelet restructuring passes over the binary tree
    int sah_depth_cap = 64;   // below this depth, object-median splits only
// Leaf box padding relative to its extent, on top of the absolute part (PRT_BVH_PAD_ABS
// of the largest scene coordinate, in bvh.cpp), which is the oracle's padding: box
// rejection must not hide a hit fp32 Moller-Trumbore accepts, and for grazing rays
// from far away those can lie slightly outside the triangle. 0 gives the product's
// leaves the oracle's margin (plus the outward rounding of the quantized planes);
// 1/128 measured 3% slower per 1080p frame (larger boxes, more visits).
#ifndef PRT_BVH_PAD_REL
#define PRT_BVH_PAD_REL 0.0f
#endif
    float pad_rel = PRT_BVH_PAD_REL;
};

// positions: n triangles x 3 vertices x xyz. abs_max: largest |coordinate| (>= 1).
void build_bvh(const float* positions, uint32_t n, const BvhParams& params, BvhLayout& out);

}  // namespace prt
