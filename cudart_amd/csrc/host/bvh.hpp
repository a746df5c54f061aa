// Binned-SAH BVH build and the child-pair layout the HIP kernels traverse.
// Replaces constructBVH / recursiveBuild / flattenBVHTree (src/geometry/bvh.cpp:65-323)
// and the texture repack of ParallelRenderer::createTextureObjects
// (src/renderer/parallel_renderer.cu:128-202).
#pragma once

#include <cstdint>
#include <vector>

namespace prt {

// Four-wide node, 128 bytes = 32 floats (one 128-byte line; a visit reads the
// first 112 bytes), boxes of the four children as structure of arrays:
//   [0..3]   lo.x of children 0..3     [4..7]   hi.x
//   [8..11]  lo.y                      [12..15] hi.y
//   [16..19] lo.z                      [20..23] hi.z
//   [24..27] child words (int32 bit patterns)      [28..31] 0
// A child word >= 0 is an inner node index, kEmptyChild an unused slot (its box
// is zero). A negative word is a leaf:
//   bit 31 set, bits 28..30 = count-1 (1..8 triangles), bits 0..27 = first record.
// Collapsed from a binary SAH tree: a node takes its two children and keeps
// opening the inner child with the largest box area until it has four.
constexpr int kNodeFloats = 32;
constexpr uint32_t kNodeReadBytes = 112;
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

struct BvhParams {
    int leaf_size = 4;        // SAH leaves hold at most this many triangles
    int sah_depth_cap = 64;   // below this depth, object-median splits only
    float pad_rel = 1.0f / 128.0f;  // leaf box padding relative to its extent
};

// positions: n triangles x 3 vertices x xyz. abs_max: largest |coordinate| (>= 1).
void build_bvh(const float* positions, uint32_t n, const BvhParams& params, BvhLayout& out);

}  // namespace prt
