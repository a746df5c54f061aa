// Binned-SAH BVH build and the child-pair layout the HIP kernels traverse.
// Replaces constructBVH / recursiveBuild / flattenBVHTree (src/geometry/bvh.cpp:65-323)
// and the texture repack of ParallelRenderer::createTextureObjects
// (src/renderer/parallel_renderer.cu:128-202).
#pragma once

#include <cstdint>
#include <vector>

namespace prt {

// Child-pair node, 64 bytes = 16 floats:
//   [0..3]   c0.lo.x c0.hi.x c0.lo.y c0.hi.y
//   [4..7]   c1.lo.x c1.hi.x c1.lo.y c1.hi.y
//   [8..11]  c0.lo.z c0.hi.z c1.lo.z c1.hi.z
//   [12..15] child0 child1 (int32 bit patterns), 0, 0
// A child word >= 0 is an inner node index. A negative word is a leaf:
//   bit 31 set, bits 28..30 = count-1 (1..8 triangles), bits 0..27 = first record.
constexpr int kMaxLeaf = 8;
constexpr uint32_t kMaxRecords = 1u << 28;
constexpr int kLeafCountShift = 28;
inline int32_t leaf_word(uint32_t first, uint32_t count) {
    return int32_t(0x80000000u | ((count - 1u) << kLeafCountShift) | first);
}

struct BvhLayout {
    std::vector<float> nodes;     // 16 floats per node
    std::vector<uint32_t> order;  // record index -> original triangle id
    uint32_t num_nodes = 0;
    uint32_t num_leaves = 0;
    uint32_t max_depth = 0;       // deepest child-pair node (root = 0)
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
