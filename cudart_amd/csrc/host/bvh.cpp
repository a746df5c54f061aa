// Binned-SAH BVH build (Wald 2007, "On fast construction of SAH-based bounding
// volume hierarchies") over triangle centroids, 64 bins on all three axes, then
// collapsed depth first into the four-wide quantized layout the HIP traversal reads.
//
// Compared with recursiveBuild (src/geometry/bvh.cpp:114-307) this keeps the same
// idea (binned SAH, small leaves, object-median fallback) but: bins all axes, uses
// true surface area (the reference multiplies the three extents, bvh.cpp:334-337),
// initialises every bound (F6), runs subtrees as OpenMP tasks, and caps the depth
// so the traversal stack is bounded (F7). The tree depends only on the input, not
// on the thread count: partitions are in place on disjoint ranges and node
// numbering is a depth-first walk after the build.
#include "bvh.hpp"

#include <omp.h>

#include <algorithm>
#include <limits>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <stdexcept>

namespace prt {
namespace {

// SAH bins per axis: 64 measured 1% faster per 1080p frame than 32 (SAH cost 0.5%
// lower; build 1.6 s instead of 0.7 s for 1M triangles on 8 cores).
#ifndef PRT_BVH_BINS
#define PRT_BVH_BINS 64
#endif
constexpr int kBins = PRT_BVH_BINS;
constexpr uint32_t kTaskGrain = 32768;

struct Prim {
    float lo[3], hi[3], c[3];
};
struct BNode {
    float lo[3], hi[3];
    int32_t child[2];
    uint32_t start, count, depth;
};

inline void box_init(float lo[3], float hi[3]) {
    for (int k = 0; k < 3; ++k) {
        lo[k] = FLT_MAX;
        hi[k] = -FLT_MAX;
    }
}
inline void box_grow(float lo[3], float hi[3], const float a[3], const float b[3]) {
    for (int k = 0; k < 3; ++k) {
        lo[k] = std::min(lo[k], a[k]);
        hi[k] = std::max(hi[k], b[k]);
    }
}
inline float box_area(const float lo[3], const float hi[3]) {
    float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    if (dx < 0.f || dy < 0.f || dz < 0.f) return 0.f;
    return 2.f * (dx * dy + dy * dz + dz * dx);
}

struct Ctx {
    const std::vector<Prim>* prims;
    uint32_t* idx;
    BNode* nodes;
    std::atomic<uint32_t> next{0};
    BvhParams p;
};

// Bin of a centroid offset already scaled to [0, kBins): truncation, clamped to the
// bins. NaN (an infinite extent times a zero scale, or non-finite input) goes to
// bin 0: converting it to int would be undefined.
inline int bin_of(float f) { return f > 0.f ? (f < float(kBins - 1) ? int(f) : kBins - 1) : 0; }

void build_rec(Ctx& c, uint32_t me, uint32_t start, uint32_t end, uint32_t depth) {
    const std::vector<Prim>& P = *c.prims;
    BNode& nd = c.nodes[me];
    nd.start = start;
    nd.count = end - start;
    nd.child[0] = nd.child[1] = -1;
    nd.depth = depth;
    float clo[3], chi[3];
    box_init(nd.lo, nd.hi);
    box_init(clo, chi);
    for (uint32_t i = start; i < end; ++i) {
        const Prim& q = P[c.idx[i]];
        box_grow(nd.lo, nd.hi, q.lo, q.hi);
        box_grow(clo, chi, q.c, q.c);
    }
    const uint32_t n = end - start;
    if (n == 1) return;

    uint32_t mid = start;
    bool have_split = false;
    if (depth < uint32_t(c.p.sah_depth_cap)) {
        int best_axis = -1, best_bin = -1;
        float best_cost = FLT_MAX;
        for (int ax = 0; ax < 3; ++ax) {
            float ext = chi[ax] - clo[ax];
            if (!(ext > 0.f)) continue;
            float scale = float(kBins) / ext;
            uint32_t cnt[kBins] = {0};
            float blo[kBins][3], bhi[kBins][3];
            for (int b = 0; b < kBins; ++b) box_init(blo[b], bhi[b]);
            for (uint32_t i = start; i < end; ++i) {
                const Prim& q = P[c.idx[i]];
                int b = bin_of((q.c[ax] - clo[ax]) * scale);
                cnt[b]++;
                box_grow(blo[b], bhi[b], q.lo, q.hi);
            }
            float rarea[kBins];
            uint32_t rcnt[kBins];
            float lo[3], hi[3];
            box_init(lo, hi);
            uint32_t acc = 0;
            for (int b = kBins - 1; b > 0; --b) {
                box_grow(lo, hi, blo[b], bhi[b]);
                acc += cnt[b];
                rarea[b] = box_area(lo, hi);
                rcnt[b] = acc;
            }
            box_init(lo, hi);
            acc = 0;
            for (int b = 0; b < kBins - 1; ++b) {
                box_grow(lo, hi, blo[b], bhi[b]);
                acc += cnt[b];
                if (acc == 0 || rcnt[b + 1] == 0) continue;
                float cost = box_area(lo, hi) * float(acc) + rarea[b + 1] * float(rcnt[b + 1]);
                if (cost < best_cost) {
                    best_cost = cost;
                    best_axis = ax;
                    best_bin = b;
                }
            }
        }
        if (best_axis >= 0) {
            float parea = box_area(nd.lo, nd.hi);
            float split_cost = 1.0f + (parea > 0.f ? best_cost / parea : 0.f);
            float leaf_cost = float(n);
            if (n <= uint32_t(c.p.leaf_size) && leaf_cost <= split_cost) return;  // SAH leaf
            const int ax = best_axis;
            const float scale = float(kBins) / (chi[ax] - clo[ax]);
            const float cmin = clo[ax];
            uint32_t* it = std::partition(c.idx + start, c.idx + end, [&](uint32_t t) {
                return bin_of((P[t].c[ax] - cmin) * scale) <= best_bin;
            });
            mid = uint32_t(it - c.idx);
            have_split = mid > start && mid < end;
        } else if (n <= uint32_t(c.p.leaf_size)) {
            return;  // all centroids coincide: one leaf
        }
    } else if (n <= uint32_t(c.p.leaf_size)) {
        return;
    }
    if (!have_split) {
        // Object median on the axis of largest centroid extent (index middle if none).
        mid = start + n / 2;
        float d[3] = {chi[0] - clo[0], chi[1] - clo[1], chi[2] - clo[2]};
        int ax = (d[0] >= d[1] && d[0] >= d[2]) ? 0 : (d[1] >= d[2] ? 1 : 2);
        if (d[ax] > 0.f)
            std::nth_element(c.idx + start, c.idx + mid, c.idx + end, [&](uint32_t a, uint32_t b) {
                if (P[a].c[ax] != P[b].c[ax]) return P[a].c[ax] < P[b].c[ax];
                return a < b;
            });
    }
    uint32_t l = c.next.fetch_add(2);
    uint32_t r = l + 1;
    nd.child[0] = int32_t(l);
    nd.child[1] = int32_t(r);
    if (n > kTaskGrain) {
#pragma omp task default(none) firstprivate(l, start, mid, depth) shared(c)
        build_rec(c, l, start, mid, depth + 1);
        build_rec(c, r, mid, end, depth + 1);
#pragma omp taskwait
    } else {
        build_rec(c, l, start, mid, depth + 1);
        build_rec(c, r, mid, end, depth + 1);
    }
}

// Treelet restructuring of the binary tree (Karras & Aila, "Fast parallel
// construction of high-quality bounding volume hierarchies", HPG 2013). For every
// inner node r, children before parents: the treelet of r is grown from r's two
// children by opening the largest-area inner node among its leaves until it has
// kTreeletLeaves leaves; then the binary topology over those leaves of least total
// inner-node area (dynamic programming over the 2^n leaf subsets) replaces the
// treelet's inner nodes when it is cheaper. Areas are of boxes grown by `pad` (the
// leaves' padding). Only nodes inside r's subtree change, so the two subtrees of a
// node are processed in parallel. Node indices lose the parent-before-child order;
// renumber() restores it and makes every subtree's triangles contiguous again.
constexpr int kTreeletLeaves = 7;
constexpr uint32_t kTreeletTaskGrain = 32768;

struct TreeletOpt {
    BNode* bn;
    float pad;

    float area(const float lo[3], const float hi[3]) const {
        const float pl[3] = {lo[0] - pad, lo[1] - pad, lo[2] - pad}, ph[3] = {hi[0] + pad, hi[1] + pad, hi[2] + pad};
        return box_area(pl, ph);
    }
    bool inner(int32_t b) const { return bn[b].child[0] >= 0; }

    void optimize(int32_t r) {
        int32_t leaves[kTreeletLeaves], internals[kTreeletLeaves];
        int nl = 2, ni = 1;
        leaves[0] = bn[r].child[0];
        leaves[1] = bn[r].child[1];
        internals[0] = r;
        while (nl < kTreeletLeaves) {
            int best = -1;
            float ba = -1.f;
            for (int k = 0; k < nl; ++k)
                if (inner(leaves[k])) {
                    const float a = area(bn[leaves[k]].lo, bn[leaves[k]].hi);
                    if (a > ba) {
                        ba = a;
                        best = k;
                    }
                }
            if (best < 0) break;
            const int32_t open = leaves[best];
            internals[ni++] = open;
            leaves[best] = bn[open].child[0];
            leaves[nl++] = bn[open].child[1];
        }
        if (nl < 3) return;
        const uint32_t full = (1u << nl) - 1u;
        float sa[1 << kTreeletLeaves], cost[1 << kTreeletLeaves];
        uint8_t split[1 << kTreeletLeaves];
        float blo[1 << kTreeletLeaves][3], bhi[1 << kTreeletLeaves][3];
        for (uint32_t s = 1; s <= full; ++s) {
            const int low = __builtin_ctz(s);
            const uint32_t rest = s & (s - 1u);
            const BNode& L = bn[leaves[low]];
            for (int k = 0; k < 3; ++k) {
                blo[s][k] = rest ? std::min(blo[rest][k], L.lo[k]) : L.lo[k];
                bhi[s][k] = rest ? std::max(bhi[rest][k], L.hi[k]) : L.hi[k];
            }
            sa[s] = area(blo[s], bhi[s]);
        }
        for (uint32_t s = 1; s <= full; ++s) {
            if ((s & (s - 1u)) == 0u) {
                cost[s] = 0.f;
                split[s] = 0;
                continue;
            }
            // Partitions (p, s ^ p) with p holding s's lowest leaf, each counted once.
            const uint32_t low = s & (~s + 1u);
            float best = std::numeric_limits<float>::infinity();
            uint32_t bp = 0;
            for (uint32_t p = (s - 1u) & s; p != 0u; p = (p - 1u) & s) {
                if (!(p & low)) continue;
                const float c = cost[p] + cost[s ^ p];
                if (c < best) {
                    best = c;
                    bp = p;
                }
            }
            cost[s] = sa[s] + best;
            split[s] = uint8_t(bp);
        }
        // Current cost: the treelet's inner nodes' areas as they stand.
        float now = 0.f;
        for (int k = 0; k < ni; ++k) now += area(bn[internals[k]].lo, bn[internals[k]].hi);
        if (!(cost[full] < now * (1.f - 1e-6f))) return;
        int next = 1;  // internals[0] stays the treelet root r
        build(full, r, leaves, split, internals, next);
    }
    // Gives node `me` the partition of subset s; returns its box through bn[me].
    void build(uint32_t s, int32_t me, const int32_t* leaves, const uint8_t* split, const int32_t* internals,
               int& next) {
        const uint32_t p = split[s], q = s ^ p;
        const uint32_t parts[2] = {p, q};
        for (int side = 0; side < 2; ++side) {
            const uint32_t t = parts[side];
            int32_t child;
            if ((t & (t - 1u)) == 0u) {
                child = leaves[__builtin_ctz(t)];
            } else {
                child = internals[next++];
                build(t, child, leaves, split, internals, next);
            }
            bn[me].child[side] = child;
        }
        const BNode& a = bn[bn[me].child[0]];
        const BNode& b = bn[bn[me].child[1]];
        for (int k = 0; k < 3; ++k) {
            bn[me].lo[k] = std::min(a.lo[k], b.lo[k]);
            bn[me].hi[k] = std::max(a.hi[k], b.hi[k]);
        }
        bn[me].count = a.count + b.count;
    }
    // Children first, then the node; subtrees larger than the grain as tasks.
    void pass(int32_t r) {
        if (!inner(r)) return;
        if (bn[r].count > kTreeletTaskGrain) {
            const int32_t c0 = bn[r].child[0], c1 = bn[r].child[1];
#pragma omp task default(none) firstprivate(c0)
            pass(c0);
            pass(c1);
#pragma omp taskwait
        } else {
            pass(bn[r].child[0]);
            pass(bn[r].child[1]);
        }
        optimize(r);
    }
};

// Renumbers the binary tree depth first (parents before children, first child
// next) and lays the triangle indices out in that leaf order, so that every node's
// triangles are idx[start, start + count).
void renumber(std::vector<BNode>& bn, uint32_t total, std::vector<uint32_t>& idx) {
    std::vector<BNode> nb;
    nb.reserve(total);
    std::vector<uint32_t> nidx;
    nidx.reserve(idx.size());
    std::vector<std::pair<int32_t, int32_t>> stack;  // (old node, new parent slot), slot = parent * 2 + side
    stack.push_back({0, -1});
    while (!stack.empty()) {
        const auto [old, slot] = stack.back();
        stack.pop_back();
        const int32_t me = int32_t(nb.size());
        nb.push_back(bn[size_t(old)]);
        if (slot >= 0) nb[size_t(slot / 2)].child[slot % 2] = me;
        BNode& n = nb.back();
        if (n.child[0] < 0) {
            const uint32_t start = uint32_t(nidx.size());
            for (uint32_t k = 0; k < n.count; ++k) nidx.push_back(idx[n.start + k]);
            n.start = start;
        } else {
            const int32_t c0 = n.child[0], c1 = n.child[1];
            stack.push_back({c1, me * 2 + 1});
            stack.push_back({c0, me * 2});
        }
    }
    // Inner ranges and depths, top down (children have larger indices now).
    nb[0].depth = 0;
    for (size_t i = 0; i < nb.size(); ++i) {
        BNode& n = nb[i];
        if (n.child[0] < 0) continue;
        nb[size_t(n.child[0])].depth = n.depth + 1;
        nb[size_t(n.child[1])].depth = n.depth + 1;
    }
    for (int64_t i = int64_t(nb.size()) - 1; i >= 0; --i) {
        BNode& n = nb[size_t(i)];
        if (n.child[0] < 0) continue;
        const BNode& a = nb[size_t(n.child[0])];
        n.start = a.start;  // the first child's leaves come first
        n.count = a.count + nb[size_t(n.child[1])].count;
    }
    bn.swap(nb);
    idx.swap(nidx);
}

// Quantizes the children's boxes (lo[3], hi[3]; null = empty slot) against their
// union. Per axis the step is the smallest power of two with 255 steps covering the
// extent; q_lo = floor((lo - origin) / step) and q_hi = ceil((hi - origin) / step),
// computed exactly in double, so fmaf(q, step, origin) (one rounding, monotonic)
// lands at or beyond the exact plane.
void write_node(float* o, const float* const boxes[4], const int32_t words[4]) {
    uint32_t* w = reinterpret_cast<uint32_t*>(o);
    std::memset(o, 0, kNodeFloats * sizeof(float));
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int c = 0; c < 4; ++c)
        if (boxes[c])
            for (int k = 0; k < 3; ++k) {
                lo[k] = std::min(lo[k], boxes[c][k]);
                hi[k] = std::max(hi[k], boxes[c][3 + k]);
            }
    uint32_t exps = 0;
    double step[3];
    for (int k = 0; k < 3; ++k) {
        const double ext = double(hi[k]) - double(lo[k]);
        int e = -126;
        while (e < 127 && std::ldexp(255.0, e) < ext) ++e;
        step[k] = std::ldexp(1.0, e);
        exps |= uint32_t(e + 127) << (8 * k);
        std::memcpy(&w[k], &lo[k], 4);
    }
    w[3] = exps;
    for (int c = 0; c < 4; ++c) std::memcpy(&w[4 + c], &words[c], 4);
    for (int c = 0; c < 4; ++c) {
        if (!boxes[c]) continue;
        for (int k = 0; k < 3; ++k) {
            const double ql = std::floor((double(boxes[c][k]) - double(lo[k])) / step[k]);
            const double qh = std::ceil((double(boxes[c][3 + k]) - double(lo[k])) / step[k]);
            const uint32_t a = uint32_t(std::max(0.0, std::min(255.0, ql)));
            const uint32_t b = uint32_t(std::max(0.0, std::min(255.0, qh)));
            w[8 + 2 * k] |= a << (8 * c);
            w[9 + 2 * k] |= b << (8 * c);
        }
    }
}

// Writes the four-wide layout depth first from the binary build tree. A node's
// inner children get consecutive indices (siblings share cache lines and DRAM
// pages): on the 1080p 1M-triangle frame wf_extend ran 0.7 %, wf_shadow 1.2 % and
// the camera packets 0.9 % faster than with each child placed just before its
// own subtree (kernel traces, two interleaved runs).
struct Emitter {
    const BNode* bn;
    const float (*pbox)[6];
    std::vector<float>& out;
    uint32_t next = 1;
    uint32_t leaves = 0, max_depth = 0, max_leaf = 0;
    double sah = 0.0;
    double root_area = 1.0;

    // SAH-optimal collapse (when set): per build node, dist[j] = least cost of the
    // subtree in at most j slots of its parent, split[j] = how many of them its first
    // child takes (0: keep the node as one slot), as_leaf = the one-slot form is a leaf.
    struct Cost {
        float dist[5];
        int8_t split[5];
        bool as_leaf;
    };
    std::vector<Cost> cost{};

    bool inner(int32_t b) const { return bn[b].child[0] >= 0; }
    double area(int32_t b) const { return double(box_area(pbox[b], pbox[b] + 3)); }
    bool as_leaf(int32_t b) const { return cost.empty() ? !inner(b) : cost[size_t(b)].as_leaf; }

    // Least-cost collapse, bottom-up (children have larger indices than parents).
    void plan(uint32_t total, float tri_cost, uint32_t merge_max) {
        cost.assign(total, Cost{});
        const float inf = std::numeric_limits<float>::infinity();
        for (int64_t i = int64_t(total) - 1; i >= 0; --i) {
            const BNode& n = bn[i];
            Cost& c = cost[size_t(i)];
            const float a = float(area(int32_t(i)) / root_area);
            const bool leafable = !inner(int32_t(i)) || n.count <= merge_max;
            const float leaf = leafable ? a * tri_cost * float(n.count) : inf;
            float node = inf;
            int8_t best = 0;
            if (inner(int32_t(i))) {
                const Cost& l = cost[size_t(n.child[0])];
                const Cost& r = cost[size_t(n.child[1])];
                for (int k = 1; k <= 3; ++k) {
                    const float v = l.dist[k] + r.dist[4 - k];
                    if (v < node) {
                        node = v;
                        best = int8_t(k);
                    }
                }
                node += a;
            }
            // A binary leaf is always a leaf (its cost may be NaN when areas overflow).
            c.as_leaf = !inner(int32_t(i)) || leaf <= node;
            c.dist[1] = std::min(leaf, node);
            c.split[1] = c.as_leaf ? 0 : best;  // the node form's own split of its four slots
            for (int j = 2; j <= 4; ++j) {
                c.dist[j] = c.dist[j - 1];
                c.split[j] = 0;
                if (!inner(int32_t(i))) continue;
                const Cost& l = cost[size_t(n.child[0])];
                const Cost& r = cost[size_t(n.child[1])];
                for (int k = 1; k < j; ++k) {
                    const float v = l.dist[k] + r.dist[j - k];
                    if (v < c.dist[j]) {
                        c.dist[j] = v;
                        c.split[j] = int8_t(k);
                    }
                }
            }
        }
    }
    // The children of b in at most j slots, by the plan.
    void expand(int32_t b, int j, int32_t* slot, int& n) const {
        const Cost& c = cost[size_t(b)];
        while (j >= 2 && c.split[j] == 0) --j;
        if (j <= 1) {
            slot[n++] = b;
            return;
        }
        const int k = c.split[j];
        expand(bn[b].child[0], k, slot, n);
        expand(bn[b].child[1], j - k, slot, n);
    }

    // Up to four build nodes that replace inner build node b: by the plan if there
    // is one, else its children with the largest inner one opened until there are four.
    int gather(int32_t b, int32_t slot[4]) const {
        if (!cost.empty()) {
            int n = 0;
            const int k = std::max<int>(1, cost[size_t(b)].split[1]);
            expand(bn[b].child[0], k, slot, n);
            expand(bn[b].child[1], 4 - k, slot, n);
            return n;
        }
        int n = 2;
        slot[0] = bn[b].child[0];
        slot[1] = bn[b].child[1];
        while (n < 4) {
            int best = -1;
            for (int k = 0; k < n; ++k)
                if (inner(slot[k]) && (best < 0 || area(slot[k]) > area(slot[best]))) best = k;
            if (best < 0) break;
            const int32_t open = slot[best];
            slot[best] = bn[open].child[0];
            slot[n++] = bn[open].child[1];
        }
        return n;
    }
    int32_t leaf_child(int32_t b) {
        const BNode& c = bn[b];
        leaves++;
        max_leaf = std::max(max_leaf, c.count);
        sah += area(b) / root_area * double(c.count);
        return leaf_word(c.start, c.count);
    }
    void emit(int32_t b, uint32_t k, uint32_t depth) {
        max_depth = std::max(max_depth, depth);
        sah += area(b) / root_area;
        int32_t slot[4];
        const int n = gather(b, slot);
        int32_t words[4] = {kEmptyChild, kEmptyChild, kEmptyChild, kEmptyChild};
        for (int c = 0; c < n; ++c) words[c] = as_leaf(slot[c]) ? leaf_child(slot[c]) : int32_t(next++);
        const float* boxes[4] = {nullptr, nullptr, nullptr, nullptr};
        for (int c = 0; c < n; ++c) boxes[c] = pbox[slot[c]];
        write_node(&out[size_t(k) * kNodeFloats], boxes, words);
        for (int c = 0; c < n; ++c)
            if (words[c] >= 0) emit(slot[c], uint32_t(words[c]), depth + 1);
    }
};

}  // namespace

void build_bvh(const float* pos, uint32_t n, const BvhParams& params, BvhLayout& out) {
    if (n == 0) throw std::invalid_argument("BVH of an empty scene");
    if (n >= kMaxRecords) throw std::invalid_argument("more than 2^28 triangles");
    if (params.leaf_size < 1 || params.leaf_size > kMaxLeaf) throw std::invalid_argument("leaf size out of range");
    if (params.merge_max < params.leaf_size || params.merge_max > kMaxLeaf)
        throw std::invalid_argument("merged leaf size out of range");
    std::vector<Prim> prims(n);
    float abs_max = 1.f;
#pragma omp parallel for reduction(max : abs_max) schedule(static)
    for (int64_t i = 0; i < int64_t(n); ++i) {
        const float* p = pos + size_t(i) * 9;
        Prim& q = prims[size_t(i)];
        for (int k = 0; k < 3; ++k) {
            q.lo[k] = std::min(p[k], std::min(p[3 + k], p[6 + k]));
            q.hi[k] = std::max(p[k], std::max(p[3 + k], p[6 + k]));
            q.c[k] = (q.lo[k] + q.hi[k]) * 0.5f;
            abs_max = std::max(abs_max, std::max(std::fabs(q.lo[k]), std::fabs(q.hi[k])));
        }
    }
    std::vector<uint32_t> idx(n);
    for (uint32_t i = 0; i < n; ++i) idx[i] = i;
    std::vector<BNode> bn(size_t(2) * n + 1);
    Ctx c;
    c.prims = &prims;
    c.idx = idx.data();
    c.nodes = bn.data();
    c.p = params;
    c.next = 1;
#pragma omp parallel
#pragma omp single
    build_rec(c, 0, 0, n, 0);
    uint32_t total = c.next.load();
#ifndef PRT_BVH_PAD_ABS
#define PRT_BVH_PAD_ABS 0x1p-16f
#endif
    const float abs_pad = abs_max * PRT_BVH_PAD_ABS;
    if (params.treelet_passes > 0 && bn[0].child[0] >= 0) {
        bn.resize(total);
        TreeletOpt opt{bn.data(), abs_pad};
        for (int pass = 0; pass < params.treelet_passes; ++pass) {
#pragma omp parallel
#pragma omp single
            opt.pass(0);
        }
        renumber(bn, total, idx);
    }

    // Padded boxes, bottom-up: a leaf box grows by pad_rel of its largest extent
    // plus 2^-16 of the largest scene coordinate, an inner box is the union of its
    // children's padded boxes. The padding keeps box rejection conservative
    // with respect to the intersection test.
    std::vector<float> pb(size_t(total) * 6);
    for (int64_t i = int64_t(total) - 1; i >= 0; --i) {
        const BNode& nd = bn[size_t(i)];
        float* b = &pb[size_t(i) * 6];
        if (nd.child[0] < 0) {
            float e = std::max(nd.hi[0] - nd.lo[0], std::max(nd.hi[1] - nd.lo[1], nd.hi[2] - nd.lo[2]));
            float pad = e * params.pad_rel + abs_pad;
            for (int k = 0; k < 3; ++k) {
                b[k] = nd.lo[k] - pad;
                b[3 + k] = nd.hi[k] + pad;
            }
        } else {
            const float* a = &pb[size_t(nd.child[0]) * 6];
            const float* d = &pb[size_t(nd.child[1]) * 6];
            for (int k = 0; k < 3; ++k) {
                b[k] = std::min(a[k], d[k]);
                b[3 + k] = std::max(a[3 + k], d[3 + k]);
            }
        }
    }
    // Children of a build node always have larger indices, so the loop above
    // visits children first.

    uint32_t inner = 0;
    for (uint32_t i = 0; i < total; ++i) inner += bn[i].child[0] >= 0 ? 1u : 0u;
    out = BvhLayout();
    if (inner == 0) {
        // Whole scene in one leaf: the root holds it in slot 0, the other slots empty.
        out.nodes.assign(kNodeFloats, 0.f);
        const float* boxes[4] = {&pb[0], nullptr, nullptr, nullptr};
        const int32_t words[4] = {leaf_word(0, n), kEmptyChild, kEmptyChild, kEmptyChild};
        write_node(out.nodes.data(), boxes, words);
        out.num_nodes = 1;
        out.num_leaves = 1;
        out.max_leaf = n;
        out.sah_cost = float(n);
    } else {
        // At most `inner` four-wide nodes (each consumes at least one inner build node).
        out.nodes.assign(size_t(inner) * kNodeFloats, 0.f);
        Emitter em{bn.data(), reinterpret_cast<const float(*)[6]>(pb.data()), out.nodes};
        em.root_area = std::max(1e-30, double(box_area(&pb[0], &pb[3])));
        if (params.collapse_sah) em.plan(total, params.tri_cost, uint32_t(params.merge_max));
        em.emit(0, 0, 0);
        out.nodes.resize(size_t(em.next) * kNodeFloats);
        out.num_nodes = em.next;
        out.num_leaves = em.leaves;
        out.max_depth = em.max_depth;
        out.max_leaf = em.max_leaf;
        out.sah_cost = float(em.sah);
    }
    out.order = std::move(idx);
}

}  // namespace prt
