// Persistent wavefront path tracer with batching over views and scenes.
//
// Work is cut into chunks of at most `queue_slots` pixel-samples. A chunk is a
// list of tiles (job, pixel range, sample range); slots are numbered tile by tile.
// One chunk runs as a captured HIP graph:
//
//   setup -> generate -> [extend -> shade -> shadow] x max_depth -> resolve -> advance
//
//   generate camera rays (origin, direction) of the chunk's pixel-samples, queued
//            in groups of 64 that share a pixel (or a small pixel block)
//   extend   closest hit of every ray in the current queue (traversal only, low
//            register pressure, persistent waves fetching 64 rays at a time; the
//            first iteration's camera rays are traced 64 at a time as packets)
//   shade    shade_vertex() of every hit; survivors are compacted into the next
//            queue and shadow rays into the shadow queue (wave ballots + one atomic
//            per workgroup and queue), finished paths write their colour to their
//            result slot (row-major per sample within the tile)
//   shadow   occlusion test of every shadow ray; adds mask * direct
//   resolve  adds the result slots of each pixel to its running sum in sample order
//
// All paths of a chunk are at bounce `it` during iteration `it`, and the vertex
// arithmetic is the megakernel's (path.hpp), so every image is bit-identical to
// the megakernel's for the same seed, view and depth. The graph is replayed once
// per chunk; kernels find their chunk through a device-side index, and every
// counter is cleared by a kernel (no memset nodes in the graph).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../host/common.hpp"
#include "path.hpp"
#include "renderer.hpp"

namespace prt {

constexpr int kMaxIter = 64;
// Same-address atomics serialise (about 11 ns each on MI355X), so work is handed
// out by kSpread counters, one kBatch-item batch per atomic, and appends to a queue
// are aggregated per workgroup.
// Fetch counters per (kernel, iteration), items per fetch atomic, and shade
// workgroup size. With every counter on its own line, 1080p x 128 spp measured
// within noise of 64 / 64 / 512 for spread 128, batch 128 and 1024-thread shade
// workgroups; spread 32 and 256-thread shade workgroups were 2-4 % slower.
#ifndef PRT_FETCH_SPREAD
#define PRT_FETCH_SPREAD 64
#endif
#ifndef PRT_FETCH_BATCH
#define PRT_FETCH_BATCH 64
#endif
#ifndef PRT_SHADE_BLOCK
#define PRT_SHADE_BLOCK 512
#endif
constexpr uint32_t kSpread = PRT_FETCH_SPREAD;
constexpr uint32_t kBatch = PRT_FETCH_BATCH;  // items per fetch atomic
constexpr int kShadeBlock = PRT_SHADE_BLOCK;
// Waves of wf_shade per compute unit in its grid (its occupancy limit, 4 waves per SIMD).
#ifndef PRT_SHADE_GRID_WAVES
#define PRT_SHADE_GRID_WAVES 16
#endif
constexpr int kShadeGridWaves = PRT_SHADE_GRID_WAVES;
// Bytes of queues and path state per slot: 2 queues x 64, hit 16 + 4, shadow 64, result 16.
constexpr size_t kSlotBytes = 2 * 64 + 16 + 4 + 64 + 16;
// Counter words per 128-byte cache line. Each queue counter and each fetch counter
// has a line of its own: atomics on one line serialize in its L2 channel, and the
// appends of wf_shade and the batch fetches of the traversal kernels would otherwise
// share a few lines.
#ifndef PRT_CTR_LINE
#define PRT_CTR_LINE 32
#endif
constexpr uint32_t kCtrLine = PRT_CTR_LINE;
enum WfCounter : uint32_t {
    kExt0 = 0,                 // ray count of queue 0
    kExt1 = kCtrLine,          // ray count of queue 1
    kShadow = 2 * kCtrLine,    // shadow ray count
    kChunk = 3 * kCtrLine,     // index of the chunk being rendered
    kFetch = 8 * kCtrLine,     // per (kernel, iteration) fetch counters: 3 * kMaxIter groups of kSpread,
                               // counter c of group g at kFetch + (g * kSpread + c) * kCtrLine
    kStats = kFetch + 3 * kMaxIter * kSpread * kCtrLine,  // kCntNum 64-bit counters (two words each)
    kNumCounters = kStats + 2 * kCntNum,
};
// First fetch counter of group g (3 * iteration + kernel: 0 extend, 2 shadow).
constexpr uint32_t fetch_group(uint32_t g) { return kFetch + g * kSpread * kCtrLine; }

struct WfJob {
    SceneArgs sc;
    CamArgs cam;
    PathArgs path;
    uint32_t width, height;
    float4* accum;
    float4* samples;       // nullptr, or each sample's radiance at [sample * width * height + pixel]
    uint32_t samples_cap;  // samples stored per pixel
};

struct WfTile {
    uint32_t job, p0, p1, s0, s1;
    uint32_t slot0;  // first slot of the tile in its chunk
    uint32_t pix0;   // first resolve thread of the tile in its chunk
    uint32_t pad;
};

struct WfChunk {
    uint32_t tile0, ntiles, paths, pixels;
};

// Queue entry (structure of arrays, one float4 each, coalesced by queue index):
//   A  origin.xyz, pixel | job << 24
//   B  direction.xyz, sample index
//   C  throughput.xyz, previous bsdf
//   D  colour.xyz, slot (result index)
// Shadow entry: E position.xyz, target | F normal.xyz, lu | G mask.xyz, light record |
//   H lv, final flag. `target` is the path's entry in the next queue, or its result
//   slot when the path ended at this vertex.
struct WfArgs {
    const WfJob* jobs;
    const WfTile* tiles;
    const WfChunk* chunks;
    uint32_t* ctr;
    float4* qa[2];
    float4* qb[2];
    float4* qc[2];
    float4* qd[2];
    float4* hits;    // t, u, v, record (0xFFFFFFFF = miss)
    uint32_t* hmat;  // material of the hit record
    float4* se;
    float4* sf;
    float4* sg;
    float4* sh;
    float4* result;  // per slot final colour
    int* spill;
    uint32_t spill_stride;
    uint32_t max_depth;
    uint32_t group_lg;  // log2 of the forced camera-path block side, or kNone: choose per tile
};

struct WavefrontState {
    int device = -1;
    uint32_t capacity = 0;       // slots
    uint32_t persistent = 0;     // threads of the persistent kernels (spill stride)
    void* mem = nullptr;
    size_t bytes = 0;
    int* spill = nullptr;
    uint32_t* ctr = nullptr;
    WfArgs args{};
    // plan buffers
    WfJob* d_jobs = nullptr;
    WfTile* d_tiles = nullptr;
    WfChunk* d_chunks = nullptr;
    size_t cap_jobs = 0, cap_tiles = 0, cap_chunks = 0;
};

namespace dev {

// Hands out 64-item batches of [0, n). Batch b belongs to counter b % kSpread. A
// wave draws from counter (wave % kSpread); issue() sends the atomic for its next
// batch before the current batch is processed, so the round trip overlaps the
// traversal. When its counter is exhausted the wave probes the others, with a
// coherent load first so exhausted counters cost no atomic. resolve() returns the
// batch's first item or 0xFFFFFFFF when every batch has been handed out.
struct SpreadFetch {
    uint32_t* ctr;
    uint32_t k, nb;
    __device__ __forceinline__ uint32_t limit(uint32_t c) const { return (nb + kSpread - 1 - c) / kSpread; }
    __device__ __forceinline__ uint32_t issue() const {
        uint32_t idx = 0xFFFFFFFFu;
        if ((threadIdx.x & 63) == 0) idx = atomicAdd(ctr + k * kCtrLine, 1u);
        return idx;
    }
    __device__ __forceinline__ uint32_t resolve(uint32_t raw) {
        uint32_t idx = __shfl(raw, 0, 64);
        if (idx < limit(k)) return (idx * kSpread + k) * kBatch;
        for (uint32_t probe = 1; probe < kSpread; ++probe) {
            k = (k + 1) % kSpread;
            idx = 0xFFFFFFFFu;
            if ((threadIdx.x & 63) == 0) {
                const uint32_t lim = limit(k);
                const uint32_t seen = __hip_atomic_load(ctr + k * kCtrLine, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (seen < lim) idx = atomicAdd(ctr + k * kCtrLine, 1u);
                if (idx >= lim) idx = 0xFFFFFFFFu;
            }
            idx = __shfl(idx, 0, 64);
            if (idx != 0xFFFFFFFFu) return (idx * kSpread + k) * kBatch;
        }
        return 0xFFFFFFFFu;
    }
};
__device__ __forceinline__ SpreadFetch make_fetch(uint32_t* ctr, uint32_t n) {
    const uint32_t wave = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    return SpreadFetch{ctr, wave % kSpread, (n + kBatch - 1) / kBatch};
}

constexpr uint32_t kNone = 0xFFFFFFFFu;

// Hands single items of [0, n) to individual lanes: the wave holds one 64-item
// batch of a SpreadFetch at a time (one atomic per 64 items, the next one already
// in flight) and gives its items to whichever lanes ask, so a lane whose ray has
// finished takes a new one while the others keep traversing. Every issued atomic
// is resolved before more() turns false. Calls must be wave-uniform.
struct LaneFeeder {
    SpreadFetch fetch;
    uint32_t n, base, used, pending;
    __device__ __forceinline__ void start(uint32_t* ctr, uint32_t count) {
        fetch = make_fetch(ctr, count);
        n = count;
        base = n ? fetch.resolve(fetch.issue()) : kNone;
        used = 0;
        pending = base != kNone ? fetch.issue() : kNone;
    }
    __device__ __forceinline__ bool more() const { return base != kNone; }
    // Lanes with `want` get the next item, or kNone once every item is handed out.
    __device__ __forceinline__ uint32_t take(bool want) {
        const uint32_t lane = threadIdx.x & 63;
        const unsigned long long below = (1ull << lane) - 1ull;
        uint32_t item = kNone;
        unsigned long long m = __ballot(want);
        while (m != 0ull && base != kNone) {
            const uint32_t size = min(kBatch, n - base);
            const uint32_t avail = size - used;
            const uint32_t rank = uint32_t(__popcll(m & below));
            if (((m >> lane) & 1ull) && rank < avail) item = base + used + rank;
            used += min(uint32_t(__popcll(m)), avail);
            if (used == size) {
                base = fetch.resolve(pending);
                used = 0;
                pending = base != kNone ? fetch.issue() : kNone;
            }
            m = __ballot(want && item == kNone);
        }
        return item;
    }
};

// Idle lanes of a wave that trigger a refill (the rest keep traversing meanwhile).
// kRefill <= 64, so a wave with no active lane always refills or, with no work
// left, leaves the loop. The finished rays' writes and the new rays' set-up run at
// that point with these lanes only; on 1080p x 128 spp over 1M triangles, 32 took
// 125.7 ms per frame, 16 129.0, 40 128.3 and 48 139.5.
#ifndef PRT_REFILL_LANES
#define PRT_REFILL_LANES 32
#endif
constexpr uint32_t kRefill = PRT_REFILL_LANES;
static_assert(kRefill >= 1 && kRefill <= 64, "refill threshold out of range");

__device__ __forceinline__ uint32_t idle_lanes(bool active) { return uint32_t(__popcll(__ballot(!active))); }

// A pointer that every lane of the wave holds, moved to scalar registers. Read from
// the job table with vector loads (the compiler cannot prove that memory read-only),
// the scene's node and triangle bases would otherwise take four VGPRs and a 64-bit
// add per fetch; from SGPRs the fetches use the scalar base + vector offset form.
template <class T>
__device__ __forceinline__ const T* wave_uniform(const T* p) {
    const uint64_t v = reinterpret_cast<uint64_t>(p);
    const uint32_t lo = uint32_t(__builtin_amdgcn_readfirstlane(int(uint32_t(v))));
    const uint32_t hi = uint32_t(__builtin_amdgcn_readfirstlane(int(uint32_t(v >> 32))));
    return reinterpret_cast<const T*>((uint64_t(hi) << 32) | lo);
}

// Phase thresholds of the two traversal kernels (node_phase_for). On 1080p x 128
// spp over 1M triangles, 24 measured best for both (extend 109 -> 92 ms, shadow
// 54 -> 43 ms per frame against no phases; 8 and 32 to 48 were slower).
#ifndef PRT_EXTEND_PHASE_MIN
#define PRT_EXTEND_PHASE_MIN 24
#endif
#ifndef PRT_SHADOW_PHASE_MIN
#define PRT_SHADOW_PHASE_MIN 24
#endif
static_assert(PRT_EXTEND_PHASE_MIN >= 0 && PRT_EXTEND_PHASE_MIN <= 64, "phase threshold out of range");
static_assert(PRT_SHADOW_PHASE_MIN >= 0 && PRT_SHADOW_PHASE_MIN <= 64, "phase threshold out of range");

// Waves per SIMD of the traversal kernels. With the scene bases in SGPRs,
// wf_extend fits 8 (63 VGPRs; the one spill is a constant reloaded only for
// zero-direction rays) and wf_shadow 7 (72 VGPRs; its spill is reloaded only on the
// stack's overflow path). 1080p x 128 spp over 1M triangles: 153.7 -> 151.6 ms per
// frame against 7 and 6; before the SGPR bases, forced occupancy spilled in the loop.
#ifndef PRT_SHADOW_WAVES
#define PRT_SHADOW_WAVES __attribute__((amdgpu_waves_per_eu(7)))
#endif
#ifndef PRT_EXTEND_WAVES
#define PRT_EXTEND_WAVES __attribute__((amdgpu_waves_per_eu(8)))
#endif
// wf_shade keeps the compiler's occupancy (104 VGPRs, 4 waves per SIMD): forcing 5
// or 6 waves spills 36 / 104 bytes per lane and measured 2% / 12% slower per frame;
// with 256-thread workgroups at 5 waves (24 bytes spilled) and 128-thread ones at 6,
// each with a grid of 20 / 24 waves per compute unit, 4% and 25% slower.
#ifndef PRT_SHADE_WAVES
#define PRT_SHADE_WAVES
#endif

// Adds the wave's counts to the counter slots; ANYHIT also books nodes and
// triangles as the shadow (any-hit) kernel's share.
template <bool ANYHIT>
__device__ __forceinline__ void flush_stats(uint32_t* ctr, const Counts& c) {
    unsigned long long v[6] = {c.primary, c.extension, c.shadow, c.nodes, c.tris, c.hits};
    unsigned long long* out = reinterpret_cast<unsigned long long*>(ctr + kStats);
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        unsigned long long x = v[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        if ((threadIdx.x & 63) == 0 && x) {
            atomicAdd(out + k, x);
            if (ANYHIT && k == kCntNodes) atomicAdd(out + kCntAnyNodes, x);
            if (ANYHIT && k == kCntTris) atomicAdd(out + kCntAnyTris, x);
        }
    }
}

__device__ __forceinline__ uint32_t find_tile(const WfTile* t, uint32_t first, uint32_t n, uint32_t key, bool by_pix) {
    uint32_t lo = first, hi = first + n - 1;
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1) / 2;
        uint32_t v = by_pix ? t[mid].pix0 : t[mid].slot0;
        if (v <= key) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ void wf_setup(WfArgs a) {
    for (uint32_t k = threadIdx.x; k < 3 * kMaxIter * kSpread; k += blockDim.x) a.ctr[kFetch + k * kCtrLine] = 0;
    if (threadIdx.x != 0) return;
    const uint32_t c = a.ctr[kChunk];
    a.ctr[kExt0] = a.chunks[c].paths;
    a.ctr[kExt1] = 0;
    a.ctr[kShadow] = 0;
}

__global__ void wf_advance(WfArgs a) {
    if (threadIdx.x == 0) a.ctr[kChunk] = a.ctr[kChunk] + 1;
}

// Path order within a tile. Where a tile covers whole rows of its image, each run
// of 64 paths (one wave's batch) is a g x g pixel block times 64 / g^2 consecutive
// samples of those pixels: blocks band by band and left to right, sample groups
// outermost. g is the smallest of 1, 2, 4, 8 that divides the tile (its sample count
// by 64 / g^2, its width and rows by g), so at 128 spp a wave traces 64 samples of
// one pixel: nearly the same primary ray in every lane, then shadow and extension
// rays from nearly the same point. Other tiles keep the row-major order per sample.
// On 1080p x 128 spp over 1M triangles, one pixel x 64 samples took 147.3 ms per
// frame, 4 x 4 x 4 148.1, 8 x 8 x 1 148.6 and row order 151.7. Only the queue is
// reordered: a path's result slot (carried in its queue entry) stays row-major per
// sample, so wf_resolve's loads stay coalesced across adjacent pixels, and every
// pixel still adds its samples in sample order.
struct PathGroup {
    uint32_t lg;  // log2 of the block side g; kNone: row order
};
// force_lg (not kNone) asks for g = 2^force_lg; a tile it does not divide takes the
// smallest g that does.
__device__ __forceinline__ PathGroup tile_group(const WfTile& t, uint32_t width, uint32_t force_lg) {
    const uint32_t w = t.p1 - t.p0, ns = t.s1 - t.s0;
    if (t.p0 % width != 0u || w % width != 0u) return PathGroup{kNone};
    const uint32_t rows = w / width;
    auto fits = [&](uint32_t lg) {
        const uint32_t g = 1u << lg, gs = 64u >> (2u * lg);
        return ns % gs == 0u && width % g == 0u && rows % g == 0u;
    };
    if (force_lg <= 3u && fits(force_lg)) return PathGroup{force_lg};
    for (uint32_t lg = 0; lg <= 3; ++lg)
        if (fits(lg)) return PathGroup{lg};
    return PathGroup{kNone};
}
// Position r of a grouped tile of w pixels -> (pixel offset from p0, sample offset from s0).
__device__ __forceinline__ uint2 group_to_linear(uint32_t r, uint32_t w, uint32_t width, PathGroup pg) {
    const uint32_t g = 1u << pg.lg, gs = 64u >> (2u * pg.lg);
    const uint32_t per_sg = w * gs, sg = r / per_sg, rem = r % per_sg;
    const uint32_t blk = rem >> 6, lane = rem & 63u, per_band = width >> pg.lg;
    const uint32_t x = ((blk % per_band) << pg.lg) + (lane & (g - 1u));
    const uint32_t y = ((blk / per_band) << pg.lg) + ((lane >> pg.lg) & (g - 1u));
    return make_uint2(y * width + x, sg * gs + (lane >> (2u * pg.lg)));
}

__global__ __launch_bounds__(256) void wf_generate(WfArgs a) {
    const WfChunk ch = a.chunks[a.ctr[kChunk]];
    for (uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x; slot < ch.paths; slot += gridDim.x * blockDim.x) {
        const WfTile t = a.tiles[find_tile(a.tiles, ch.tile0, ch.ntiles, slot, false)];
        const WfJob& j = a.jobs[t.job];
        const uint32_t w = t.p1 - t.p0;
        const uint32_t r = slot - t.slot0;
        const PathGroup pg = tile_group(t, j.width, a.group_lg);
        const uint2 ps = pg.lg != kNone ? group_to_linear(r, w, j.width, pg) : make_uint2(r % w, r / w);
        const uint32_t pixel = t.p0 + ps.x, sample = t.s0 + ps.y;
        const int x = int(pixel % j.width), y = int(pixel / j.width);
        Rng rng{j.path.seed, j.path.view, pixel, sample};
        V3 ro, rd;
        camera_ray(j.cam, x, y, rng.block(0), ro, rd);
        // C (throughput 1, previous bsdf 0) and D (colour 0, result slot) of a new path
        // are not written: wf_shade(0) makes them from the tile (first_bounce_cd).
        a.qa[0][slot] = make_float4(ro.x, ro.y, ro.z, __uint_as_float(pixel | (t.job << 24)));
        a.qb[0][slot] = make_float4(rd.x, rd.y, rd.z, __uint_as_float(sample));
    }
}

// The C and D entries of a new path (queue position i in the first extend queue of
// chunk ch): throughput 1, previous bsdf 0, colour 0, and the path's result slot,
// row-major per sample within its tile.
__device__ __forceinline__ void first_bounce_cd(const WfArgs& a, const WfChunk& ch, uint32_t i, uint32_t pixel,
                                                uint32_t sample, float4& C, float4& D) {
    const WfTile t = a.tiles[find_tile(a.tiles, ch.tile0, ch.ntiles, i, false)];
    C = make_float4(1.f, 1.f, 1.f, __int_as_float(0));
    D = make_float4(0.f, 0.f, 0.f, __uint_as_float(t.slot0 + (sample - t.s0) * (t.p1 - t.p0) + (pixel - t.p0)));
}

// Runs f(job) with the job index made wave-uniform (scalar loads of the job
// descriptor) when every valid lane of the wave belongs to the same job.
template <class F>
__device__ __forceinline__ void with_job(bool valid, uint32_t job, F&& f) {
    const uint32_t j0 = __builtin_amdgcn_readfirstlane(job);
    if (__ballot(valid && job != j0) == 0ull) {
        if (valid) f(j0);
    } else {
        if (valid) f(job);
    }
}

// ONE_SCENE: every job of the batch traverses the same scene, so the node and
// triangle arrays are wave-uniform (scalar registers).
template <bool COUNT, bool ONE_SCENE>
__global__ __launch_bounds__(kBlock) PRT_EXTEND_WAVES void wf_extend(WfArgs a, uint32_t it) {
    __shared__ int s_stack[kStackLdsQueue * kBlock];
    const uint32_t q = it & 1u;
    const uint32_t n = a.ctr[kExt0 + q];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.ctr[kExt0 + (q ^ 1u)] = 0;  // the queue shade(it) fills
        a.ctr[kShadow] = 0;
    }
    StackT<kStackLdsQueue> st;
    st.bind(s_stack, a.spill + size_t(blockIdx.x) * kBlock, a.spill_stride);
    Counts cnt = {0, 0, 0, 0, 0, 0};
    LaneFeeder feed;
    feed.start(&a.ctr[fetch_group(3 * it)], n);
    bool active = false;
    // A lane whose ray has finished keeps its hit in registers until the next refill
    // point, where every such lane writes it at once: hit_details() (a triangle load
    // and a full intersection) then runs with at least kRefill lanes instead of the
    // few that finish in any one step.
    bool finished = false;
    uint32_t i = 0;
    TravRay ray;
    ray.leaf = kNoChild;
    HitRec h;
    const float4* nodes = ONE_SCENE ? wave_uniform(a.jobs[0].sc.nodes) : nullptr;
    const float4* tris = ONE_SCENE ? wave_uniform(a.jobs[0].sc.tris) : nullptr;
    while (true) {
        const uint32_t idle = idle_lanes(active);
        if (idle >= kRefill) {
            if (finished) {
                if (h.rec != 0xFFFFFFFFu) hit_details(tris, ray, h);
                a.hits[i] = make_float4(h.t, h.u, h.v, __uint_as_float(h.rec));
                a.hmat[i] = h.mat;
                finished = false;
            }
            if (!feed.more()) {
                if (idle == 64) break;
            } else if (const uint32_t got = feed.take(!active); got != kNone) {
                i = got;
                const float4 o = a.qa[q][i], d = a.qb[q][i];
                if (!ONE_SCENE) {
                    const SceneArgs& sc = a.jobs[__float_as_uint(o.w) >> 24].sc;
                    nodes = sc.nodes;
                    tris = sc.tris;
                }
                if (COUNT) {
                    if (it == 0) cnt.primary++; else cnt.extension++;
                }
                h.t = kFltMax;
                h.id = 0;
                h.rec = 0xFFFFFFFFu;
                h.mat = 0;
                h.u = h.v = 0.f;
                st.sp = 0;
                active = trav_begin(ray, v3(o), v3(d));
                if (!active) {  // zero direction: no triangle can be hit
                    a.hits[i] = make_float4(h.t, h.u, h.v, __uint_as_float(0xFFFFFFFFu));
                    a.hmat[i] = 0;
                }
            }
        }
        // One node phase or leaf phase per step; a refill is considered every step.
        const bool node_phase = node_phase_for<PRT_EXTEND_PHASE_MIN>(active && trav_wants_nodes(ray), active);
        if (active) {
            if (trav_step<false, COUNT, true>(node_phase, nodes, tris, ray, kEps, h, st, cnt) != kTravRunning) {
                active = false;
                finished = true;
            }
        }
    }
    if (COUNT) flush_stats<false>(a.ctr, cnt);
}

// Runs f(lanes, job) once per distinct job among the wave's `live` lanes, with
// `lanes` true for the lanes of that job and the job index wave-uniform: a packet
// traversal needs one scene per pass. With ONE_SCENE there is one pass.
template <bool ONE_SCENE, class F>
__device__ __forceinline__ void per_scene_pass(bool live, uint32_t job, F&& f) {
    if (ONE_SCENE) {
        f(live, 0u);
        return;
    }
    bool todo = live;
    unsigned long long m = __ballot(todo);
    while (m != 0ull) {
        const uint32_t jl = uint32_t(__builtin_amdgcn_readlane(int(job), __builtin_ctzll(m)));
        const bool mine = todo && job == jl;
        f(mine, jl);
        todo = todo && !mine;
        m = __ballot(todo);
    }
}

// wf_extend for the first iteration, whose queue holds the camera rays in 64-ray
// batches of 64 samples of a pixel or pixel block (wf_generate): nearly the same
// ray in every lane. Each wave takes one batch at a time and traces it as a packet
// (packet_traverse). On the 1080p 128-spp frame over 1M triangles this launch took
// 15.8 ms with per-lane traversal and 9.9 ms as packets. (Packets measured 5x
// slower for the first iteration's occlusion rays, whose batches fan out over the
// lights, and later iterations' rays leave diffuse bounces in all directions.)
static_assert(kBatch == 64 && kBlock == 64, "a packet is one fetch batch traced by one wave");
template <bool COUNT, bool ONE_SCENE>
__global__ __launch_bounds__(kBlock) PRT_EXTEND_WAVES void wf_extend_packet(WfArgs a, uint32_t it) {
    __shared__ int s_pk[kPacketStack];
    const uint32_t q = it & 1u;
    const uint32_t n = a.ctr[kExt0 + q];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.ctr[kExt0 + (q ^ 1u)] = 0;  // the queue shade(it) fills
        a.ctr[kShadow] = 0;
    }
    Counts cnt = {0, 0, 0, 0, 0, 0};
    const uint32_t lane = threadIdx.x & 63u;
    SpreadFetch fetch = make_fetch(&a.ctr[fetch_group(3 * it)], n);
    uint32_t base = n ? fetch.resolve(fetch.issue()) : kNone;
    uint32_t pending = base != kNone ? fetch.issue() : kNone;
    while (base != kNone) {
        const uint32_t i = base + lane;
        const bool valid = i < n;
        float4 o = make_float4(0.f, 0.f, 0.f, 0.f), d = o;
        if (valid) {
            o = a.qa[q][i];
            d = a.qb[q][i];
        }
        if (COUNT && valid) {
            if (it == 0) cnt.primary++; else cnt.extension++;
        }
        TravRay ray;
        const bool began = trav_begin(ray, v3(o), v3(d));  // false for the zero ray of an empty lane
        const bool live = valid && began;
        HitRec h;
        h.t = kFltMax;
        h.id = 0;
        h.rec = 0xFFFFFFFFu;
        h.mat = 0;
        h.u = h.v = 0.f;
        per_scene_pass<ONE_SCENE>(live, __float_as_uint(o.w) >> 24, [&](bool mine, uint32_t job) {
            const float4* nodes = wave_uniform(a.jobs[job].sc.nodes);
            const float4* tris = wave_uniform(a.jobs[job].sc.tris);
            bool occ;
            packet_traverse<false, COUNT>(nodes, tris, mine, ray, kEps, h, (lds_int*)s_pk, cnt, occ);
        });
        if (valid) {
            a.hits[i] = make_float4(h.t, h.u, h.v, __uint_as_float(h.rec));
            a.hmat[i] = h.mat;
        }
        base = fetch.resolve(pending);
        pending = base != kNone ? fetch.issue() : kNone;
    }
    if (COUNT) flush_stats<false>(a.ctr, cnt);
}

// Appends to two queues for a whole workgroup with one atomic per queue: lanes
// with more / nee get consecutive positions. Returns (next-queue, shadow) positions.
// (Grouping a workgroup's shadow rays by the light they aim at, so that 64-ray
// batches mostly share a light, measured 2% slower per frame on the 1M-triangle
// scene with its two lights 10 units apart: the queue order of shade, pixel by
// pixel, is the better one for the L2.)
__device__ __forceinline__ uint2 block_append(bool more, bool nee, uint32_t* ctr_next, uint32_t* ctr_shadow) {
    __shared__ uint32_t s_off[2][kShadeBlock / 64];
    __shared__ uint32_t s_base[2];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x / 64;
    const unsigned long long m0 = __ballot(more), m1 = __ballot(nee);
    if (lane == 0) {
        s_off[0][wave] = uint32_t(__popcll(m0));
        s_off[1][wave] = uint32_t(__popcll(m1));
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t0 = 0, t1 = 0;
        for (int w = 0; w < kShadeBlock / 64; ++w) {
            uint32_t c0 = s_off[0][w], c1 = s_off[1][w];
            s_off[0][w] = t0;
            s_off[1][w] = t1;
            t0 += c0;
            t1 += c1;
        }
        s_base[0] = t0 ? atomicAdd(ctr_next, t0) : 0u;
        s_base[1] = t1 ? atomicAdd(ctr_shadow, t1) : 0u;
    }
    __syncthreads();
    const unsigned long long lt = (1ull << lane) - 1ull;
    uint2 r;
    r.x = s_base[0] + s_off[0][wave] + uint32_t(__popcll(m0 & lt));
    r.y = s_base[1] + s_off[1][wave] + uint32_t(__popcll(m1 & lt));
    __syncthreads();  // s_off / s_base are reused by the next iteration
    return r;
}

template <bool COUNT>
__global__ __launch_bounds__(kShadeBlock) PRT_SHADE_WAVES void wf_shade(WfArgs a, uint32_t it) {
    const uint32_t q = it & 1u;
    const uint32_t n = a.ctr[kExt0 + q];
    // Read once, before this kernel's queue atomics start on the counters' cache line.
    const uint32_t chunk = __builtin_amdgcn_readfirstlane(a.ctr[kChunk]);
    const WfChunk ch = a.chunks[chunk];
    Counts cnt = {0, 0, 0, 0, 0, 0};
    // Static, workgroup-uniform partition of the queue (shading cost varies little).
    for (uint32_t base = blockIdx.x * kShadeBlock; base < n; base += gridDim.x * kShadeBlock) {
        const uint32_t i = base + threadIdx.x;
        const bool valid = i < n;
        bool more = false, want_nee = false;
        PathState ps;
        NeeSample ns;
        float4 A = make_float4(0.f, 0.f, 0.f, 0.f), B = A, Cc = A, D = A, hh = A;
        uint32_t hm = 0;
        if (valid) {
            A = a.qa[q][i];
            B = a.qb[q][i];
            if (it == 0) {
                first_bounce_cd(a, ch, i, __float_as_uint(A.w) & 0xFFFFFFu, __float_as_uint(B.w), Cc, D);
            } else {
                Cc = a.qc[q][i];
                D = a.qd[q][i];
            }
            hh = a.hits[i];
            hm = a.hmat[i];
        }
        const uint32_t pj = __float_as_uint(A.w);
        const uint32_t slot = __float_as_uint(D.w);
        ps.ro = v3(A);
        ps.rd = v3(B);
        ps.mask = v3(Cc);
        ps.color = v3(D);
        ps.prev = __float_as_int(Cc.w);
        ps.b = it;
        ps.input = it == 0 ? v3(0.f, 0.f, 0.f) : ps.rd;  // at b >= 1 the stale input is the ray direction
        with_job(valid, pj >> 24, [&](uint32_t jb) {
            const WfJob& j = a.jobs[jb];
            const uint32_t rec = __float_as_uint(hh.w);
            if (rec == 0xFFFFFFFFu) {
                path_miss(ps);
            } else {
                HitRec h;
                h.t = hh.x;
                h.u = hh.y;
                h.v = hh.z;
                h.rec = rec;
                h.id = 0;
                h.mat = hm;
                Rng rng{j.path.seed, j.path.view, pj & 0xFFFFFFu, __float_as_uint(B.w)};
                more = shade_vertex<COUNT, true>(j.sc, j.path, rng, ps, h, &want_nee, ns, cnt);
            }
        });
        if (valid && !more) a.result[slot] = make_float4(ps.color.x, ps.color.y, ps.color.z, 0.f);
        const uint2 pos = block_append(more, want_nee, &a.ctr[kExt0 + (q ^ 1u)], &a.ctr[kShadow]);
        const uint32_t qi = pos.x;
        if (more) {
            a.qa[q ^ 1u][qi] = make_float4(ps.ro.x, ps.ro.y, ps.ro.z, A.w);
            a.qb[q ^ 1u][qi] = make_float4(ps.rd.x, ps.rd.y, ps.rd.z, B.w);
            a.qc[q ^ 1u][qi] = make_float4(ps.mask.x, ps.mask.y, ps.mask.z, __int_as_float(ps.prev));
            a.qd[q ^ 1u][qi] = make_float4(ps.color.x, ps.color.y, ps.color.z, D.w);
        }
        const uint32_t si = pos.y;
        if (want_nee) {
            const uint32_t target = more ? qi : slot;
            a.se[si] = make_float4(ns.position.x, ns.position.y, ns.position.z, __uint_as_float(target));
            a.sf[si] = make_float4(ns.normal.x, ns.normal.y, ns.normal.z, ns.lu);
            a.sg[si] = make_float4(ns.mask.x, ns.mask.y, ns.mask.z, __uint_as_float(ns.li));
            a.sh[si] = make_float4(ns.lv, __uint_as_float(more ? 0u : 1u), __uint_as_float(pj >> 24), 0.f);
        }
    }
    if (COUNT) flush_stats<false>(a.ctr, cnt);
}

template <bool COUNT, bool ONE_SCENE>
__global__ __launch_bounds__(kBlock) PRT_SHADOW_WAVES void wf_shadow(WfArgs a, uint32_t it) {
    __shared__ int s_stack[kStackLdsQueue * kBlock];
    const uint32_t q1 = (it & 1u) ^ 1u;  // the queue shade(it) filled
    const uint32_t n = a.ctr[kShadow];
    StackT<kStackLdsQueue> st;
    st.bind(s_stack, a.spill + size_t(blockIdx.x) * kBlock, a.spill_stride);
    Counts cnt = {0, 0, 0, 0, 0, 0};
    LaneFeeder feed;
    feed.start(&a.ctr[fetch_group(3 * it + 2)], n);
    bool active = false;
    TravRay ray;
    ray.leaf = kNoChild;
    HitRec occ;
    const float4* nodes = ONE_SCENE ? wave_uniform(a.jobs[0].sc.nodes) : nullptr;
    const float4* tris = ONE_SCENE ? wave_uniform(a.jobs[0].sc.tris) : nullptr;
    float4* dst = nullptr;  // the path's colour: next-queue entry or result slot
    V3 mask = v3(0.f, 0.f, 0.f), direct = mask;
    // As in wf_extend: a finished ray's colour update waits for the next refill point.
    bool finished = false, lit = false;
    while (true) {
        const uint32_t idle = idle_lanes(active);
        if (idle >= kRefill) {
            if (finished) {
                const float4 c = *dst;
                const V3 s = add_direct(v3(c), mask, direct, lit);
                *dst = make_float4(s.x, s.y, s.z, c.w);
                finished = false;
            }
            if (!feed.more()) {
                if (idle == 64) break;
            } else if (const uint32_t got = feed.take(!active); got != kNone) {
                const float4 e = a.se[got], f = a.sf[got], g = a.sg[got], hx = a.sh[got];
                const SceneArgs& sc = a.jobs[__float_as_uint(hx.z)].sc;
                NeeSample ns;
                ns.position = v3(e);
                ns.normal = v3(f);
                ns.mask = v3(g);
                ns.li = __float_as_uint(g.w);
                ns.lu = f.w;
                ns.lv = hx.x;
                ShadowReq sr;
                sr.direct = v3(0.f, 0.f, 0.f);
                const bool reach = nee_prepare<COUNT>(sc, ns, sr, cnt);
                const uint32_t target = __float_as_uint(e.w);
                dst = __float_as_uint(hx.y) ? a.result + target : a.qd[q1] + target;
                mask = ns.mask;
                direct = sr.direct;
                if (!ONE_SCENE) {
                    nodes = sc.nodes;
                    tris = sc.tris;
                }
                occ.t = sr.tL;
                occ.id = sr.lid;
                occ.rec = 0;
                occ.mat = 0;
                occ.u = occ.v = 0.f;
                st.sp = 0;
                active = reach && trav_begin(ray, sr.so, sr.sd);
                if (!active) {  // the light sample cannot be seen: add mask * 0
                    const float4 c = *dst;
                    const V3 s = add_direct(v3(c), mask, direct, false);
                    *dst = make_float4(s.x, s.y, s.z, c.w);
                }
            }
        }
        const bool node_phase = node_phase_for<PRT_SHADOW_PHASE_MIN>(active && trav_wants_nodes(ray), active);
        if (active) {
            const int r = trav_step<true, COUNT, true>(node_phase, nodes, tris, ray, kEps, occ, st, cnt);
            if (r != kTravRunning) {
                active = false;
                finished = true;
                lit = r == kTravDone;
            }
        }
    }
    if (COUNT) flush_stats<true>(a.ctr, cnt);
}

__global__ __launch_bounds__(256) void wf_resolve(WfArgs a) {
    const WfChunk ch = a.chunks[a.ctr[kChunk]];
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < ch.pixels; k += gridDim.x * blockDim.x) {
        const WfTile t = a.tiles[find_tile(a.tiles, ch.tile0, ch.ntiles, k, true)];
        const uint32_t w = t.p1 - t.p0;
        const uint32_t p = t.p0 + (k - t.pix0);
        float4* acc = a.jobs[t.job].accum + p;
        float4 s = *acc;
        V3 sum = v3(s.x, s.y, s.z);
        const WfJob& jb = a.jobs[t.job];
        const size_t npx = size_t(jb.width) * jb.height;
        for (uint32_t smp = t.s0; smp < t.s1; ++smp) {
            const float4 r = a.result[t.slot0 + (smp - t.s0) * w + (p - t.p0)];
            sum = sum + v3(r);
            if (jb.samples && smp < jb.samples_cap) jb.samples[size_t(smp) * npx + p] = make_float4(r.x, r.y, r.z, 0.f);
        }
        *acc = make_float4(sum.x, sum.y, sum.z, 0.f);
    }
}

}  // namespace dev

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
namespace {

#define WF_TRY(expr)                                        \
    do {                                                    \
        hipError_t e_ = (expr);                             \
        if (e_ != hipSuccess) return hip_fail(e_, #expr);   \
    } while (0)

struct PlanJob {
    uint32_t pixels;
    uint32_t s_begin, s_end;
};

// Tiles of at most `cap` slots, packed in order into chunks of at most `cap` slots.
void make_plan(const std::vector<PlanJob>& jobs, uint32_t cap, std::vector<WfTile>& tiles,
               std::vector<WfChunk>& chunks) {
    tiles.clear();
    chunks.clear();
    std::vector<WfTile> raw;
    for (uint32_t j = 0; j < jobs.size(); ++j) {
        const PlanJob& pj = jobs[j];
        if (pj.s_end <= pj.s_begin || pj.pixels == 0) continue;
        if (pj.pixels >= cap) {
            for (uint32_t s = pj.s_begin; s < pj.s_end; ++s)
                for (uint32_t p0 = 0; p0 < pj.pixels; p0 += cap)
                    raw.push_back(WfTile{j, p0, std::min(pj.pixels, p0 + cap), s, s + 1, 0, 0, 0});
        } else {
            const uint32_t k = cap / pj.pixels;
            for (uint32_t s = pj.s_begin; s < pj.s_end; s += k)
                raw.push_back(WfTile{j, 0, pj.pixels, s, std::min(pj.s_end, s + k), 0, 0, 0});
        }
    }
    WfChunk cur{0, 0, 0, 0};
    for (WfTile t : raw) {
        const uint32_t n = (t.p1 - t.p0) * (t.s1 - t.s0);
        // A pixel may appear in only one tile of a chunk: resolve adds each tile's
        // samples with one thread per pixel.
        bool overlap = false;
        for (uint32_t k = cur.tile0; k < tiles.size() && !overlap; ++k)
            overlap = tiles[k].job == t.job && tiles[k].p0 < t.p1 && t.p0 < tiles[k].p1;
        if (cur.ntiles && (cur.paths + n > cap || overlap)) {
            chunks.push_back(cur);
            cur = WfChunk{uint32_t(tiles.size()), 0, 0, 0};
        }
        t.slot0 = cur.paths;
        t.pix0 = cur.pixels;
        tiles.push_back(t);
        cur.ntiles++;
        cur.paths += n;
        cur.pixels += t.p1 - t.p0;
    }
    if (cur.ntiles) chunks.push_back(cur);
}

int persistent_blocks(const void* kernel, uint32_t blocks_per_cu, int device) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return 1024;
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, kBlock, 0) != hipSuccess || occ <= 0) occ = 8;
    uint32_t per = blocks_per_cu ? blocks_per_cu : uint32_t(occ);
    return int(std::max<uint32_t>(1, per * uint32_t(prop.multiProcessorCount)));
}

template <class T>
int ensure_buffer(T*& p, size_t& cap, size_t n) {
    if (cap >= n && p) return PRT_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    size_t want = std::max<size_t>(n, 16);
    WF_TRY(hipMalloc(reinterpret_cast<void**>(&p), want * sizeof(T)));
    cap = want;
    return PRT_OK;
}

int ensure_state(WavefrontState& w, int device, uint32_t capacity, uint32_t persistent, hipStream_t stream) {
    if (w.mem && w.capacity >= capacity && w.persistent >= persistent && w.device == device) return PRT_OK;
    if (w.mem) (void)hipFree(w.mem);
    if (w.spill) (void)hipFree(w.spill);
    if (w.ctr) (void)hipFree(w.ctr);
    w.mem = nullptr;
    w.spill = nullptr;
    w.ctr = nullptr;
    // per slot: 2 queues x 64 B, hit 16 + 4, shadow 64, result 16
    w.bytes = kSlotBytes * size_t(capacity);
    WF_TRY(hipMalloc(&w.mem, w.bytes));
    WF_TRY(hipMalloc(reinterpret_cast<void**>(&w.ctr), kNumCounters * sizeof(uint32_t)));
    WF_TRY(fill_wait(w.ctr, 0, kNumCounters * sizeof(uint32_t), stream));
    const size_t spill_entries = size_t(kStackMax - kStackLdsQueue) * persistent;
    WF_TRY(hipMalloc(reinterpret_cast<void**>(&w.spill), spill_entries * sizeof(int)));
    char* p = static_cast<char*>(w.mem);
    auto take = [&](size_t bytes) {
        void* r = p;
        p += bytes;
        return r;
    };
    const size_t C = capacity;
    WfArgs& a = w.args;
    for (int k = 0; k < 2; ++k) {
        a.qa[k] = static_cast<float4*>(take(16 * C));
        a.qb[k] = static_cast<float4*>(take(16 * C));
        a.qc[k] = static_cast<float4*>(take(16 * C));
        a.qd[k] = static_cast<float4*>(take(16 * C));
    }
    a.hits = static_cast<float4*>(take(16 * C));
    a.se = static_cast<float4*>(take(16 * C));
    a.sf = static_cast<float4*>(take(16 * C));
    a.sg = static_cast<float4*>(take(16 * C));
    a.sh = static_cast<float4*>(take(16 * C));
    a.result = static_cast<float4*>(take(16 * C));
    a.hmat = static_cast<uint32_t*>(take(4 * C));
    a.ctr = w.ctr;
    a.spill = w.spill;
    a.spill_stride = persistent;
    w.capacity = capacity;
    w.persistent = persistent;
    w.device = device;
    return PRT_OK;
}

void free_state(WavefrontState& w) {
    void* ptrs[] = {w.mem, w.spill, w.ctr, w.d_jobs, w.d_tiles, w.d_chunks};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    w = WavefrontState();
}

struct RunStats {
    unsigned long long c[kCntNum];
    double ms;
    double kernel_ms[3];  // profile: extend, shade, shadow
    uint32_t launches[3];
};

// Launches one chunk's kernels on `stream`: setup, generate, then per iteration
// extend (wf_extend_packet at the first), shade and shadow, then resolve and
// advance. With `ev` (4 per iteration) the extend / shade / shadow launches are
// bracketed by events.
template <bool COUNT, bool ONE_SCENE>
void launch_chunk(const WfArgs& a, hipStream_t stream, uint32_t max_depth, int pblocks, int gblocks, int sblocks,
                  hipEvent_t* ev) {
    hipLaunchKernelGGL(dev::wf_setup, dim3(1), dim3(64), 0, stream, a);
    hipLaunchKernelGGL(dev::wf_generate, dim3(gblocks), dim3(256), 0, stream, a);
    for (uint32_t it = 0; it < max_depth; ++it) {
        if (ev) (void)hipEventRecord(ev[4 * it], stream);
        if (it == 0)
            hipLaunchKernelGGL((dev::wf_extend_packet<COUNT, ONE_SCENE>), dim3(pblocks), dim3(kBlock), 0, stream, a,
                               it);
        else
            hipLaunchKernelGGL((dev::wf_extend<COUNT, ONE_SCENE>), dim3(pblocks), dim3(kBlock), 0, stream, a, it);
        if (ev) (void)hipEventRecord(ev[4 * it + 1], stream);
        hipLaunchKernelGGL(dev::wf_shade<COUNT>, dim3(sblocks), dim3(kShadeBlock), 0, stream, a, it);
        if (ev) (void)hipEventRecord(ev[4 * it + 2], stream);
        hipLaunchKernelGGL((dev::wf_shadow<COUNT, ONE_SCENE>), dim3(pblocks), dim3(kBlock), 0, stream, a, it);
        if (ev) (void)hipEventRecord(ev[4 * it + 3], stream);
    }
    hipLaunchKernelGGL(dev::wf_resolve, dim3(gblocks), dim3(256), 0, stream, a);
    hipLaunchKernelGGL(dev::wf_advance, dim3(1), dim3(64), 0, stream, a);
}

void launch_chunk_any(bool count, bool one_scene, const WfArgs& a, hipStream_t stream, uint32_t max_depth,
                      int pblocks, int gblocks, int sblocks, hipEvent_t* ev) {
    if (count && one_scene) launch_chunk<true, true>(a, stream, max_depth, pblocks, gblocks, sblocks, ev);
    else if (count) launch_chunk<true, false>(a, stream, max_depth, pblocks, gblocks, sblocks, ev);
    else if (one_scene) launch_chunk<false, true>(a, stream, max_depth, pblocks, gblocks, sblocks, ev);
    else launch_chunk<false, false>(a, stream, max_depth, pblocks, gblocks, sblocks, ev);
}

// Profile mode: every chunk launched directly, each extend / shade / shadow launch
// timed with events; per-kernel device time summed into rs.
int run_profiled(const WfArgs& a, hipStream_t stream, size_t nchunks, uint32_t max_depth, int pblocks, int gblocks,
                 int sblocks, bool count, bool one_scene, RunStats* rs) {
    std::vector<hipEvent_t> ev(4 * max_depth, nullptr);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = PRT_OK;
    double total = 0.0;
    double kms[3] = {0, 0, 0};
    hipError_t e = hipSuccess;
    for (auto& x : ev)
        if (e == hipSuccess) e = hipEventCreate(&x);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    for (size_t c = 0; c < nchunks && e == hipSuccess; ++c) {
        e = hipEventRecord(e0, stream);
        if (e != hipSuccess) break;
        launch_chunk_any(count, one_scene, a, stream, max_depth, pblocks, gblocks, sblocks, ev.data());
        e = hipGetLastError();
        if (e == hipSuccess) e = hipEventRecord(e1, stream);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        total += ms;
        for (uint32_t it = 0; it < max_depth && e == hipSuccess; ++it)
            for (int k = 0; k < 3 && e == hipSuccess; ++k) {
                e = hipEventElapsedTime(&ms, ev[4 * it + k], ev[4 * it + k + 1]);
                kms[k] += ms;
            }
    }
    if (e != hipSuccess) rc = hip_fail(e, "profiled wavefront launch");
    for (auto& x : ev)
        if (x) (void)hipEventDestroy(x);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (rc == PRT_OK && rs) {
        rs->ms = total;
        for (int k = 0; k < 3; ++k) {
            rs->kernel_ms[k] = kms[k];
            rs->launches[k] = uint32_t(nchunks * max_depth);
        }
    }
    return rc;
}

// Renders every job of the plan on `stream`. Returns PRT_OK or an error.
// group_side: camera-path block side g for wf_generate (1, 2, 4, 8; 0 = choose per tile).
int run_wavefront(WavefrontState& w, int device, hipStream_t stream, const std::vector<WfJob>& jobs,
                  const std::vector<PlanJob>& pjobs, uint32_t max_depth, uint32_t capacity, uint32_t blocks_per_cu,
                  uint32_t group_side, bool count, RunStats* rs, bool profile = false) {
    if (max_depth == 0 || max_depth > uint32_t(kMaxIter)) return fail(PRT_ERR_INVALID_ARG, "max_depth out of range");
    uint32_t group_lg = dev::kNone;
    if (group_side != 0) {
        if (group_side != 1 && group_side != 2 && group_side != 4 && group_side != 8)
            return fail(PRT_ERR_INVALID_ARG, "wavefront group side must be 0, 1, 2, 4 or 8");
        group_lg = group_side == 1 ? 0u : group_side == 2 ? 1u : group_side == 4 ? 2u : 3u;
    }
    // Queue entries pack pixel | job << 24.
    if (jobs.size() > 256) return fail(PRT_ERR_LIMIT, "more than 256 jobs in one batch");
    for (const PlanJob& pj : pjobs)
        if (pj.pixels > (1u << 24)) return fail(PRT_ERR_LIMIT, "a view with more than 2^24 pixels");
    std::vector<WfTile> tiles;
    std::vector<WfChunk> chunks;
    make_plan(pjobs, capacity, tiles, chunks);
    if (chunks.empty()) return PRT_OK;
    uint32_t need = 0;
    for (auto& c : chunks) need = std::max(need, c.paths);
    bool one_scene = true;
    for (const WfJob& j : jobs)
        one_scene = one_scene && j.sc.nodes == jobs[0].sc.nodes && j.sc.tris == jobs[0].sc.tris;
    const void* kext = count ? (one_scene ? reinterpret_cast<const void*>(dev::wf_extend<true, true>)
                                          : reinterpret_cast<const void*>(dev::wf_extend<true, false>))
                             : (one_scene ? reinterpret_cast<const void*>(dev::wf_extend<false, true>)
                                          : reinterpret_cast<const void*>(dev::wf_extend<false, false>));
    const int pblocks = persistent_blocks(kext, blocks_per_cu, device);
    int rc = ensure_state(w, device, need, uint32_t(pblocks) * kBlock, stream);
    if (rc != PRT_OK) return rc;
    rc = ensure_buffer(w.d_jobs, w.cap_jobs, jobs.size());
    if (rc == PRT_OK) rc = ensure_buffer(w.d_tiles, w.cap_tiles, tiles.size());
    if (rc == PRT_OK) rc = ensure_buffer(w.d_chunks, w.cap_chunks, chunks.size());
    if (rc != PRT_OK) return rc;
    WF_TRY(hipMemcpyAsync(w.d_jobs, jobs.data(), jobs.size() * sizeof(WfJob), hipMemcpyHostToDevice, stream));
    WF_TRY(hipMemcpyAsync(w.d_tiles, tiles.data(), tiles.size() * sizeof(WfTile), hipMemcpyHostToDevice, stream));
    WF_TRY(hipMemcpyAsync(w.d_chunks, chunks.data(), chunks.size() * sizeof(WfChunk), hipMemcpyHostToDevice, stream));
    WF_TRY(hipMemsetAsync(w.ctr, 0, kNumCounters * sizeof(uint32_t), stream));
    WfArgs a = w.args;
    a.jobs = w.d_jobs;
    a.tiles = w.d_tiles;
    a.chunks = w.d_chunks;
    a.max_depth = max_depth;
    a.group_lg = group_lg;
    a.spill_stride = uint32_t(pblocks) * kBlock;

    hipDeviceProp_t prop;
    WF_TRY(hipGetDeviceProperties(&prop, device));
    const int gblocks = prop.multiProcessorCount * 8;
    const int sblocks = prop.multiProcessorCount * kShadeGridWaves * 64 / kShadeBlock;

    if (profile) {
        rc = run_profiled(a, stream, chunks.size(), max_depth, pblocks, gblocks, sblocks, count, one_scene, rs);
        if (rc == PRT_OK && rs) WF_TRY(copy_wait(rs->c, w.ctr + kStats, sizeof rs->c, hipMemcpyDeviceToHost, stream));
        return rc;
    }
    // Capture one chunk's launch sequence.
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    WF_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    launch_chunk_any(count, one_scene, a, stream, max_depth, pblocks, gblocks, sblocks, nullptr);
    hipError_t e = hipStreamEndCapture(stream, &graph);
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(graph);
        return hip_fail(e, "hipGraphInstantiate");
    }
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e0, stream);
    for (size_t c = 0; c < chunks.size() && e == hipSuccess; ++c) e = hipGraphLaunch(exec, stream);
    (void)hipEventRecord(e1, stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return hip_fail(e, "wavefront launch");
    if (rs) {
        rs->ms = ms;
        for (int k = 0; k < 3; ++k) {
            rs->kernel_ms[k] = 0.0;
            rs->launches[k] = 0;
        }
        WF_TRY(copy_wait(rs->c, w.ctr + kStats, sizeof rs->c, hipMemcpyDeviceToHost, stream));
    }
    return PRT_OK;
}

// The most path slots worth holding: a quarter of the device's free memory (an
// MI355X's 288 GB holds a whole 1080p x 128-spp frame, 265M paths, in one chunk),
// at most 2^28. Fewer, larger chunks leave fewer nearly empty late iterations.
uint64_t max_slots() {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = size_t(8) << 30;
    return std::max<uint64_t>(1u << 20, std::min<uint64_t>(1ull << 28, free_b / 4 / kSlotBytes));
}

uint32_t default_capacity(uint64_t total_paths) {
    return uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(max_slots(), total_paths)));
}

}  // namespace

int render_wavefront(prt_renderer* r, uint32_t n) {
    if (!r->wf) r->wf = new WavefrontState();
    WfJob j;
    j.sc = r->sc;
    j.cam = cam_args(r->cam);
    j.path = PathArgs{r->cfg.seed, r->cfg.view_id, r->cfg.max_depth, r->cfg.rr_start};
    j.width = r->w;
    j.height = r->h;
    j.accum = r->accum;
    j.samples = r->samples;
    j.samples_cap = r->samples ? r->spp : 0;
    std::vector<WfJob> jobs = {j};
    std::vector<PlanJob> pj = {PlanJob{r->w * r->h, r->samples_done, r->samples_done + n}};
    const uint64_t total = uint64_t(r->w) * r->h * n;
    const uint32_t cap = r->cfg.wavefront_queue ? r->cfg.wavefront_queue : default_capacity(total);
    RunStats rs;
    int rc = run_wavefront(*r->wf, r->cfg.device, r->stream, jobs, pj, r->cfg.max_depth, cap,
                           r->cfg.wavefront_blocks_per_cu, r->cfg.wavefront_group, r->cfg.count_rays != 0, &rs,
                           r->cfg.profile_kernels != 0);
    if (rc != PRT_OK) return rc;
    // Fold the wavefront's counters into the renderer's.
    unsigned long long host[kCntNum];
    WF_TRY(copy_wait(host, r->counters, sizeof host, hipMemcpyDeviceToHost, r->stream));
    for (int k = 0; k < kCntNum; ++k) host[k] += rs.c[k];
    WF_TRY(copy_wait(r->counters, host, sizeof host, hipMemcpyHostToDevice, r->stream));
    r->last_ms = rs.ms;
    for (int k = 0; k < 3; ++k) {
        r->kernel_ms[k] = rs.kernel_ms[k];
        r->kernel_launches[k] = rs.launches[k];
    }
    r->samples_done += n;
    r->samples_total += total;
    return PRT_OK;
}

int wavefront_samples(prt_renderer* r, uint32_t spp, float4* accum, float4* samples) {
    const bool had_state = r->wf != nullptr;
    if (!r->wf) r->wf = new WavefrontState();
    WfJob j;
    j.sc = r->sc;
    j.cam = cam_args(r->cam);
    j.path = PathArgs{r->cfg.seed, r->cfg.view_id, r->cfg.max_depth, r->cfg.rr_start};
    j.width = r->w;
    j.height = r->h;
    j.accum = accum;
    j.samples = samples;
    j.samples_cap = spp;
    std::vector<WfJob> jobs = {j};
    std::vector<PlanJob> pj = {PlanJob{r->w * r->h, 0, spp}};
    const uint64_t total = uint64_t(r->w) * r->h * spp;
    const uint32_t cap = r->cfg.wavefront_queue ? r->cfg.wavefront_queue : default_capacity(total);
    RunStats rs;
    const int rc = run_wavefront(*r->wf, r->cfg.device, r->stream, jobs, pj, r->cfg.max_depth, cap,
                                 r->cfg.wavefront_blocks_per_cu, r->cfg.wavefront_group, false, &rs);
    // A megakernel renderer does not keep the queues of a one-off forward pass.
    if (!had_state && r->cfg.integrator != PRT_INTEGRATOR_WAVEFRONT) release_wavefront(r);
    return rc;
}

void release_wavefront(prt_renderer* r) {
    if (!r->wf) return;
    free_state(*r->wf);
    delete r->wf;
    r->wf = nullptr;
}

namespace {

struct BatchCtx {
    std::vector<WfJob> jobs;
    std::vector<PlanJob> pjobs;
    std::vector<float4*> accum;
    uint32_t max_depth = 0;
    uint64_t total = 0;
};

int batch_prepare(prt_job* jobs, uint32_t n, const prt_render_config& c, BatchCtx& b, hipStream_t stream) {
    if (!jobs || n == 0) return fail(PRT_ERR_INVALID_ARG, "no jobs");
    b.max_depth = c.max_depth;
    for (uint32_t i = 0; i < n; ++i) {
        const prt_job& jb = jobs[i];
        if (!jb.scene) return fail(PRT_ERR_INVALID_ARG, "job without a scene");
        if (jb.scene->lights.empty()) return fail(PRT_ERR_NO_EMITTERS, "job scene has no emitters");
        if (jb.camera.xpixels <= 0 || jb.camera.ypixels <= 0) return fail(PRT_ERR_INVALID_ARG, "job camera size");
        if (jb.spp == 0) return fail(PRT_ERR_INVALID_ARG, "job with zero samples per pixel");
        if (uint64_t(jb.camera.xpixels) * uint64_t(jb.camera.ypixels) > (1u << 24))
            return fail(PRT_ERR_LIMIT, "a view with more than 2^24 pixels");
        WfJob w;
        int rc = device_scene(jb.scene, c.device, w.sc, stream);
        if (rc != PRT_OK) return rc;
        w.cam = cam_args(jb.camera);
        w.path = PathArgs{c.seed, jb.view_id, c.max_depth, c.rr_start};
        w.width = uint32_t(jb.camera.xpixels);
        w.height = uint32_t(jb.camera.ypixels);
        const size_t px = size_t(w.width) * w.height;
        float4* acc = nullptr;
        WF_TRY(hipMalloc(reinterpret_cast<void**>(&acc), px * sizeof(float4)));
        b.accum.push_back(acc);
        WF_TRY(fill_wait(acc, 0, px * sizeof(float4), stream));
        w.accum = acc;
        w.samples = nullptr;
        w.samples_cap = 0;
        b.jobs.push_back(w);
        b.pjobs.push_back(PlanJob{uint32_t(px), 0, jb.spp});
        b.total += uint64_t(px) * jb.spp;
    }
    return PRT_OK;
}

void batch_release(BatchCtx& b) {
    for (float4* p : b.accum)
        if (p) (void)hipFree(p);
    b.accum.clear();
}

}  // namespace

}  // namespace prt

using namespace prt;

extern "C" int prt_render_batch(prt_job* jobs, uint32_t n, const prt_render_config* cfg, const prt_tune* tune,
                                prt_stats* stats) {
    clear_error();
    if (tune && tune->block_size != 0 && tune->block_size != uint32_t(kBlock))
        return fail(PRT_ERR_INVALID_ARG, "prt_tune.block_size: the traversal kernels are built for 64 threads");
    prt_render_config c;
    if (cfg) c = *cfg;
    else prt_render_config_default(&c);
    int rc = set_device(c.device);
    if (rc != PRT_OK) return rc;
    PRT_TRY_BEGIN
    BatchCtx b;
    hipStream_t s;
    hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e != hipSuccess) return hip_fail(e, "stream");
    rc = batch_prepare(jobs, n, c, b, s);
    if (rc != PRT_OK) {
        batch_release(b);
        (void)hipStreamDestroy(s);
        return rc;
    }
    WavefrontState w;
    const uint32_t cap = tune && tune->queue_slots ? tune->queue_slots : default_capacity(b.total);
    RunStats rs;
    rc = run_wavefront(w, c.device, s, b.jobs, b.pjobs, c.max_depth, cap, tune ? tune->blocks_per_cu : 0,
                       tune ? tune->group_side : c.wavefront_group, c.count_rays != 0, &rs);
    if (rc == PRT_OK) {
        for (uint32_t i = 0; i < n && rc == PRT_OK; ++i) {
            if (!jobs[i].out_rgb) continue;
            const size_t px = size_t(b.jobs[i].width) * b.jobs[i].height;
            float* tmp = nullptr;
            e = hipMalloc(reinterpret_cast<void**>(&tmp), px * 3 * sizeof(float));
            if (e == hipSuccess) e = launch_mean(b.accum[i], uint32_t(px), jobs[i].spp, tmp, s);
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            if (e == hipSuccess) e = copy_wait(jobs[i].out_rgb, tmp, px * 3 * sizeof(float), hipMemcpyDeviceToHost, s);
            if (tmp) (void)hipFree(tmp);
            if (e != hipSuccess) rc = hip_fail(e, "batch readback");
        }
    }
    if (rc == PRT_OK && stats) {
        stats->primary_rays = rs.c[kCntPrimary];
        stats->extension_rays = rs.c[kCntExtension];
        stats->shadow_rays = rs.c[kCntShadow];
        stats->nodes_visited = rs.c[kCntNodes];
        stats->tris_tested = rs.c[kCntTris];
        stats->hits = rs.c[kCntHits];
        stats->samples = b.total;
        stats->last_render_ms = rs.ms;
    }
    free_state(w);
    (void)hipStreamDestroy(s);
    batch_release(b);
    return rc;
    PRT_TRY_END
}

extern "C" int prt_autotune(prt_job* jobs, uint32_t n, const prt_render_config* cfg, prt_tune* best,
                            double* best_ms) {
    clear_error();
    if (!best) return fail(PRT_ERR_INVALID_ARG, "null argument");
    prt_render_config c;
    if (cfg) c = *cfg;
    else prt_render_config_default(&c);
    c.count_rays = 0;
    int rc = set_device(c.device);
    if (rc != PRT_OK) return rc;
    PRT_TRY_BEGIN
    // Trial workload: every job at its own samples per pixel and width, but with only
    // kTrialRows rows (the camera's vertical field kept, so the rows spread over the
    // whole view): the camera-path grouping, whose candidates depend on the sample
    // count and the row count, then behaves as in the full run, at a small fraction
    // of its paths.
    constexpr int32_t kTrialRows = 8;
    std::vector<prt_job> trial(jobs, jobs + n);
    uint64_t paths = 0;
    for (auto& t : trial) {
        t.camera.ypixels = std::min<int32_t>(t.camera.ypixels, kTrialRows);
        t.out_rgb = nullptr;
        paths += uint64_t(t.camera.xpixels) * t.camera.ypixels * t.spp;
    }
    BatchCtx b;
    hipStream_t s;
    WF_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    rc = batch_prepare(trial.data(), n, c, b, s);
    if (rc != PRT_OK) {
        batch_release(b);
        (void)hipStreamDestroy(s);
        return rc;
    }
    uint64_t full = 0;
    for (uint32_t i = 0; i < n; ++i) full += uint64_t(jobs[i].camera.xpixels) * jobs[i].camera.ypixels * jobs[i].spp;
    WavefrontState w;
    double bms = 1e30;
    prt_tune bt{};
    // Every candidate is timed (best of two) on the trial cut into as many chunks as
    // the full run at that chunk size would take. Occupancy candidates stay near full
    // occupancy: wf_extend fits 32 one-wave workgroups per CU (8 per SIMD), wf_shadow
    // 28. Chunk counts: the fewest that fit the path slots (larger chunks were faster
    // at every size measured on one view, since every chunk ends in nearly empty
    // iterations), and twice and four times as many.
    const uint64_t limit = max_slots();
    const uint32_t full_cap = uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(full, limit)));
    // The trial's path slots for a full-run chunk size: the trial is cut into as many
    // chunks as the full run would take, so its queues drain like the full run's.
    auto trial_cap_for = [&](uint32_t cap) {
        const uint64_t chunks = (full + cap - 1) / cap;
        return uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(paths, (paths + chunks - 1) / chunks)));
    };
    auto time_trial = [&](const prt_tune& t) -> double {
        double best_t = 1e30;
        for (int rep = 0; rep < 2 && rc == PRT_OK; ++rep) {  // best of two
            RunStats rs;
            rc = run_wavefront(w, c.device, s, b.jobs, b.pjobs, c.max_depth, trial_cap_for(t.queue_slots),
                               t.blocks_per_cu, t.group_side, false, &rs);
            best_t = std::min(best_t, rs.ms);
        }
        return best_t;
    };
    // A candidate replaces the best so far only when it is at least 2% faster on the
    // trial: smaller differences are within the trial's run-to-run spread and need
    // not carry over to the full run.
    auto consider = [&](const prt_tune& t) {
        const double ms = time_trial(t);
        if (rc == PRT_OK && (bms >= 1e30 || ms < 0.98 * bms)) {
            bms = ms;
            bt = t;
        }
    };
    // Coordinate search from the defaults (fewest chunks, smallest dividing group):
    // occupancy, then the camera-path grouping, then the chunk count.
    bt = prt_tune{full_cap, 0, kBlock, 0};
    consider(bt);
    for (uint32_t bpc : {20u, 24u, 28u}) {
        prt_tune t = bt;
        t.blocks_per_cu = bpc;
        if (rc == PRT_OK) consider(t);
    }
    const prt_tune occ = bt;
    for (uint32_t g : {1u, 2u, 4u, 8u}) {
        prt_tune t = occ;
        t.group_side = g;
        if (rc == PRT_OK) consider(t);
    }
    const prt_tune grp = bt;
    for (uint32_t div : {2u, 4u}) {
        prt_tune t = grp;
        t.queue_slots = std::max<uint32_t>(1u << 16, full_cap / div);
        if (t.queue_slots < full_cap && rc == PRT_OK) consider(t);
    }
    free_state(w);
    (void)hipStreamDestroy(s);
    batch_release(b);
    if (rc != PRT_OK) return rc;
    *best = bt;
    if (best_ms) *best_ms = bms;
    return PRT_OK;
    PRT_TRY_END
}

