// Path-replay backpropagation (Vicini, Speierer, Jakob: "Path Replay
// Backpropagation: Differentiating Light Paths using Constant Memory and Linear
// Time", SIGGRAPH 2021) for the reference's estimator.
//
// Given dLoss/dI for every pixel (I = mean of `spp` samples), each sample's
// radiance L has adjoint delta = dLoss/dI / spp. For each pixel-sample the kernel
// takes L from the forward pass when the renderer kept it (keep_samples), or traces
// the path once (primal, the megakernel's arithmetic) to get it, then replays the
// path with the same Philox slots, keeping the radiance still to be collected,
// rem = L - (what the path collected so far). Parameters:
//   kd_m (RGB)  a diffuse vertex multiplies everything collected after it by
//               kd_c, so it contributes delta_c * rem_c / kd_c (rem taken after
//               that vertex's emission, before its next-event estimate);
//   e_m         MTL Ka = e_m (1,1,1), ka = 3 e_m A_L: emission terms contribute
//               delta . mask * 3 A_L / A_L and next-event terms delta . mask * 3 A_L wf.
// Russian roulette (p = max(mask) after the vertex's weight) is differentiated
// as well: where p >= 1 the path always survives and mask /= p is a
// deterministic rescale; where p < 1 the survival probability depends on the
// albedos, and the boundary (score) term of that decision is weighted by the
// loss change it causes, which for a loss with curvature hess includes
// -(hess / 2) |R / spp|^2 (R = radiance collected after the decision). With
// both, the result is an unbiased estimate of the derivative of the expected
// loss, the quantity a central difference of renders with common random
// numbers estimates. Textured albedos get no kd gradient.
// Gradients accumulate in fp64: LDS per workgroup, then a fixed-order reduction.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "../host/common.hpp"
#include "path.hpp"
#include "renderer.hpp"

namespace prt {

struct AdjArgs {
    SceneArgs sc;
    CamArgs cam;
    PathArgs path;
    uint32_t width, height, spp;
    uint32_t tiles_x, tiles;
    uint32_t parts;       // workgroups per tile, each replaying a 1/parts range of its pixel-samples
    const float* grad;    // W*H*3 dLoss / dI
    const float4* samples;  // nullptr, or each sample's radiance from the forward pass [s * W * H + p]
    double norm;          // 1 / spp: dI / dL of one sample
    double half_hess;     // (d2 Loss / dI2) / (2 spp^2): loss curvature per sample radiance
    uint32_t num_params;  // 4 * materials: kd (M*3) then emission (M)
    uint32_t num_materials;
    double* partial;      // grid x num_params (grid: tiles x parts rounded up to 8)
    int* spill;
    uint32_t spill_stride;
};

namespace dev {

// The replay's RNG: Philox with each product's low and high halves taken by
// separate instructions. The 64-bit-product form takes this kernel from 167 to 172
// VGPRs, 2 waves per SIMD instead of 3 (45.2 against 34.1 ms on 512^2 x 64 samples
// of the 100k-triangle scene), or, held at 3 waves, spills two VGPRs to scratch
// memory, which the allocation-free entry point should not need. Same bits.
using AdjRng = RngT<false>;

struct D3 {
    double x, y, z;
};
__device__ __forceinline__ D3 d3(V3 v) { return D3{double(v.x), double(v.y), double(v.z)}; }
__device__ __forceinline__ double ddot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ D3 dsub(D3 a, D3 b) { return D3{a.x - b.x, a.y - b.y, a.z - b.z}; }

// Tie-aware weight of each channel in d log max(mask): 1 for a unique maximum,
// 1/2 for each of two or three equal maxima (the mean of the one-sided
// derivatives of max, which is what a central difference measures).
__device__ __forceinline__ D3 rr_weight(uint32_t ties) {
    const double w = __popc(ties) == 1 ? 1.0 : 0.5;
    return D3{(ties & 1u) ? w : 0.0, (ties & 2u) ? w : 0.0, (ties & 4u) ? w : 0.0};
}

// Traces the shadow ray of a next-event sample; true when the light is visible.
template <class StackType>
__device__ __forceinline__ bool shadow_visible(const SceneArgs& sc, const NeeSample& ns, ShadowReq& sr, StackType& st,
                                               Counts& cnt) {
    sr.direct = v3(0.f, 0.f, 0.f);
    if (!nee_prepare<false>(sc, ns, sr, cnt)) return false;
    HitRec occ;
    occ.t = sr.tL;
    occ.id = sr.lid;
    occ.rec = 0;
    occ.mat = 0;
    occ.u = occ.v = 0.f;
    return !traverse<true, false>(sc, sr.so, sr.sd, kEps, occ, st, cnt);
}

// samplePixel's radiance (the megakernel's arithmetic).
template <class StackType>
__device__ V3 primal(const AdjArgs& a, int x, int y, uint32_t sample, StackType& st, Counts& cnt) {
    const SceneArgs& sc = a.sc;
    AdjRng rng{a.path.seed, a.path.view, uint32_t(y) * uint32_t(a.cam.xpixels) + uint32_t(x), sample};
    V3 ro, rd;
    camera_ray(a.cam, x, y, rng.block(0), ro, rd);
    PathState ps;
    path_init(ps, ro, rd);
    while (true) {
        HitRec h;
        h.t = kFltMax;
        h.id = 0;
        h.rec = 0;
        h.mat = 0;
        h.u = h.v = 0.f;
        if (!traverse<false, false>(sc, ps.ro, ps.rd, kEps, h, st, cnt)) {
            path_miss(ps);
            break;
        }
        bool want_nee;
        NeeSample ns;
        const bool more = shade_vertex<false, false>(sc, a.path, rng, ps, h, &want_nee, ns, cnt);
        if (want_nee) {
            ShadowReq sr;
            const bool visible = shadow_visible(sc, ns, sr, st, cnt);
            ps.color = add_direct(ps.color, ns.mask, sr.direct, visible);
        }
        if (!more) break;
    }
    return ps.color;
}

// Materials of the path's untextured diffuse vertices, 8 bits each (at most 16).
struct VertexList {
    uint64_t lo = 0, hi = 0;
    uint32_t n = 0;
    __device__ __forceinline__ void push(uint32_t m) {
        if (n < 8) lo |= uint64_t(m) << (8 * n);
        else hi |= uint64_t(m) << (8 * (n - 8));
        ++n;
    }
    __device__ __forceinline__ uint32_t at(uint32_t i) const {
        return uint32_t(((i < 8 ? lo : hi) >> (8 * (i & 7))) & 0xffu);
    }
};

// The replay state of one pixel-sample, advanced a vertex at a time (replay_hit,
// replay_after_nee) so that a lane whose path has ended can start the next sample
// while the others go on. L comes from the primal pass of the same sample.
struct Replay {
    PathState ps;
    AdjRng rng;
    D3 rem, delta;
    V3 L;
    VertexList vl;
    uint32_t prev_ties;  // channels at the maximum at the previous roulette step (0: none)
};

__device__ __forceinline__ void replay_begin(const AdjArgs& a, int x, int y, uint32_t sample, D3 delta, V3 L,
                                             Replay& r) {
    r.rem = d3(L);
    r.delta = delta;
    r.L = L;
    r.vl = VertexList();
    r.prev_ties = 0;
    r.rng = AdjRng{a.path.seed, a.path.view, uint32_t(y) * uint32_t(a.cam.xpixels) + uint32_t(x), sample};
    V3 ro, rd;
    camera_ray(a.cam, x, y, r.rng.block(0), ro, rd);
    path_init(r.ps, ro, rd);
}

// A lane of the replay kernel traces either its path's extension ray or the shadow
// ray of the vertex it has just shaded; the vertex's work is split at the shadow
// ray: replay_hit() once the extension ray's closest hit is known, replay_after_nee()
// once the shadow ray's visibility is known (or at once when there is none).
// The *Done modes hold a lane whose traversal has ended until its vertex's work runs.
enum : int { kLaneIdle = 0, kLaneExt = 1, kLaneShadow = 2, kLaneExtDone = 3, kLaneShadowDone = 4 };

// What replay_after_nee() needs of the vertex that replay_hit() shaded.
struct Pending {
    V3 mask;         // the light sample's throughput
    V3 direct;       // the light's contribution if visible
    float wf;        // its geometric weight
    uint32_t lmat;   // the light triangle's material
    RrInfo rr;
    uint32_t b, before;  // bounce index and diffuse vertices ahead of the vertex
    bool more, want_nee;
};

__device__ __forceinline__ void ray_reset(HitRec& h, float t, uint32_t id) {
    h.t = t;
    h.id = id;
    h.rec = 0;
    h.mat = 0;
    h.u = h.v = 0.f;
}

// Starts the closest-hit traversal of the path's current ray; false when its
// direction cannot hit any triangle (the path then ends, as at a miss).
template <class StackType>
__device__ __forceinline__ bool start_ext(const Replay& r, TravRay& tr, HitRec& h, StackType& st) {
    ray_reset(h, kFltMax, 0);
    st.sp = 0;
    return trav_begin(tr, r.ps.ro, r.ps.rd);
}

// The light term and the roulette terms of the vertex; then the next ray. Returns
// the lane's next mode.
template <class StackType>
__device__ int replay_after_nee(const AdjArgs& a, Replay& r, const Pending& pd, bool visible, TravRay& tr, HitRec& h,
                                double* g, StackType& st) {
    const SceneArgs& sc = a.sc;
    const double gain = 3.0 * double(sc.lights_area);  // d ka / d e
    PathState& ps = r.ps;
    if (pd.want_nee) {
        if (visible) {
            r.rem = dsub(r.rem, d3(pd.mask * pd.direct));
            atomicAdd(&g[3 * a.num_materials + pd.lmat], ddot(r.delta, d3(pd.mask)) * gain * double(pd.wf));
        }
        ps.color = add_direct(ps.color, pd.mask, pd.direct, visible);
    }
    if (pd.rr.p >= 0.f && ps.b == pd.b + 1) {
        // Survived roulette step j. Everything collected after it, R_j = L - C_j,
        // carries 1/p_j, and with S_k[c] = sum over diffuse vertices v <= k of
        // d log kd_{m_v,c}: d log p_j = w_j . S_j - w_{j-1} . S_{j-1}. The step adds
        // -kappa_j d log p_j, with kappa_j = delta . R_j where p_j >= 1 (the path
        // always survives; mask /= p_j is a deterministic rescale), and where
        // p_j < 1 kappa_j = (hess / 2) |R_j / spp|^2: the score term of the survival
        // probability times the loss change of that decision, delta . R_j -
        // (hess / 2) |R_j / spp|^2, less the derivative of 1/p_j.
        const D3 R = dsub(d3(r.L), d3(ps.color));
        const double kappa = pd.rr.p >= 1.f ? ddot(r.delta, R) : a.half_hess * ddot(R, R);
        if (kappa != 0.0) {
            const D3 w = rr_weight(pd.rr.ties), wp = rr_weight(r.prev_ties);
            for (uint32_t i = 0; i < r.vl.n; ++i) {
                const uint32_t m = r.vl.at(i);
                const float4 k4 = sc.mats[size_t(m) * 4];
                const double kd[3] = {k4.x, k4.y, k4.z};
                const double cf[3] = {i < pd.before ? wp.x - w.x : -w.x, i < pd.before ? wp.y - w.y : -w.y,
                                      i < pd.before ? wp.z - w.z : -w.z};
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    if (cf[c] != 0.0 && kd[c] != 0.0) atomicAdd(&g[3 * m + c], kappa * cf[c] / kd[c]);
            }
        }
        r.prev_ties = pd.rr.ties;
    }
    if (!pd.more) return kLaneIdle;
    return start_ext(r, tr, h, st) ? kLaneExt : kLaneIdle;
}

// The vertex at the extension ray's closest hit h (hit = false: the sky, which
// carries no parameter): emission and diffuse terms, shade_vertex and the light
// sample, whose shadow ray the lane then traces. Returns the lane's next mode.
template <class StackType>
__device__ int replay_hit(const AdjArgs& a, Replay& r, HitRec& h, bool hit, Pending& pd, TravRay& tr, double* g,
                          StackType& st, Counts& cnt) {
    if (!hit) return kLaneIdle;
    const SceneArgs& sc = a.sc;
    const uint32_t M = a.num_materials;
    const double gain = 3.0 * double(sc.lights_area);  // d ka / d e
    PathState& ps = r.ps;
    const float4 mkd = sc.mats[size_t(h.mat) * 4], mka = sc.mats[size_t(h.mat) * 4 + 1];
    if (ps.b == 0 || ps.prev == 1 || ps.prev == 4) {
        // the emission term shade_vertex is about to add
        const V3 e = ps.mask * v3(mka) * sc.inv_lights_area;
        atomicAdd(&g[3 * M + h.mat], ddot(r.delta, d3(ps.mask)) * gain * double(sc.inv_lights_area));
        r.rem = dsub(r.rem, d3(e));
    }
    NeeSample ns;
    pd.b = ps.b;
    pd.more = shade_vertex<false, false>(sc, a.path, r.rng, ps, h, &pd.want_nee, ns, cnt, &pd.rr);
    pd.before = r.vl.n;
    if (ps.prev == 0 /*DIFFUSE*/ && __float_as_int(mka.w) < 0) {
        const double kd[3] = {mkd.x, mkd.y, mkd.z};
        const double rc[3] = {r.rem.x, r.rem.y, r.rem.z};
        const double dc[3] = {r.delta.x, r.delta.y, r.delta.z};
#pragma unroll
        for (int c = 0; c < 3; ++c)
            if (kd[c] != 0.0) atomicAdd(&g[3 * h.mat + c], dc[c] * rc[c] / kd[c]);
        r.vl.push(h.mat);
    }
    if (pd.want_nee) {
        ShadowReq sr;
        sr.direct = v3(0.f, 0.f, 0.f);
        const bool reach = nee_prepare<false>(sc, ns, sr, cnt);
        pd.mask = ns.mask;
        pd.direct = sr.direct;
        pd.wf = sr.wf;
        pd.lmat = sr.lmat;
        if (reach) {
            // The occlusion bound: the light sample itself, (t, id) as the closest hit.
            ray_reset(h, sr.tL, sr.lid);
            st.sp = 0;
            if (trav_begin(tr, sr.so, sr.sd)) return kLaneShadow;
            return replay_after_nee(a, r, pd, true, tr, h, g, st);  // no triangle can occlude it
        }
        return replay_after_nee(a, r, pd, false, tr, h, g, st);
    }
    return replay_after_nee(a, r, pd, false, tr, h, g, st);
}

// Occupancy of the replay kernel (empty: the compiler's choice, 3 waves per SIMD).
#ifndef PRT_ADJOINT_WAVES
#define PRT_ADJOINT_WAVES
#endif
// Idle lanes that make the wave start new pixel-samples. On 512^2 x 64 samples of
// the 100k-triangle scene, one workgroup per tile: 1 lane 71.7 / 73.1 ms, 8: 72.7 /
// 74.6, 16: 72.5 / 72.9, 32: 79.8 / 79.0 (two runs each); 6 workgroups per tile:
// 1 lane 34.31 / 34.29 ms, 8: 34.04 / 34.11 (interleaved).
#ifndef PRT_ADJOINT_REFILL
#define PRT_ADJOINT_REFILL 8
#endif
constexpr uint32_t kAdjRefill = PRT_ADJOINT_REFILL;
// Lanes whose traversal has ended wait until this many have (or none is still
// traversing), then do their vertices' work together (512^2 x 64 samples of the
// 100k-triangle scene: 16 lanes 73.8 ms, 32: 66.3, 48: 68.8; each vertex's work at
// the step its traversal ended: 77.9; with 6 workgroups per tile 16: 37.7, 32: 34.3,
// 48: 36.8).
#ifndef PRT_ADJOINT_BATCH
#define PRT_ADJOINT_BATCH 32
#endif
constexpr uint32_t kAdjBatch = PRT_ADJOINT_BATCH;
// Lanes still wanting nodes below which the wave's step is a leaf phase. With the
// lanes tracing independently most of them are mid-traversal at any step (512^2 x
// 64 samples: 1 lane 66.7 ms, 8: 55.1, 12: 54.1, 16: 52.6, 20: 52.8, 24: 53.9; with
// 6 workgroups per tile 8: 34.6, 16: 34.3, 24: 35.4).
#ifndef PRT_ADJOINT_PHASE_MIN
#define PRT_ADJOINT_PHASE_MIN 16
#endif
static_assert(kAdjRefill >= 1 && kAdjRefill <= 64, "refill threshold out of range");
__global__ __launch_bounds__(kBlock) PRT_ADJOINT_WAVES void adjoint_kernel(AdjArgs a) {
    __shared__ int s_stack[kStackLds * kBlock];
    __shared__ double s_g[4 * PRT_MAX_MATERIALS];
    for (uint32_t k = threadIdx.x; k < a.num_params; k += kBlock) s_g[k] = 0.0;
    __syncthreads();
    const uint32_t units = a.tiles * a.parts;
    const uint32_t unit = xcd_tile(blockIdx.x, units);
    const int lane = int(threadIdx.x);
    if (unit < units) {
        const uint32_t tile = unit / a.parts, part = unit - tile * a.parts;
        Stack st;
        st.bind(s_stack, a.spill + size_t(blockIdx.x) * kBlock, a.spill_stride);
        Counts cnt = {0, 0, 0, 0, 0, 0};
        // The wave works through its tile's pixel-samples in pixel-major order (64
        // consecutive samples of one pixel are 64 nearly equal camera rays). A lane
        // whose path has ended takes the next item, so the wave does not wait for
        // each group's longest path (512^2 x 64 samples: 77.9 ms with one sample per
        // lane per pixel and whole traversals in lockstep).
        const uint32_t tx = (tile % a.tiles_x) * 8, ty = (tile / a.tiles_x) * 8;
        const uint32_t cols = min(8u, a.width - tx), rows = min(8u, a.height - ty);
        // This workgroup's share: items [i0, i1) of the tile's pixel-samples.
        const uint64_t items = uint64_t(cols) * rows * a.spp;
        const uint32_t i0 = uint32_t(items * part / a.parts), i1 = uint32_t(items * (part + 1) / a.parts);
        const unsigned long long below = (1ull << lane) - 1ull;
        uint32_t next = i0;  // the first item not yet taken (the same in every lane)
        int mode = kLaneIdle;
        Replay r;
        TravRay tr;
        tr.leaf = kNoChild;
        HitRec h;
        Pending pd;
        // Each step advances every busy lane's traversal by one node or leaf phase;
        // a lane whose traversal ends does that vertex's work and starts its next ray
        // (shadow or extension) at once, so lanes do not wait for one another's
        // traversals, and a lane whose path has ended takes the next pixel-sample.
        bool found = false;  // the ended traversal's result: a hit (extension) or visibility (shadow)
        while (true) {
            const unsigned long long done = __ballot(mode >= kLaneExtDone);
            if (done != 0ull && (uint32_t(__popcll(done)) >= kAdjBatch ||
                                 __ballot(mode == kLaneExt || mode == kLaneShadow) == 0ull)) {
                if (mode == kLaneExtDone) mode = replay_hit(a, r, h, found, pd, tr, s_g, st, cnt);
                else if (mode == kLaneShadowDone) mode = replay_after_nee(a, r, pd, found, tr, h, s_g, st);
            }
            const unsigned long long idle = __ballot(mode == kLaneIdle);
            const uint32_t nidle = uint32_t(__popcll(idle));
            if ((nidle >= kAdjRefill || nidle == 64u) && next < i1) {
                const uint32_t j = next + uint32_t(__popcll(idle & below));
                if (mode == kLaneIdle && j < i1) {
                    const uint32_t p = j / a.spp, s = j - p * a.spp;
                    const int x = int(tx + p % cols), y = int(ty + p / cols);
                    const size_t pix = size_t(y) * a.width + size_t(x);
                    const D3 delta{a.norm * double(a.grad[pix * 3]), a.norm * double(a.grad[pix * 3 + 1]),
                                   a.norm * double(a.grad[pix * 3 + 2])};
                    const V3 L = a.samples ? v3(a.samples[size_t(s) * a.width * a.height + pix])
                                           : primal(a, x, y, s, st, cnt);
                    replay_begin(a, x, y, s, delta, L, r);
                    mode = start_ext(r, tr, h, st) ? kLaneExt : kLaneIdle;
                }
                next += nidle;
            }
            if (__ballot(mode != kLaneIdle) == 0ull && next >= i1) break;
            const bool busy = mode == kLaneExt || mode == kLaneShadow;
            const bool node_phase = node_phase_for<PRT_ADJOINT_PHASE_MIN>(busy && trav_wants_nodes(tr), busy);
            if (busy) {
                // Both ray kinds take the folded box test here: 33.4 against 34.0 ms on
                // 512^2 x 64 samples of the 100k-triangle scene.
                const int res = trav_step<false, false, false, true>(node_phase, a.sc.nodes, a.sc.tris, tr, kEps, h,
                                                                     st, cnt, mode == kLaneShadow);
                if (res != kTravRunning) {
                    found = mode == kLaneExt ? (h.t != kFltMax || h.id != 0u) : res != kTravOccluded;
                    mode = mode == kLaneExt ? kLaneExtDone : kLaneShadowDone;
                }
            }
        }
    }
    __syncthreads();
    double* out = a.partial + size_t(blockIdx.x) * a.num_params;
    for (uint32_t k = threadIdx.x; k < a.num_params; k += kBlock) out[k] = s_g[k];
}

// Sums the per-workgroup partials of parameter blockIdx.x in a fixed order: thread t
// adds partials t, t + 256, t + 512, ... in turn, then the 256 sums are added in a
// fixed pairwise tree. The result depends only on the partials, not on scheduling.
constexpr uint32_t kReduceBlock = 256;
__global__ __launch_bounds__(kReduceBlock) void adjoint_reduce(const double* partial, uint32_t blocks,
                                                              uint32_t num_params, double* out) {
    __shared__ double s[kReduceBlock];
    const uint32_t k = blockIdx.x;
    double v = 0.0;
    for (uint32_t b = threadIdx.x; b < blocks; b += kReduceBlock) v += partial[size_t(b) * num_params + k];
    s[threadIdx.x] = v;
    __syncthreads();
    for (uint32_t w = kReduceBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[k] = s[0];
}

}  // namespace dev

// Scratch of prt_adjoint_grad_static, the entry point that allocates nothing: the
// per-workgroup partial gradients (grid x P doubles), then the P reduced
// gradients. 16 MiB on every device the library runs that entry point on, held
// until the process ends. In host code `scratch` names a host-side stand-in, not
// device memory: host code uses only the address hipGetSymbolAddress returns.
constexpr size_t kStaticScratchDoubles = size_t(1) << 21;
static __device__ __align__(256) double scratch[kStaticScratchDoubles];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch_ptr[kMaxScratchDevices] = {};

// The device address of `scratch` on `device` (the current device), looked up once
// per device on that device's first call.
int static_scratch(int device, double** out) {
    if (device < 0 || device >= kMaxScratchDevices) return fail(PRT_ERR_LIMIT, "device ordinal above 63");
    std::lock_guard<std::mutex> g(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        void* ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        if (!ptr) return fail(PRT_ERR_HIP, "hipGetSymbolAddress(scratch) returned a null address");
        g_scratch_ptr[device] = static_cast<double*>(ptr);
    }
    *out = g_scratch_ptr[device];
    return PRT_OK;
}
}  // namespace

namespace {
#define ADJ_TRY(expr)                                           \
    do {                                                        \
        hipError_t e_ = (expr);                                 \
        if (e_ != hipSuccess) {                                 \
            rc = hip_fail(e_, #expr);                           \
            goto done;                                          \
        }                                                       \
    } while (0)
// PRT_ADJOINT_PRIMAL_IN_REPLAY=1: without kept radiance, trace each sample's primal
// path inside the replay kernel even where the per-sample buffer would fit (tests).
#define HIP_CHECK(expr)                                    \
    do {                                                   \
        hipError_t e_ = (expr);                            \
        if (e_ != hipSuccess) return hip_fail(e_, #expr);  \
    } while (0)
bool primal_in_replay() {
    const char* v = std::getenv("PRT_ADJOINT_PRIMAL_IN_REPLAY");
    return v && v[0] == '1';
}

// The replay's arguments for the renderer's view: seed, view, depth and the scene
// as the renderer holds them; grad = W*H*3 dLoss/dI on the device.
// Workgroups per 8x8 tile of the replay. One workgroup per tile left most of the GPU
// idle for the second half of the kernel: 512^2 x 64 samples of the 100k-triangle
// scene is 4,096 tiles against 3,072 resident waves (3 per SIMD), and the tiles'
// path costs differ, so the kernel ran 2.5 times as long as a tile's replay
// (6.4 of 12 wave slots per compute unit occupied on average). Splitting each tile's
// pixel-samples into `parts` ranges gives about 8 workgroups per resident wave slot.
// The partials (ceil8(tiles x parts) x 4 x materials doubles) must also fit the
// allocation-free entry point's scratch, so both entry points pick the same split and
// give the same sums.
uint32_t adjoint_parts(const prt_renderer* r, uint32_t spp) {
    constexpr uint64_t kTargetUnits = 8ull * 3 * 4 * 256;  // 8 x (3 waves x 4 SIMDs x 256 CUs)
    const uint64_t tiles = std::max<uint32_t>(1, r->tiles);
    const uint64_t P = 4ull * r->scene->num_materials;
    uint64_t parts = std::max<uint64_t>(1, (kTargetUnits + tiles - 1) / tiles);
    parts = std::min<uint64_t>(parts, std::max<uint32_t>(1, spp));
    while (parts > 1 && ((tiles * parts + 7) / 8) * 8 * P + P > kStaticScratchDoubles) --parts;
    return uint32_t(parts);
}

AdjArgs adjoint_args(const prt_renderer* r, const float* grad, float hess, uint32_t spp, double* partial) {
    AdjArgs a;
    a.sc = r->sc;
    a.cam = cam_args(r->cam);
    a.path = PathArgs{r->cfg.seed, r->cfg.view_id, r->cfg.max_depth, r->cfg.rr_start};
    a.width = r->w;
    a.height = r->h;
    a.spp = spp;
    a.tiles_x = r->tiles_x;
    a.tiles = r->tiles;
    a.parts = adjoint_parts(r, spp);
    a.grad = grad;
    // The forward pass's per-sample radiance, when it was kept for exactly these
    // samples with the scene's current materials (same seed, view and depth: the
    // renderer's configuration).
    a.samples = r->samples && r->samples_ok && r->samples_done >= spp && spp <= r->spp &&
                        r->samples_version == r->scene->mats_version
                    ? r->samples
                    : nullptr;
    a.norm = 1.0 / double(spp);
    a.half_hess = double(hess) / (2.0 * double(spp) * double(spp));
    a.num_params = 4 * r->scene->num_materials;
    a.num_materials = r->scene->num_materials;
    a.partial = partial;
    a.spill = r->spill;
    a.spill_stride = r->spill_threads;
    return a;
}
}  // namespace

}  // namespace prt

using namespace prt;

extern "C" int prt_adjoint_grad(prt_renderer* r, const float* grad_rgb, float hess, uint32_t spp, float* d_albedo,
                                float* d_emission) {
    clear_error();
    if (!r || !grad_rgb || !d_albedo || !d_emission || spp == 0 || !(hess >= 0.f))
        return fail(PRT_ERR_INVALID_ARG, "bad argument");
    if (r->cfg.max_depth > 16) return fail(PRT_ERR_LIMIT, "the adjoint supports max_depth <= 16");
    int rc = set_device(r->cfg.device);
    if (rc != PRT_OK) return rc;
    rc = device_scene(r->scene, r->cfg.device, r->sc, r->stream);
    if (rc != PRT_OK) return rc;
    const size_t px = size_t(r->w) * r->h;
    const uint32_t M = r->scene->num_materials;
    const uint32_t P = 4 * M;
    const uint32_t grid = ((r->tiles * adjoint_parts(r, spp) + 7) / 8) * 8;
    float* d_grad = nullptr;
    double* d_part = nullptr;
    double* d_out = nullptr;
    float4* d_samp = nullptr;  // this call's per-sample radiance, when none was kept
    float4* d_acc = nullptr;
    std::vector<double> g(P);
    float ms = 0.f;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    AdjArgs a;
    ADJ_TRY(hipMalloc(reinterpret_cast<void**>(&d_grad), px * 3 * sizeof(float)));
    ADJ_TRY(hipMalloc(reinterpret_cast<void**>(&d_part), size_t(grid) * P * sizeof(double)));
    ADJ_TRY(hipMalloc(reinterpret_cast<void**>(&d_out), P * sizeof(double)));
    ADJ_TRY(copy_wait(d_grad, grad_rgb, px * 3 * sizeof(float), hipMemcpyHostToDevice, r->stream));
    a = adjoint_args(r, d_grad, hess, spp, d_part);
    // Without kept radiance, a wavefront forward pass first writes each sample's
    // radiance to a buffer of this call (when that fits in a quarter of the free
    // memory), and the replay kernel only replays; otherwise each lane traces its
    // sample's primal path before replaying it. The radiance is the same bit for
    // bit either way, and so are the gradients.
    if (!a.samples && !primal_in_replay()) {
        size_t free_b = 0, total_b = 0;
        const size_t need = px * (size_t(spp) + 1) * sizeof(float4);
        if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && need <= free_b / 4) {
            ADJ_TRY(hipMalloc(reinterpret_cast<void**>(&d_samp), px * spp * sizeof(float4)));
            ADJ_TRY(hipMalloc(reinterpret_cast<void**>(&d_acc), px * sizeof(float4)));
            ADJ_TRY(hipMemsetAsync(d_acc, 0, px * sizeof(float4), r->stream));
        }
    }
    ADJ_TRY(hipEventCreate(&e0));
    ADJ_TRY(hipEventCreate(&e1));
    ADJ_TRY(hipEventRecord(e0, r->stream));
    if (d_samp) {
        rc = wavefront_samples(r, spp, d_acc, d_samp);
        if (rc != PRT_OK) goto done;
        a.samples = d_samp;
    }
    hipLaunchKernelGGL(dev::adjoint_kernel, dim3(grid), dim3(kBlock), 0, r->stream, a);
    ADJ_TRY(hipGetLastError());
    hipLaunchKernelGGL(dev::adjoint_reduce, dim3(P), dim3(dev::kReduceBlock), 0, r->stream, d_part, grid, P, d_out);
    ADJ_TRY(hipGetLastError());
    ADJ_TRY(hipEventRecord(e1, r->stream));
    ADJ_TRY(hipEventSynchronize(e1));
    ADJ_TRY(hipEventElapsedTime(&ms, e0, e1));
    ADJ_TRY(copy_wait(g.data(), d_out, P * sizeof(double), hipMemcpyDeviceToHost, r->stream));
    for (uint32_t k = 0; k < 3 * M; ++k) d_albedo[k] = float(g[k]);
    for (uint32_t k = 0; k < M; ++k) d_emission[k] = float(g[3 * M + k]);
    r->last_ms = ms;
done:
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    void* ptrs[] = {d_grad, d_part, d_out, d_samp, d_acc};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    return rc;
}

extern "C" int prt_adjoint_grad_static(prt_renderer* r, const float* grad_rgb_device, float hess, uint32_t spp,
                                       float* d_albedo, float* d_emission) {
    clear_error();
    if (!r || !grad_rgb_device || !d_albedo || !d_emission || spp == 0 || !(hess >= 0.f))
        return fail(PRT_ERR_INVALID_ARG, "bad argument");
    if (r->cfg.max_depth > 16) return fail(PRT_ERR_LIMIT, "the adjoint supports max_depth <= 16");
    int rc = set_device(r->cfg.device);
    if (rc != PRT_OK) return rc;
    {
        hipPointerAttribute_t at;
        hipError_t e = hipPointerGetAttributes(&at, grad_rgb_device);
        if (e != hipSuccess) (void)hipGetLastError();
        if (e != hipSuccess || at.type != hipMemoryTypeDevice || at.device != r->cfg.device)
            return fail(PRT_ERR_INVALID_ARG, "grad_rgb_device is not device memory of the renderer's device");
    }
    // The renderer's creation uploaded the scene to this device: this refreshes at
    // most the material table in place (a copy, no allocation).
    rc = device_scene(r->scene, r->cfg.device, r->sc, r->stream);
    if (rc != PRT_OK) return rc;
    const uint32_t M = r->scene->num_materials;
    const uint32_t P = 4 * M;
    const uint32_t grid = ((r->tiles * adjoint_parts(r, spp) + 7) / 8) * 8;
    if (size_t(grid) * P + P > kStaticScratchDoubles)
        return fail(PRT_ERR_LIMIT, "image too large for the static scratch: (8x8 tiles rounded up to 8) x 4 x "
                                   "materials + 4 x materials must be at most 2^21 doubles");
    double* base = nullptr;
    rc = static_scratch(r->cfg.device, &base);
    if (rc != PRT_OK) return rc;
    // Every workgroup writes all P of its partials and the reduction writes all P
    // results, so nothing in the scratch needs clearing between calls.
    AdjArgs a = adjoint_args(r, grad_rgb_device, hess, spp, base);
    double* d_out = base + size_t(grid) * P;
    double g[4 * PRT_MAX_MATERIALS];
    float ms = 0.f;
    HIP_CHECK(hipEventRecord(r->ev0, r->stream));
    hipLaunchKernelGGL(dev::adjoint_kernel, dim3(grid), dim3(kBlock), 0, r->stream, a);
    HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(dev::adjoint_reduce, dim3(P), dim3(dev::kReduceBlock), 0, r->stream,
                       static_cast<const double*>(base), grid, P, d_out);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipEventRecord(r->ev1, r->stream));
    HIP_CHECK(hipEventSynchronize(r->ev1));
    HIP_CHECK(hipEventElapsedTime(&ms, r->ev0, r->ev1));
    HIP_CHECK(copy_wait(g, d_out, P * sizeof(double), hipMemcpyDeviceToHost, r->stream));
    for (uint32_t k = 0; k < 3 * M; ++k) d_albedo[k] = float(g[k]);
    for (uint32_t k = 0; k < M; ++k) d_emission[k] = float(g[3 * M + k]);
    r->last_ms = ms;
    return PRT_OK;
}

extern "C" int prt_adjoint(prt_renderer* r, const float* target_rgb, const float* image_rgb, uint32_t spp,
                           float* d_albedo, float* d_emission, float* loss) {
    clear_error();
    if (!r || !target_rgb || !d_albedo || !d_emission || spp == 0) return fail(PRT_ERR_INVALID_ARG, "bad argument");
    const size_t n = size_t(r->w) * r->h * 3;
    std::vector<float> img(n), grad(n);
    int rc = PRT_OK;
    if (image_rgb) {
        std::memcpy(img.data(), image_rgb, n * sizeof(float));
    } else {
        // Forward pass at the same seed: the renderer's accumulation is replaced.
        rc = prt_renderer_reset(r);
        if (rc == PRT_OK) rc = prt_render_spp(r, spp);
        if (rc == PRT_OK) rc = prt_get_radiance(r, img.data());
        if (rc != PRT_OK) return rc;
    }
    // d/dI of sum (I - T)^2 / n
    double l = 0.0;
    for (size_t k = 0; k < n; ++k) {
        const double d = double(img[k]) - double(target_rgb[k]);
        l += d * d;
        grad[k] = float(2.0 * d / double(n));
    }
    if (loss) *loss = float(l / double(n));
    return prt_adjoint_grad(r, grad.data(), float(2.0 / double(n)), spp, d_albedo, d_emission);
}
