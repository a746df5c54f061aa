// Device-side building blocks of the path tracer for gfx950 (wave64).
//
// Arithmetic that decides which triangle is hit and what radiance a path carries
// is written operation by operation in the order of the reference
// (src/renderer/renderer.cu, src/geometry/camera.cu, src/geometry/geometry.cu,
// src/include/linalg.h) and compiled with -ffp-contract=off, correctly rounded
// division and square root, so the result is the same IEEE-754 value the CPU
// restatement computes. Box tests only cull, so they use FMA freely.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernel_api.hpp"

namespace prt {
namespace dev {

constexpr float kEps = 0.00001f;   // EPSILON, src/include/geometry.h:9
constexpr float kEps2 = 0.000005f; // EPSILON_2, geometry.h:10
constexpr float kFltMax = 3.402823466e+38f;

struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 v3(float4 a) { return V3{a.x, a.y, a.z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ V3 operator*(V3 a, float b) { return {a.x * b, a.y * b, a.z * b}; }
__device__ __forceinline__ V3 operator/(V3 a, float b) { return {a.x / b, a.y / b, a.z / b}; }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ V3 normalize(V3 v) {
    float n = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
    return {v.x / n, v.y / n, v.z / n};
}
__device__ __forceinline__ float max_component(V3 v) {
    float m = v.x > v.y ? v.x : v.y;
    return m > v.z ? m : v.z;
}
__device__ __forceinline__ V3 ld3(const float* p) { return V3{p[0], p[1], p[2]}; }

// ---- Philox4x32-10 (Salmon et al., SC'11) --------------------------------
// WIDE: each round's two 32 x 32 -> 64-bit products are taken whole (one
// v_mad_u64_u32 each) rather than as separate low and high halves (v_mul_lo_u32 +
// v_mul_hi_u32): fewer instructions (wf_generate, bound by instruction issue, 3 %
// faster), but the 64-bit results hold two more registers. Same bits either way.
template <bool WIDE = true>
__device__ __forceinline__ uint4 philox(uint4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        if (WIDE) {
            const uint64_t p0 = uint64_t(0xD2511F53u) * c.x, p1 = uint64_t(0xCD9E8D57u) * c.z;
            c = make_uint4(uint32_t(p1 >> 32) ^ c.y ^ k0, uint32_t(p1), uint32_t(p0 >> 32) ^ c.w ^ k1, uint32_t(p0));
        } else {
            const uint32_t lo0 = 0xD2511F53u * c.x, hi0 = __umulhi(0xD2511F53u, c.x);
            const uint32_t lo1 = 0xCD9E8D57u * c.z, hi1 = __umulhi(0xCD9E8D57u, c.z);
            c = make_uint4(hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0);
        }
    }
    return c;
}
__device__ __forceinline__ float u01(uint32_t x) { return float(x >> 8) * 0x1p-24f; }

// Fixed draw slots: counter (pixel, sample, block, 0), key (seed, view).
// block 0 camera; bounce b: block 1+2b {lobe, r1, r2, light pick},
// block 2+2b {light u, light v, roulette, spare}.
template <bool WIDE = true>
struct RngT {
    uint32_t k0, k1, pixel, sample;
    __device__ __forceinline__ uint4 block(uint32_t b) const {
        return philox<WIDE>(make_uint4(pixel, sample, b, 0u), k0, k1);
    }
};
using Rng = RngT<true>;

// ---- sin/cos: Cody-Waite by pi/2 + Taylor, same bits on host and device ----
__device__ __forceinline__ void own_sincosf(float x, float* s, float* c) {
    const float kTwoOverPi = 0x1.45f306p-1f;
    const float P1 = 0x1.92p+0f, P2 = 0x1.fb6p-12f, P3 = -0x1.777a5cp-25f;
    float kf = floorf(x * kTwoOverPi + 0.5f);
    float r = ((x - kf * P1) - kf * P2) - kf * P3;
    float r2 = r * r;
    float sp = 0x1.71de3ap-19f;
    sp = sp * r2 + -0x1.a01a02p-13f;
    sp = sp * r2 + 0x1.111112p-7f;
    sp = sp * r2 + -0x1.555556p-3f;
    float sr = r + (r * r2) * sp;
    float cp = -0x1.27e4fcp-22f;
    cp = cp * r2 + 0x1.a01a02p-16f;
    cp = cp * r2 + -0x1.6c16c2p-10f;
    cp = cp * r2 + 0x1.555556p-5f;
    cp = cp * r2 + -0.5f;
    float cr = 1.0f + r2 * cp;
    int q = int(kf) & 3;
    float s0 = (q & 1) ? cr : sr;
    float c0 = (q & 1) ? sr : cr;
    *s = (q & 2) ? -s0 : s0;
    *c = ((q + 1) & 2) ? -c0 : c0;
}

// ---- camera (Camera::computeCameraRay + tentFilter, src/geometry/camera.cu:13-57)
__device__ __forceinline__ float tent(float r) { return r < 1.0f ? sqrtf(r) - 1.0f : 1.0f - sqrtf(2.0f - r); }

__device__ __forceinline__ void camera_ray(const CamArgs& cam, int i, int j, uint4 cb, V3& o, V3& d) {
    float dx = tent(2.0f * u01(cb.x));
    float dy = tent(2.0f * u01(cb.y));
    float ni = float(double((float(i) + dx) / float(cam.xpixels)) - 0.5);
    float nj = 0.5f - ((float(j) + dy) / float(cam.ypixels));
    V3 right = ld3(cam.right), up = ld3(cam.up), eye = ld3(cam.eye);
    V3 dir = ld3(cam.dir);
    dir = dir + right * cam.fov * cam.aspect * ni;
    dir = dir + up * cam.fov * nj;
    V3 origin = eye;
    if (cam.aperture_width > 0.0f) {
        float r1 = float(double(u01(cb.z)) - 1.0 * M_PI);
        float r2 = u01(cb.w) - 0.5f;
        float s, c;
        own_sincosf(r1, &s, &c);
        float rj = c * r2 * cam.aperture_width;
        float uj = s * r2 * cam.aperture_width;
        origin = origin + ((right * rj) + (up * uj));
        V3 focal = eye + dir * cam.focus_distance;
        dir = focal - origin;
    }
    o = origin;
    d = normalize(dir);
}

// ---- counters ------------------------------------------------------------
struct Counts {
    uint32_t primary, extension, shadow, nodes, tris, hits;
};

// ---- traversal -----------------------------------------------------------
// Short stack in LDS (kStackLds entries per lane, lane-interleaved so any mix of
// depths is bank-conflict free), continued in a global spill area.
// Explicit address spaces: scene arrays reached through a pointer loaded from
// memory are otherwise generic (flat) accesses, which take 64-bit VGPR addresses
// and wait on the LDS counter as well.
template <class T>
using gptr = const __attribute__((address_space(1))) T*;
// p[i] as a global-memory load (the host compilation pass sees a plain load).
template <class T>
__device__ __forceinline__ T gload(const T* p, size_t i) {
#if defined(__HIP_DEVICE_COMPILE__)
    return ((gptr<T>)p)[i];
#else
    return p[i];
#endif
}
using lds_int = __attribute__((address_space(3))) int;
using glob_int = __attribute__((address_space(1))) int;

template <int LDS>
struct StackT {
    static constexpr int kLds = LDS;
    lds_int* lds;          // the workgroup's LDS stack block; lane l uses lds[k * kBlock + l]
    glob_int* spill;       // the workgroup's spill block; lane l uses spill[k * stride + l]
    uint32_t spill_stride; // threads sharing the spill area
    int sp;
    // Both bases are workgroup-uniform; the lane offset is added per access.
    __device__ __forceinline__ void bind(int* lds_block, int* spill_block, uint32_t stride) {
        lds = (lds_int*)lds_block;
        spill = (glob_int*)spill_block;
        spill_stride = stride;
        sp = 0;
    }
    __device__ __forceinline__ void push(int v) {
        const uint32_t lane = threadIdx.x;
        if (sp < LDS) lds[sp * kBlock + lane] = v;
        else spill[size_t(sp - LDS) * spill_stride + lane] = v;
        ++sp;
    }
    __device__ __forceinline__ int pop() {
        const uint32_t lane = threadIdx.x;
        --sp;
        int v;
        if (sp < LDS) v = lds[sp * kBlock + lane];
        else v = spill[size_t(sp - LDS) * spill_stride + lane];
        return v;
    }
};
using Stack = StackT<kStackLds>;

struct HitRec {
    float t, u, v;
    uint32_t rec;  // record index (BVH order)
    uint32_t id;   // original triangle id, 0 = none yet
    uint32_t mat;  // material of the record
};

// Moller-Trumbore exactly as Triangle::intersect (src/geometry/geometry.cu:38-65).
__device__ __forceinline__ bool mt_intersect(float4 a, float4 b, float4 c, V3 o, V3 d, float& t, float& u, float& v) {
    V3 v1 = v3(a), e1 = v3(b), e2 = v3(c);
    V3 P = cross(d, e2);
    float det = dot(e1, P);
    if (fabsf(det) < kEps) return false;
    float inv = 1.f / det;
    V3 T = o - v1;
    u = dot(T, P) * inv;
    if (u < 0.f || u > 1.f) return false;
    V3 Q = cross(T, e1);
    v = dot(d, Q) * inv;
    if (v < 0.f || u + v > 1.f) return false;
    t = dot(e2, Q) * inv;
    return t > kEps;
}

// One four-wide node (64 bytes, quantized child boxes): decodes each child plane as
// fmaf(q, 2^e, origin) (exactly as the builder bounded it; with FOLD the decode is
// folded into the slab FMA), slab-tests the four boxes against [tmin, tfar], and
// orders the children hit nearest first. Pushes the farther ones (farthest first,
// so the nearest of them is popped next) and returns the nearest, or kNoChild.
//
// The near and far plane of each axis are picked once per node from the sign of
// the inverse direction, instead of a min and a max per child and axis: for i > 0,
// q_lo <= q_hi gives t(q_lo) <= t(q_hi) (each rounding is monotonic), so the picked
// planes are the values min / max would give. i is finite and nonzero
// (slab_inverse), so no plane is NaN for a finite origin.
constexpr int kNoChild = 0x7FFFFFFF;  // also the word of an empty slot
__device__ __forceinline__ void cswap(float& ka, int& wa, float& kb, int& wb) {
    const bool sw = kb < ka;
    const float k = sw ? kb : ka;
    kb = sw ? ka : kb;
    ka = k;
    const int w = sw ? wb : wa;
    wb = sw ? wa : wb;
    wa = w;
}
__device__ __forceinline__ float qbyte(uint32_t w, int c) { return float((w >> (8 * c)) & 0xFFu); }
// FOLD: each plane's decode folded into its slab FMA, t = q (2^e i) + (origin i - o i),
// the two factors made once per node (6 FMAs instead of 12 per child; see
// packet_traverse for why culling stays conservative). Occlusion rays use it
// (wf_shadow 2 % faster); closest-hit rays do not (wf_extend 3 % slower with it).
template <bool FOLD, class StackType>
__device__ __forceinline__ int visit_node(const float4* nodes, int node, float ix, float iy, float iz, float ox,
                                          float oy, float oz, float tmin, float tfar, StackType& st) {
    const size_t nb = size_t(node) * 4;
    const uint4* un = reinterpret_cast<const uint4*>(nodes);
    const float4 org = gload(nodes, nb);
    const uint4 w = gload(un, nb + 1), qxy = gload(un, nb + 2), qz = gload(un, nb + 3);
    const uint32_t ex = __float_as_uint(org.w);
    const float sx = __uint_as_float((ex & 0xFFu) << 23), sy = __uint_as_float(((ex >> 8) & 0xFFu) << 23),
                sz = __uint_as_float(((ex >> 16) & 0xFFu) << 23);
    float key[4];
    int word[4] = {int(w.x), int(w.y), int(w.z), int(w.w)};
    const bool fx = ix < 0.f, fy = iy < 0.f, fz = iz < 0.f;  // the hi plane is the near one
    const uint32_t nqx = fx ? qxy.y : qxy.x, fqx = fx ? qxy.x : qxy.y;
    const uint32_t nqy = fy ? qxy.w : qxy.z, fqy = fy ? qxy.z : qxy.w;
    const uint32_t nqz = fz ? qz.y : qz.x, fqz = fz ? qz.x : qz.y;
    const float ax = sx * ix, ay = sy * iy, az = sz * iz;
    const float bx = __builtin_fmaf(org.x, ix, -ox), by = __builtin_fmaf(org.y, iy, -oy), bz = __builtin_fmaf(org.z, iz, -oz);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        float tn, tf;
        if (FOLD) {
            tn = fmaxf(fmaxf(__builtin_fmaf(qbyte(nqx, c), ax, bx), __builtin_fmaf(qbyte(nqy, c), ay, by)),
                       fmaxf(__builtin_fmaf(qbyte(nqz, c), az, bz), tmin));
            tf = fminf(fminf(__builtin_fmaf(qbyte(fqx, c), ax, bx), __builtin_fmaf(qbyte(fqy, c), ay, by)),
                       fminf(__builtin_fmaf(qbyte(fqz, c), az, bz), tfar));
        } else {
            const float nx = __builtin_fmaf(qbyte(nqx, c), sx, org.x), fx_ = __builtin_fmaf(qbyte(fqx, c), sx, org.x);
            const float ny = __builtin_fmaf(qbyte(nqy, c), sy, org.y), fy_ = __builtin_fmaf(qbyte(fqy, c), sy, org.y);
            const float nz = __builtin_fmaf(qbyte(nqz, c), sz, org.z), fz_ = __builtin_fmaf(qbyte(fqz, c), sz, org.z);
            tn = fmaxf(fmaxf(__builtin_fmaf(nx, ix, -ox), __builtin_fmaf(ny, iy, -oy)),
                       fmaxf(__builtin_fmaf(nz, iz, -oz), tmin));
            tf = fminf(fminf(__builtin_fmaf(fx_, ix, -ox), __builtin_fmaf(fy_, iy, -oy)),
                       fminf(__builtin_fmaf(fz_, iz, -oz), tfar));
        }
        const bool hit = tn <= tf && word[c] != kNoChild;
        key[c] = hit ? tn : __builtin_inff();
        word[c] = hit ? word[c] : kNoChild;
        // One child at a time: interleaving the four decodes would hold all 24
        // decoded planes live (register pressure, hence occupancy).
        __builtin_amdgcn_sched_barrier(0);
    }
    // Occlusion rays use the same order: the nearest boxes hold occluders most often.
    cswap(key[0], word[0], key[1], word[1]);
    cswap(key[2], word[2], key[3], word[3]);
    cswap(key[0], word[0], key[2], word[2]);
    cswap(key[1], word[1], key[3], word[3]);
    cswap(key[1], word[1], key[2], word[2]);
    if (word[3] != kNoChild) st.push(word[3]);
    if (word[2] != kNoChild) st.push(word[2]);
    if (word[1] != kNoChild) st.push(word[1]);
    return word[0];
}

// Resumable form of traverse() for kernels that refill finished lanes with new
// rays: the same culling and (t, id) ordering, one wave-wide phase per call.
//
// Leaves are postponed (Aila & Laine, "Understanding the efficiency of ray
// traversal on GPUs", HPG 2009): a lane that reaches a leaf sets it aside and
// goes on visiting inner nodes; it stops only at a second leaf. The caller picks
// the phase for the whole wave: a node phase while some lane has no leaf set
// aside, else a leaf phase, in which every lane tests its leaf. Node decoding and
// triangle tests then never share a step, so a step waits on one kind of load.
// A set-aside leaf is tested after nodes that a tested leaf might have culled;
// that costs visits, never results: the (t, id) minimum does not depend on order.
// Passing node_phase = trav_wants_nodes(r) per lane gives the unphased order: a
// lane tests each leaf at the step after it reaches it.
//
// trav_begin() sets up a ray; trav_step() returns kTravRunning, kTravDone
// (closest hit: best holds the result; any-hit: no occluder) or kTravOccluded
// (any-hit only).
struct TravRay {
    V3 o, d;
    float ix, iy, iz;  // o * i is recomputed per node (the same fp32 products as traverse's)
    int node;          // the inner node to visit next, a second leaf, or kNoChild
    int leaf;          // the leaf set aside for the next leaf phase, or kNoChild
};
enum { kTravRunning = 0, kTravDone = 1, kTravOccluded = 2 };

// Inverse direction component for the slab tests. A component below 2^-64 in
// magnitude (exactly 0 happens: a diffuse sample with r1 or r2 = 0 on an
// axis-aligned surface) is taken as +-2^-64: with 1/0 = inf, p * i - o * i is
// inf - inf = NaN, which either drops the slab (the ray then visits most of the
// tree) or, compared against a finite plane, culls a box the ray runs inside.
// Tilted by 2^-64 the ray moves less than 2^-40 over t <= 2^24, far inside the
// boxes' padding (2^-16 of the largest scene coordinate), so every hit closer
// than 2^24 stays inside its boxes' slabs. The triangle tests use d itself.
__device__ __forceinline__ float slab_inverse(float c) {
    return fabsf(c) < 0x1p-64f ? copysignf(0x1p64f, c) : 1.f / c;
}

// Returns false when the ray cannot hit anything: a zero direction (det == 0 for
// every triangle) or a NaN component (every comparison of the triangle test fails).
__device__ __forceinline__ bool trav_begin(TravRay& r, V3 o, V3 d) {
    r.o = o;
    r.d = d;
    r.ix = slab_inverse(d.x);
    r.iy = slab_inverse(d.y);
    r.iz = slab_inverse(d.z);
    r.node = 0;
    r.leaf = kNoChild;
    const bool nan = d.x != d.x || d.y != d.y || d.z != d.z;
    return !nan && !(d.x == 0.f && d.y == 0.f && d.z == 0.f);
}

// True for a lane that a node phase can advance without a leaf set aside.
__device__ __forceinline__ bool trav_wants_nodes(const TravRay& r) { return r.leaf == kNoChild; }

// LEAN keeps only t, rec and id of the best hit (u, v and the material are not
// written; hit_details() recomputes them from the record afterwards).
// `any` (default ANY) may also be set per lane at run time, for kernels whose lanes
// trace closest-hit and occlusion rays in the same steps.
// FOLD: the box test visit_node<FOLD> uses (by default the folded one for occlusion
// rays only).
template <bool ANY, bool COUNT, bool LEAN, bool FOLD = ANY, class StackType>
__device__ __forceinline__ int trav_step(bool node_phase, const float4* nodes, const float4* tris, TravRay& r,
                                         float tmin, HitRec& best, StackType& st, Counts& cnt, bool any = ANY) {
    if (node_phase) {
        if (uint32_t(r.node) < uint32_t(kNoChild)) {  // an inner node
            if (COUNT) cnt.nodes++;
            int next = visit_node<FOLD>(nodes, r.node, r.ix, r.iy, r.iz, r.o.x * r.ix, r.o.y * r.iy, r.o.z * r.iz,
                                        tmin, best.t, st);
            if (next == kNoChild && st.sp != 0) next = st.pop();
            r.node = next;
        }
    } else if (r.leaf != kNoChild) {
        const uint32_t w = uint32_t(r.leaf);
        const uint32_t first = w & 0x0FFFFFFFu;
        const uint32_t count = ((w >> 28) & 7u) + 1u;
        r.leaf = kNoChild;
        for (uint32_t k = 0; k < count; ++k) {
            const size_t tb = size_t(first + k) * 3;
            float4 a = gload(tris, tb), b = gload(tris, tb + 1), c = gload(tris, tb + 2);
            if (COUNT) cnt.tris++;
            float t, u, v;
            if (!mt_intersect(a, b, c, r.o, r.d, t, u, v)) continue;
            uint32_t id = __float_as_uint(a.w);
            if (!(t > tmin)) continue;
            if (t < best.t || (t == best.t && id < best.id)) {
                if (any) {
                    st.sp = 0;
                    r.node = kNoChild;
                    return kTravOccluded;
                }
                best.t = t;
                best.rec = first + k;
                best.id = id;
                if (!LEAN) {
                    best.u = u;
                    best.v = v;
                    best.mat = __float_as_uint(b.w);
                }
            }
        }
    }
    // Set a reached leaf aside (at most one) and continue with the next entry.
    if (r.leaf == kNoChild && r.node < 0) {
        r.leaf = r.node;
        r.node = st.sp != 0 ? st.pop() : kNoChild;
    }
    return (r.leaf == kNoChild && r.node == kNoChild) ? kTravDone : kTravRunning;
}

// Phase choice for a traversal step, over the lanes of the wave that make the call.
// MIN = 0: no phases, each lane tests its leaf at the step after it reaches it.
// MIN >= 1: a node phase while at least MIN lanes still want nodes (no leaf set
// aside), or while no lane holds a leaf; otherwise a leaf phase. Waiting for every
// lane (MIN = 1) costs closest-hit rays culling.
template <uint32_t MIN>
__device__ __forceinline__ bool node_phase_for(bool wants_nodes, bool active) {
    if (MIN == 0) return wants_nodes;
    const unsigned long long want = __ballot(wants_nodes);
    if (MIN == 1) return want != 0ull;
    return uint32_t(__popcll(want)) >= MIN || __ballot(active && !wants_nodes) == 0ull;
}

// traverse()'s threshold. The megakernel and the adjoint call it with the lanes whose
// paths are still alive; on the 1080p 1M-triangle frame, 1 measured 665 ms per frame
// for the megakernel (no phases: 689; 8: 661; 24: 714) and the best adjoint rate.
#ifndef PRT_LOOP_PHASE_MIN
#define PRT_LOOP_PHASE_MIN 1
#endif
static_assert(PRT_LOOP_PHASE_MIN >= 0 && PRT_LOOP_PHASE_MIN <= 64, "phase threshold out of range");

// Closest hit (ANY = false) or occlusion test against a known best (ANY = true),
// run to completion by each calling lane (trav_step in a loop; lanes that finish
// leave the loop and the phases are chosen among the rest). Accepts t in
// (tmin, best.t) or t == best.t with a lower original id: the closest hit is the
// lexicographic minimum of (t, original id), independent of tree shape and
// visiting order. Returns true when best changed (closest) or when an occluder was
// found (any). The stack must be empty on entry, and is empty on return.
template <bool ANY, bool COUNT, class StackType>
__device__ __forceinline__ bool traverse(const SceneArgs& sc, V3 o, V3 d, float tmin, HitRec& best, StackType& st,
                                         Counts& cnt) {
    TravRay r;
    if (!trav_begin(r, o, d)) return false;  // det == 0 for every triangle
    const float t0 = best.t;
    const uint32_t id0 = best.id;
    while (true) {
        const bool node_phase = node_phase_for<PRT_LOOP_PHASE_MIN>(trav_wants_nodes(r), true);
        const int s = trav_step<ANY, COUNT, false>(node_phase, sc.nodes, sc.tris, r, tmin, best, st, cnt);
        if (s == kTravOccluded) return true;
        if (s == kTravDone) break;
    }
    return !ANY && (best.t != t0 || best.id != id0);
}

// ---- wave-wide (packet) traversal -----------------------------------------
// For rays that start together and run close to each other (a wave's 64 camera
// samples of one pixel or pixel block), the wave walks ONE depth-first path
// through the tree: a child is entered when any live lane's ray hits its box, and
// every live lane tests every triangle of every leaf entered. Node and triangle
// indices are then the same in every lane, so their 64 / 48 bytes are read once
// for the wave with scalar loads (scalar data cache, SGPR operands of the box and
// triangle tests) and the stack is a single wave-wide LDS array. Each lane still
// culls with its own [tmin, best.t] and accepts hits by the (t, original id) rule,
// and the leaves entered are a superset of those its own traversal would reach, so
// each lane's closest hit (or occlusion) is the one traverse() finds.
template <class T>
using kptr = const __attribute__((address_space(4))) T*;
// p[i] with a wave-uniform i: one scalar load for the wave.
template <class T>
__device__ __forceinline__ T sload(const T* p, uint32_t i) {
#if defined(__HIP_DEVICE_COMPILE__)
    return ((kptr<T>)p)[i];
#else
    return p[i];
#endif
}
__device__ __forceinline__ uint32_t uniform_u32(uint32_t v) { return uint32_t(__builtin_amdgcn_readfirstlane(int(v))); }

// Entries of the wave-wide stack: at most three per level below the root, and
// scenes are limited to 3 (depth + 1) <= kStackMax.
constexpr int kPacketStack = kStackMax;
constexpr uint32_t kPkNone = 0xFFFFFFFFu;  // sort key of a child no live lane enters

// ANY = false: closest hit into best (t, rec, id, u, v, mat). ANY = true: occlusion
// test against best (t, id), `occluded` set per lane. `live` lanes take part (a
// lane with no ray, or an occluded one, tests nothing); nodes and tris must be the
// same in every live lane. COUNT books each live lane's node and triangle tests.
template <bool ANY, bool COUNT>
__device__ __forceinline__ void packet_traverse(const float4* nodes, const float4* tris, bool live, const TravRay& r,
                                                float tmin, HitRec& best, lds_int* stk, Counts& cnt, bool& occluded) {
    occluded = false;
    unsigned long long lm = __ballot(live);
    if (lm == 0ull) return;
    const float ox = r.o.x * r.ix, oy = r.o.y * r.iy, oz = r.o.z * r.iz;
    const bool fx = r.ix < 0.f, fy = r.iy < 0.f, fz = r.iz < 0.f;  // the hi plane is the near one
    const uint4* un = reinterpret_cast<const uint4*>(nodes);
    int node = 0;
    int sp = 0;
    while (true) {
        if (node >= 0) {  // an inner node
            if (COUNT && live) cnt.nodes++;
            const uint32_t nb = uint32_t(node) * 4u;
            const float4 org = sload(nodes, nb);
            const uint4 w = sload(un, nb + 1), qxy = sload(un, nb + 2), qz = sload(un, nb + 3);
            const uint32_t ex = __float_as_uint(org.w);
            const float sx = __uint_as_float((ex & 0xFFu) << 23), sy = __uint_as_float(((ex >> 8) & 0xFFu) << 23),
                        sz = __uint_as_float(((ex >> 16) & 0xFFu) << 23);
            const uint32_t nqx = fx ? qxy.y : qxy.x, fqx = fx ? qxy.x : qxy.y;
            const uint32_t nqy = fy ? qxy.w : qxy.z, fqy = fy ? qxy.z : qxy.w;
            const uint32_t nqz = fz ? qz.y : qz.x, fqz = fz ? qz.x : qz.y;
            const int rep = __builtin_ctzll(lm);  // the first live lane orders the children
            uint32_t key[4];
            int word[4] = {int(w.x), int(w.y), int(w.z), int(w.w)};
            // Decode and slab test folded into one FMA per plane, t = q (2^e i) + (origin i
            // - o i), with the two factors made once per node: 30 of the visit's ~120 VALU
            // instructions fewer, and this loop is bound by instruction issue (camera-ray
            // packets 8 % faster). The fold rounds differently from decoding the plane
            // first, by about 2^-23 |coordinate * i| in t, far inside the leaf boxes'
            // padding (2^-16 of the largest scene coordinate, times |i| in t), so culling
            // stays conservative. t stays monotonic in q for a fixed sign of i, so the
            // near / far plane choice holds.
            const float ax = sx * r.ix, ay = sy * r.iy, az = sz * r.iz;
            const float bx = __builtin_fmaf(org.x, r.ix, -ox), by = __builtin_fmaf(org.y, r.iy, -oy),
                        bz = __builtin_fmaf(org.z, r.iz, -oz);
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const float tn = fmaxf(fmaxf(__builtin_fmaf(qbyte(nqx, c), ax, bx), __builtin_fmaf(qbyte(nqy, c), ay, by)),
                                       fmaxf(__builtin_fmaf(qbyte(nqz, c), az, bz), tmin));
                const float tf = fminf(fminf(__builtin_fmaf(qbyte(fqx, c), ax, bx), __builtin_fmaf(qbyte(fqy, c), ay, by)),
                                       fminf(__builtin_fmaf(qbyte(fqz, c), az, bz), best.t));
                const bool hit = live && tn <= tf && word[c] != kNoChild;
                // tn >= tmin > 0: the bit patterns of the keys order like the values.
                const uint32_t mine = hit ? __float_as_uint(tn) : 0x7F800000u;
                key[c] = __ballot(hit) != 0ull ? uint32_t(__builtin_amdgcn_readlane(int(mine), rep)) : kPkNone;
            }
            // Nearest first for the representative lane; children only other lanes
            // enter follow, then children no lane enters (kPkNone).
            auto cs = [](uint32_t& ka, int& wa, uint32_t& kb, int& wb) {
                if (kb < ka) {
                    const uint32_t k = ka;
                    ka = kb;
                    kb = k;
                    const int t = wa;
                    wa = wb;
                    wb = t;
                }
            };
            cs(key[0], word[0], key[1], word[1]);
            cs(key[2], word[2], key[3], word[3]);
            cs(key[0], word[0], key[2], word[2]);
            cs(key[1], word[1], key[3], word[3]);
            cs(key[1], word[1], key[2], word[2]);
            if (key[3] != kPkNone) stk[sp++] = word[3];
            if (key[2] != kPkNone) stk[sp++] = word[2];
            if (key[1] != kPkNone) stk[sp++] = word[1];
            if (key[0] != kPkNone) node = word[0];
            else node = sp ? int(uniform_u32(uint32_t(stk[--sp]))) : kNoChild;
        } else {  // a leaf
            const uint32_t wl = uint32_t(node);
            const uint32_t first = wl & 0x0FFFFFFFu;
            const uint32_t count = ((wl >> 28) & 7u) + 1u;
            for (uint32_t k = 0; k < count; ++k) {
                const uint32_t tb = (first + k) * 3u;
                const float4 a = sload(tris, tb), b = sload(tris, tb + 1), c = sload(tris, tb + 2);
                if (COUNT && live) cnt.tris++;
                float t, u, v;
                if (live && mt_intersect(a, b, c, r.o, r.d, t, u, v)) {
                    const uint32_t id = __float_as_uint(a.w);
                    if (t > tmin && (t < best.t || (t == best.t && id < best.id))) {
                        if (ANY) {
                            occluded = true;
                            live = false;
                        } else {
                            best.t = t;
                            best.rec = first + k;
                            best.id = id;
                            best.u = u;
                            best.v = v;
                            best.mat = __float_as_uint(b.w);
                        }
                    }
                }
            }
            if (ANY) {
                lm = __ballot(live);
                if (lm == 0ull) return;
            }
            node = sp ? int(uniform_u32(uint32_t(stk[--sp]))) : kNoChild;
        }
        if (node == kNoChild) return;
    }
}

// u, v and material of best.rec, by the intersection test that selected it
// (identical inputs, so identical results).
__device__ __forceinline__ void hit_details(const float4* tris, const TravRay& r, HitRec& best) {
    const size_t tb = size_t(best.rec) * 3;
    const float4 a = gload(tris, tb), b = gload(tris, tb + 1), c = gload(tris, tb + 2);
    float t, u, v;
    mt_intersect(a, b, c, r.o, r.d, t, u, v);
    best.u = u;
    best.v = v;
    best.mat = __float_as_uint(b.w);
}

// Interpolated shading normal (Triangle::getNormal, geometry.cu:67-71); shading
// record = {n1.xyz n2.x}{n2.yz n3.xy}{n3.z uv1.xy uv2.x}{uv2.y uv3.xy area}.
__device__ __forceinline__ V3 shading_normal(float4 s0, float4 s1, float4 s2, float u, float v) {
    V3 n1 = v3(s0.x, s0.y, s0.z), n2 = v3(s0.w, s1.x, s1.y), n3 = v3(s1.z, s1.w, s2.x);
    float w = 1.f - u - v;
    return normalize(n1 * w + n2 * u + n3 * v);
}

// Software bilinear texture with wrap on linearised RGBA8 texels.
__device__ __forceinline__ V3 tex_sample(const SceneArgs& sc, int tex, float u, float v) {
    const uint4 info = sc.texinfo[tex];  // offset (texels), width, height
    const uint8_t* px = sc.texels + size_t(info.x) * 4;
    const int W = int(info.y), H = int(info.z);
    u = u - floorf(u);
    v = v - floorf(v);
    float x = u * float(W) - 0.5f, y = v * float(H) - 0.5f;
    float fx = floorf(x), fy = floorf(y);
    float ax = x - fx, ay = y - fy;
    int i0 = int(fx), j0 = int(fy);
    int i1 = (i0 + 1) % W, j1 = (j0 + 1) % H;
    i0 %= W;
    j0 %= H;
    if (i1 < 0) i1 += W;
    if (j1 < 0) j1 += H;
    if (i0 < 0) i0 += W;
    if (j0 < 0) j0 += H;
    auto texel = [&](int i, int j) {
        const uint8_t* p = px + (size_t(j) * W + i) * 4;
        return V3{sc.lin[p[0]], sc.lin[p[1]], sc.lin[p[2]]};
    };
    V3 a = texel(i0, j0) * (1.f - ax) + texel(i1, j0) * ax;
    V3 b = texel(i0, j1) * (1.f - ax) + texel(i1, j1) * ax;
    return a * (1.f - ay) + b * ay;
}

}  // namespace dev
}  // namespace prt
