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
__device__ __forceinline__ uint4 philox(uint4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint32_t lo0 = 0xD2511F53u * c.x, hi0 = __umulhi(0xD2511F53u, c.x);
        uint32_t lo1 = 0xCD9E8D57u * c.z, hi1 = __umulhi(0xCD9E8D57u, c.z);
        c = make_uint4(hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0);
    }
    return c;
}
__device__ __forceinline__ float u01(uint32_t x) { return float(x >> 8) * 0x1p-24f; }

// Fixed draw slots: counter (pixel, sample, block, 0), key (seed, view).
// block 0 camera; bounce b: block 1+2b {lobe, r1, r2, light pick},
// block 2+2b {light u, light v, roulette, spare}.
struct Rng {
    uint32_t k0, k1, pixel, sample;
    __device__ __forceinline__ uint4 block(uint32_t b) const { return philox(make_uint4(pixel, sample, b, 0u), k0, k1); }
};

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
template <int LDS>
struct StackT {
    static constexpr int kLds = LDS;
    int* lds;              // this lane's column: lds[k * kBlock]
    int* spill;            // this thread's column in the spill area
    uint32_t spill_stride; // threads sharing the spill area
    int sp;
    __device__ __forceinline__ void push(int v) {
        if (sp < LDS) lds[sp * kBlock] = v;
        else spill[size_t(sp - LDS) * spill_stride] = v;
        ++sp;
    }
    __device__ __forceinline__ int pop() {
        --sp;
        return sp < LDS ? lds[sp * kBlock] : spill[size_t(sp - LDS) * spill_stride];
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

// Closest hit (ANY = false) or occlusion test against a known best (ANY = true).
// Accepts t in (tmin, best.t) or t == best.t with a lower original id: the
// closest hit is the lexicographic minimum of (t, original id), independent of
// tree shape and visiting order. Returns true when best changed (closest) or
// when an occluder was found (any).
template <bool ANY, bool COUNT, class StackType>
__device__ __forceinline__ bool traverse(const SceneArgs& sc, V3 o, V3 d, float tmin, HitRec& best, StackType& st,
                                         Counts& cnt) {
    bool found = false;
    if (d.x == 0.f && d.y == 0.f && d.z == 0.f) return false;  // det == 0 for every triangle
    const float ix = 1.f / d.x, iy = 1.f / d.y, iz = 1.f / d.z;
    const float ox = o.x * ix, oy = o.y * iy, oz = o.z * iz;
    const int base = st.sp;
    int node = 0;
    while (true) {
        if (node >= 0) {
            const float4* np = sc.nodes + size_t(node) * 4;
            float4 n0 = np[0], n1 = np[1], n2 = np[2];
            int4 n3 = reinterpret_cast<const int4*>(np)[3];
            if (COUNT) cnt.nodes++;
            float a0 = __builtin_fmaf(n0.x, ix, -ox), b0 = __builtin_fmaf(n0.y, ix, -ox);
            float a1 = __builtin_fmaf(n0.z, iy, -oy), b1 = __builtin_fmaf(n0.w, iy, -oy);
            float a2 = __builtin_fmaf(n2.x, iz, -oz), b2 = __builtin_fmaf(n2.y, iz, -oz);
            float tn0 = fmaxf(fmaxf(fminf(a0, b0), fminf(a1, b1)), fmaxf(fminf(a2, b2), tmin));
            float tf0 = fminf(fminf(fmaxf(a0, b0), fmaxf(a1, b1)), fminf(fmaxf(a2, b2), best.t));
            float c0 = __builtin_fmaf(n1.x, ix, -ox), d0 = __builtin_fmaf(n1.y, ix, -ox);
            float c1 = __builtin_fmaf(n1.z, iy, -oy), d1 = __builtin_fmaf(n1.w, iy, -oy);
            float c2 = __builtin_fmaf(n2.z, iz, -oz), d2 = __builtin_fmaf(n2.w, iz, -oz);
            float tn1 = fmaxf(fmaxf(fminf(c0, d0), fminf(c1, d1)), fmaxf(fminf(c2, d2), tmin));
            float tf1 = fminf(fminf(fmaxf(c0, d0), fmaxf(c1, d1)), fminf(fmaxf(c2, d2), best.t));
            bool h0 = tn0 <= tf0, h1 = tn1 <= tf1;
            if (h0 && h1) {
                bool first0 = tn0 <= tn1;
                st.push(first0 ? n3.y : n3.x);
                node = first0 ? n3.x : n3.y;
                continue;
            }
            if (h0 || h1) {
                node = h0 ? n3.x : n3.y;
                continue;
            }
        } else {
            const uint32_t w = uint32_t(node);
            const uint32_t first = w & 0x0FFFFFFFu;
            const uint32_t count = ((w >> 28) & 7u) + 1u;
            for (uint32_t k = 0; k < count; ++k) {
                const float4* tp = sc.tris + size_t(first + k) * 3;
                float4 a = tp[0], b = tp[1], c = tp[2];
                if (COUNT) cnt.tris++;
                float t, u, v;
                if (!mt_intersect(a, b, c, o, d, t, u, v)) continue;
                uint32_t id = __float_as_uint(a.w);
                if (!(t > tmin)) continue;
                if (t < best.t || (t == best.t && id < best.id)) {
                    if (ANY) {
                        st.sp = base;
                        return true;
                    }
                    best.t = t;
                    best.u = u;
                    best.v = v;
                    best.rec = first + k;
                    best.id = id;
                    best.mat = __float_as_uint(b.w);
                    found = true;
                }
            }
        }
        if (st.sp == base) break;
        node = st.pop();
    }
    return found;
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
