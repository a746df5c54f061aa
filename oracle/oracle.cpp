/*
 * ORACLE — test infrastructure only (see oracle_api.h).
 *
 * A plain, scalar CPU restatement of CudaRT's per-pixel path tracer, templated on
 * the arithmetic type: float is the parity oracle and CPU baseline, double is the
 * high-precision reference used for image RMSE. It shares no code with the HIP
 * product: its BVH follows the reference builder (src/geometry/bvh.cpp) and its
 * traversal the reference stack loop (src/renderer/renderer.cu:283-330).
 *
 * Deviations from the reference, each one a fix of undefined or racy behaviour:
 *   - RNG: Philox4x32-10 with a fixed slot per draw instead of curand/rand() (F8).
 *   - BVH: centroid/bucket bounds initialised, real surface area in the SAH (F6),
 *     unbounded traversal stack (F7), node boxes padded so that box rejection can
 *     never hide a triangle the intersection test would accept.
 *   - Closest hit: ties in t go to the lowest original triangle id (Q13), which
 *     makes the result independent of BVH shape and traversal order.
 *   - Light list in original triangle order; total light area summed in that order.
 *   - Tonemap clamps negative/NaN to 0 (Q12).
 * Everything else, including the estimator quirks Q1-Q11, follows the reference.
 *
 * How it is pinned: tests/test_oracle_kat.py checks every part against published
 * vectors (Philox), closed forms and whole-render known answers in which each
 * branch of samplePixel (emission, DIFFSPEC, REFRACTIVE with the Fresnel quirk,
 * DIFFUSE + NEE, SPECULAR, Russian roulette) has a value fixed by the reference's
 * semantics; tests/test_oracle_mutations.py plants mistakes in copies of this file
 * and requires those checks to catch each one. Parity unpinned: the reference's own
 * random streams, texture filtering and fast-math intrinsics, which are replaced
 * (nondeterministic or undefined in the reference).
 */
#include <omp.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <vector>

#include "oracle_api.h"

namespace {

// ---------------------------------------------------------------------------
// Vector arithmetic with the operator semantics of src/include/linalg.h:12-41
// (component-wise, left to right; normalize = one sqrt then three divides).
// ---------------------------------------------------------------------------
template <class T>
struct V3 {
    T x, y, z;
};
template <class T>
inline V3<T> mk(T x, T y, T z) { return V3<T>{x, y, z}; }
template <class T>
inline V3<T> operator+(const V3<T>& a, const V3<T>& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T>
inline V3<T> operator-(const V3<T>& a, const V3<T>& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T>
inline V3<T> operator*(const V3<T>& a, const V3<T>& b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class T>
inline V3<T> operator/(const V3<T>& a, const V3<T>& b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
template <class T>
inline V3<T> operator*(const V3<T>& a, T b) { return {a.x * b, a.y * b, a.z * b}; }
template <class T>
inline V3<T> operator/(const V3<T>& a, T b) { return {a.x / b, a.y / b, a.z / b}; }
template <class T>
inline T dot(const V3<T>& a, const V3<T>& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T>
inline V3<T> cross(const V3<T>& a, const V3<T>& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T>
inline V3<T> normalize(const V3<T>& v) {
    T n = std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
    return {v.x / n, v.y / n, v.z / n};
}
template <class T>
inline T maxComponent(const V3<T>& v) {
    T m = v.x > v.y ? v.x : v.y;
    return m > v.z ? m : v.z;
}
template <class T>
inline V3<T> load3(const float* p) { return {T(p[0]), T(p[1]), T(p[2])}; }

constexpr float kEps = 0.00001f;   // EPSILON, src/include/geometry.h:9
constexpr float kEps2 = 0.000005f; // EPSILON_2, geometry.h:10
constexpr float kLightsGain = 3.0f;  // LIGHTS_GAIN, src/geometry/scene.cpp:19
constexpr float kSky[3] = {0.0f, 0.025f, 0.05f};  // skybox, renderer.cu:64-70

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw: "Parallel random numbers: as easy
// as 1, 2, 3", SC'11): 10 rounds of two 32x32->64 multiplies, Weyl key schedule.
// ---------------------------------------------------------------------------
void philox(const uint32_t c_in[4], const uint32_t k_in[2], uint32_t out[4]) {
    uint32_t c0 = c_in[0], c1 = c_in[1], c2 = c_in[2], c3 = c_in[3];
    uint32_t k0 = k_in[0], k1 = k_in[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint64_t p0 = uint64_t(0xD2511F53u) * c0;
        uint64_t p1 = uint64_t(0xCD9E8D57u) * c2;
        uint32_t n0 = uint32_t(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = uint32_t(p1);
        uint32_t n2 = uint32_t(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = uint32_t(p0);
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// Fixed draw slots. Counter = (pixel, sample, block, 0), key = (seed, view).
// Block 0: camera {tent x, tent y, lens angle, lens radius}.
// Bounce b, block 1+2b: {lobe choice, bsdf r1, bsdf r2, light pick}.
// Bounce b, block 2+2b: {light point u, light point v, russian roulette, spare}.
// Uniform float = (x >> 8) * 2^-24 in [0, 1).
struct Sampler {
    uint32_t key[2];
    uint32_t pixel, sample;
    uint32_t cur = 0xFFFFFFFFu;
    uint32_t buf[4];
    float draw(uint32_t block, uint32_t slot) {
        if (block != cur) {
            uint32_t ctr[4] = {pixel, sample, block, 0u};
            philox(ctr, key, buf);
            cur = block;
        }
        return float(buf[slot] >> 8) * 0x1p-24f;
    }
};
inline uint32_t blockA(uint32_t b) { return 1u + 2u * b; }
inline uint32_t blockB(uint32_t b) { return 2u + 2u * b; }

// ---------------------------------------------------------------------------
// sin/cos shared by every fp32 path: Cody-Waite reduction by pi/2 with a
// three-part constant (exact first two products for |k| <= 8), then Taylor
// polynomials of degree 9 (sin) and 10 (cos) on [-pi/4, pi/4]. Only + - * and
// floor, so every IEEE-754 binary32 machine gives the same bits.
// ---------------------------------------------------------------------------
void own_sincosf(float x, float* s, float* c) {
    const float kTwoOverPi = 0x1.45f306p-1f;
    const float P1 = 0x1.92p+0f, P2 = 0x1.fb6p-12f, P3 = -0x1.777a5cp-25f;
    float kf = std::floor(x * kTwoOverPi + 0.5f);
    float r = ((x - kf * P1) - kf * P2) - kf * P3;
    float r2 = r * r;
    float sp = 0x1.71de3ap-19f;            // 1/9!
    sp = sp * r2 + -0x1.a01a02p-13f;       // -1/7!
    sp = sp * r2 + 0x1.111112p-7f;         // 1/5!
    sp = sp * r2 + -0x1.555556p-3f;        // -1/3!
    float sr = r + (r * r2) * sp;
    float cp = -0x1.27e4fcp-22f;           // -1/10!
    cp = cp * r2 + 0x1.a01a02p-16f;        // 1/8!
    cp = cp * r2 + -0x1.6c16c2p-10f;       // -1/6!
    cp = cp * r2 + 0x1.555556p-5f;         // 1/4!
    cp = cp * r2 + -0.5f;
    float cr = 1.0f + r2 * cp;
    int q = int(kf) & 3;
    switch (q) {
        case 0: *s = sr; *c = cr; break;
        case 1: *s = cr; *c = -sr; break;
        case 2: *s = -sr; *c = -cr; break;
        default: *s = -cr; *c = sr; break;
    }
}
template <class T>
inline void sincosT(T x, T* s, T* c);
template <>
inline void sincosT<float>(float x, float* s, float* c) { own_sincosf(x, s, c); }
template <>
inline void sincosT<double>(double x, double* s, double* c) { *s = std::sin(x); *c = std::cos(x); }

// ---------------------------------------------------------------------------
// Scene in the reference's shape: Triangle (src/include/geometry.h:30-51) kept in
// BVH leaf order, LinearBVHNode (src/include/bvh.h:18-27), packed materials
// (sequential_renderer.cu:48-64).
// ---------------------------------------------------------------------------
struct Tri {
    float v1[3], e1[3], e2[3];
    float n1[3], n2[3], n3[3];
    float uv1[2], uv2[2], uv3[2];
    float area;
    uint32_t id;   // original triangle index
    uint32_t mat;
};
struct Node {
    float lo[3], hi[3];
    int32_t offset;  // trianglesOffset (leaf) or secondChildOffset (inner)
    int16_t ntris;
    int16_t axis;
};
struct Mat {
    float kd[3], ka[3], ks[3];
    float ns, ni, dcoef;
    int bsdf, tex;
    float ka_mtl[3];
};
struct Tex {
    uint32_t w, h;
    std::vector<float> lin;  // w*h*3 linearised texels
};

}  // namespace

struct oracle_scene {
    std::vector<Tri> tris;
    std::vector<Node> nodes;
    std::vector<uint32_t> lights;  // indices into tris, ordered by original id
    float lightsArea = 0.f;
    std::vector<Mat> mats;
    std::vector<Tex> texs;
    uint32_t maxDepth = 0;
    float texTable[256];
};

namespace {

// ---------------------------------------------------------------------------
// BVH build, following recursiveBuild (src/geometry/bvh.cpp:114-307): binned SAH
// with 8 buckets along the axis of largest centroid extent, leaves of <= 3
// triangles, median nth_element fallback, and child 0 = the upper half.
// ---------------------------------------------------------------------------
struct BBox {
    uint32_t tri;
    float lo[3], hi[3], c[3];
};
struct BuildNode {
    float lo[3], hi[3];
    int child[2] = {-1, -1};
    uint32_t first = 0, count = 0, axis = 0;
};

inline void boxInit(float lo[3], float hi[3]) {
    for (int k = 0; k < 3; ++k) { lo[k] = FLT_MAX; hi[k] = -FLT_MAX; }
}
inline void boxGrow(float lo[3], float hi[3], const float a[3], const float b[3]) {
    for (int k = 0; k < 3; ++k) {
        lo[k] = std::min(lo[k], a[k]);
        hi[k] = std::max(hi[k], b[k]);
    }
}
inline float boxArea(const float lo[3], const float hi[3]) {
    float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    if (dx < 0 || dy < 0 || dz < 0) return 0.f;
    return 2.f * (dx * dy + dy * dz + dz * dx);
}

struct Builder {
    std::vector<BBox>& info;
    std::vector<BuildNode>& out;
    std::vector<uint32_t>& order;
    uint32_t maxDepth = 0;

    int build(uint32_t start, uint32_t end, uint32_t depth) {
        maxDepth = std::max(maxDepth, depth);
        int me = int(out.size());
        out.emplace_back();
        float lo[3], hi[3];
        boxInit(lo, hi);
        for (uint32_t i = start; i < end; ++i) boxGrow(lo, hi, info[i].lo, info[i].hi);
        uint32_t n = end - start;
        auto makeLeaf = [&]() {
            BuildNode& nd = out[me];
            nd.first = uint32_t(order.size());
            nd.count = n;
            for (uint32_t i = start; i < end; ++i) order.push_back(info[i].tri);
            std::memcpy(nd.lo, lo, sizeof lo);
            std::memcpy(nd.hi, hi, sizeof hi);
            return me;
        };
        if (n <= 3) return makeLeaf();
        float clo[3], chi[3];
        boxInit(clo, chi);
        for (uint32_t i = start; i < end; ++i) boxGrow(clo, chi, info[i].c, info[i].c);
        float d[3] = {chi[0] - clo[0], chi[1] - clo[1], chi[2] - clo[2]};
        int dim = (d[0] > d[1] && d[0] > d[2]) ? 0 : (d[1] > d[2] ? 1 : 2);  // maximumExtent
        if (chi[dim] == clo[dim]) return makeLeaf();
        uint32_t mid = (start + end) / 2;
        auto byCenter = [dim](const BBox& a, const BBox& b) { return a.c[dim] < b.c[dim]; };
        auto bucketOf = [&](const BBox& t) {
            float off = t.c[dim] - clo[dim];
            if (chi[dim] > clo[dim]) off /= (chi[dim] - clo[dim]);
            // int(8 * off), clamped to bucket 7; NaN (an extent overflowing to
            // infinity) goes to bucket 0, as converting it to int is undefined.
            if (!(off > 0.f)) return 0;
            int b = off < 1.f ? int(8 * off) : 8;
            return b >= 8 ? 7 : b;
        };
        if (n <= 4) {
            std::nth_element(info.begin() + start, info.begin() + mid, info.begin() + end, byCenter);
        } else {
            int cnt[8] = {0};
            float blo[8][3], bhi[8][3];
            for (int b = 0; b < 8; ++b) boxInit(blo[b], bhi[b]);
            for (uint32_t i = start; i < end; ++i) {
                int b = bucketOf(info[i]);
                cnt[b]++;
                boxGrow(blo[b], bhi[b], info[i].lo, info[i].hi);
            }
            float area = boxArea(lo, hi);
            float best = FLT_MAX;
            int split = 0;
            for (int s = 0; s < 7; ++s) {
                float l0[3], h0[3], l1[3], h1[3];
                boxInit(l0, h0);
                boxInit(l1, h1);
                int c0 = 0, c1 = 0;
                for (int j = 0; j <= s; ++j) { boxGrow(l0, h0, blo[j], bhi[j]); c0 += cnt[j]; }
                for (int j = s + 1; j < 8; ++j) { boxGrow(l1, h1, blo[j], bhi[j]); c1 += cnt[j]; }
                float cost = 1.f + (c0 * boxArea(l0, h0) + c1 * boxArea(l1, h1)) / (area > 0 ? area : 1.f);
                if (cost < best) { best = cost; split = s; }
            }
            auto it = std::partition(info.begin() + start, info.begin() + end,
                                     [&](const BBox& t) { return bucketOf(t) <= split; });
            mid = uint32_t(it - info.begin());
            if (mid == start || mid == end) {
                mid = (start + end) / 2;
                std::nth_element(info.begin() + start, info.begin() + mid, info.begin() + end, byCenter);
            }
        }
        // initInner(dim, build(mid, end), build(start, mid)), bvh.cpp:302-304
        int c0 = build(mid, end, depth + 1);
        int c1 = build(start, mid, depth + 1);
        BuildNode& nd = out[me];
        nd.child[0] = c0;
        nd.child[1] = c1;
        nd.axis = uint32_t(dim);
        for (int k = 0; k < 3; ++k) {
            nd.lo[k] = std::min(out[c0].lo[k], out[c1].lo[k]);
            nd.hi[k] = std::max(out[c0].hi[k], out[c1].hi[k]);
        }
        return me;
    }
};

// flattenBVHTree (bvh.cpp:309-323): depth first, child 0 at i+1.
int flatten(const std::vector<BuildNode>& bn, int i, std::vector<Node>& lin, float pad) {
    int my = int(lin.size());
    lin.emplace_back();
    const BuildNode& b = bn[i];
    for (int k = 0; k < 3; ++k) {
        lin[my].lo[k] = b.lo[k] - pad;
        lin[my].hi[k] = b.hi[k] + pad;
    }
    if (b.child[0] < 0) {
        lin[my].offset = int32_t(b.first);
        lin[my].ntris = int16_t(b.count);
        lin[my].axis = 0;
    } else {
        lin[my].ntris = 0;
        lin[my].axis = int16_t(b.axis);
        flatten(bn, b.child[0], lin, pad);
        int second = flatten(bn, b.child[1], lin, pad);
        lin[my].offset = second;
    }
    return my;
}

// ---------------------------------------------------------------------------
// Ray / triangle / box (src/include/geometry.h:17-27, src/geometry/geometry.cu:38-65,
// src/renderer/renderer.cu:333-373).
// ---------------------------------------------------------------------------
template <class T>
struct Ray {
    V3<T> o, d;
    T tMin = T(kEps);
    T tMax = T(FLT_MAX);
};

// Triangle::intersect: Moller-Trumbore, |det| < 1e-5 rejected, t > 1e-5.
template <class T>
inline bool triIntersect(const Tri& tr, const V3<T>& o, const V3<T>& d, T& t, T& u, T& v) {
    V3<T> e1 = load3<T>(tr.e1), e2 = load3<T>(tr.e2), v1 = load3<T>(tr.v1);
    V3<T> P = cross(d, e2);
    T det = dot(e1, P);
    if (std::fabs(det) < T(kEps)) return false;
    T inv = T(1) / det;
    V3<T> Tv = o - v1;
    u = dot(Tv, P) * inv;
    if (u < T(0) || u > T(1)) return false;
    V3<T> Q = cross(Tv, e1);
    v = dot(d, Q) * inv;
    if (v < T(0) || u + v > T(1)) return false;
    t = dot(e2, Q) * inv;
    return t > T(kEps);
}

// rayIntersectsBox over the infinite line; a NaN slab (0 * inf) is ignored.
template <class T>
inline bool boxHit(const V3<T>& o, const V3<T>& d, const float lo[3], const float hi[3]) {
    T t0 = -T(FLT_MAX), t1 = T(FLT_MAX);
    const T oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
    for (int k = 0; k < 3; ++k) {
        T inv = T(1) / dd[k];
        T tn = (T(lo[k]) - oo[k]) * inv;
        T tf = (T(hi[k]) - oo[k]) * inv;
        if (tn > tf) std::swap(tn, tf);
        t0 = tn > t0 ? tn : t0;
        t1 = tf < t1 ? tf : t1;
        if (t0 > t1) return false;
    }
    return true;
}

template <class T>
struct Hit {
    uint32_t idx = 0;     // index into the leaf-ordered triangle array
    uint32_t id = 0xFFFFFFFFu;  // original id of the current best
    T u = 0, v = 0;
};

// Closest hit over the whole tree: t in (tMin, tMax), ties to the lowest original id.
template <class T>
inline void considerTri(const oracle_scene& sc, uint32_t i, Ray<T>& ray, Hit<T>& h, bool& any) {
    T t, u, v;
    const Tri& tr = sc.tris[i];
    if (!triIntersect(tr, ray.o, ray.d, t, u, v)) return;
    if (!(t > ray.tMin)) return;
    if (t < ray.tMax || (t == ray.tMax && any && tr.id < h.id)) {
        ray.tMax = t;
        h.idx = i;
        h.id = tr.id;
        h.u = u;
        h.v = v;
        any = true;
    }
}

// intersectBVH (renderer.cu:283-330): near child chosen by dirIsNeg[axis].
template <class T>
bool intersectScene(const oracle_scene& sc, Ray<T>& ray, Hit<T>& h) {
    bool any = false;
    const V3<T> inv = {T(1) / ray.d.x, T(1) / ray.d.y, T(1) / ray.d.z};
    const int neg[3] = {inv.x < 0, inv.y < 0, inv.z < 0};
    int stack[512];
    int top = 0, cur = 0;
    while (true) {
        const Node& nd = sc.nodes[size_t(cur)];
        if (boxHit(ray.o, ray.d, nd.lo, nd.hi)) {
            if (nd.ntris > 0) {
                for (int i = 0; i < nd.ntris; ++i) considerTri(sc, uint32_t(nd.offset + i), ray, h, any);
                if (top == 0) break;
                cur = stack[--top];
            } else if (neg[nd.axis]) {
                stack[top++] = cur + 1;
                cur = nd.offset;
            } else {
                stack[top++] = nd.offset;
                cur = cur + 1;
            }
        } else {
            if (top == 0) break;
            cur = stack[--top];
        }
    }
    return any;
}

template <class T>
bool intersectBrute(const oracle_scene& sc, Ray<T>& ray, Hit<T>& h) {
    bool any = false;
    for (uint32_t i = 0; i < sc.tris.size(); ++i) considerTri(sc, i, ray, h, any);
    return any;
}

// Triangle::getNormal (geometry.cu:67-71): normalised barycentric vertex normals, never flipped.
template <class T>
inline V3<T> triNormal(const Tri& tr, T u, T v) {
    T w = T(1) - u - v;
    return normalize(load3<T>(tr.n1) * w + load3<T>(tr.n2) * u + load3<T>(tr.n3) * v);
}

// ---------------------------------------------------------------------------
// Camera::computeCameraRay with tentFilter (src/geometry/camera.cu:13-57).
// ---------------------------------------------------------------------------
template <class T>
inline T tent(T r) {
    return r < T(1) ? std::sqrt(r) - T(1) : T(1) - std::sqrt(T(2) - r);
}
template <class T>
Ray<T> cameraRay(const prt_camera& cam, int i, int j, float u0, float u1, float u2, float u3) {
    T dx = tent<T>(T(2) * T(u0));
    T dy = tent<T>(T(2) * T(u1));
    T ni = T(double((T(i) + dx) / T(cam.xpixels)) - 0.5);
    T nj = T(0.5) - ((T(j) + dy) / T(cam.ypixels));
    V3<T> right = load3<T>(cam.right), up = load3<T>(cam.up), eye = load3<T>(cam.eye);
    V3<T> dir = load3<T>(cam.dir);
    T fov = T(cam.fov), aspect = T(cam.aspect);
    dir = dir + right * fov * aspect * ni;
    dir = dir + up * fov * nj;
    V3<T> origin = eye;
    if (cam.aperture_width > 0.0f) {
        T r1 = T(double(u2) - double(0.5f * 2) * M_PI);
        T r2 = T(u3) - T(0.5);
        T s, c;
        sincosT<T>(r1, &s, &c);
        T rj = c * r2 * T(cam.aperture_width);
        T uj = s * r2 * T(cam.aperture_width);
        origin = origin + ((right * rj) + (up * uj));
        V3<T> focal = eye + dir * T(cam.focus_distance);
        dir = focal - origin;
    }
    Ray<T> r;
    r.o = origin;
    r.d = normalize(dir);
    return r;
}

// ---------------------------------------------------------------------------
// Texture: software bilinear with wrap on linearised RGBA8 (build definition of Q15).
// ---------------------------------------------------------------------------
template <class T>
V3<T> texSample(const Tex& tx, T u, T v) {
    u = u - std::floor(u);
    v = v - std::floor(v);
    T x = u * T(tx.w) - T(0.5), y = v * T(tx.h) - T(0.5);
    T fx = std::floor(x), fy = std::floor(y);
    T ax = x - fx, ay = y - fy;
    int i0 = int(fx), j0 = int(fy);
    auto wrap = [](int a, int n) { a %= n; return a < 0 ? a + n : a; };
    int i1 = wrap(i0 + 1, int(tx.w)), j1 = wrap(j0 + 1, int(tx.h));
    i0 = wrap(i0, int(tx.w));
    j0 = wrap(j0, int(tx.h));
    auto texel = [&](int i, int j) { return load3<T>(&tx.lin[(size_t(j) * tx.w + size_t(i)) * 3]); };
    V3<T> a = texel(i0, j0) * (T(1) - ax) + texel(i1, j0) * ax;
    V3<T> b = texel(i0, j1) * (T(1) - ax) + texel(i1, j1) * ax;
    return a * (T(1) - ay) + b * ay;
}

// ---------------------------------------------------------------------------
// getFresnelReflectance (src/renderer/renderer.cu:464-493), with its operator
// precedence kept: R0 = ((etai - etat)^2 / (etat + etai)) * (etat + etai), and
// c = 1 - dot(t, n) with the unflipped normal, so Re > 1 is possible (IOR 1.5 at
// normal incidence: R0 = 0.25, c = 2, Re = 24.25). Re == 1 marks total internal
// reflection. samplePixel and the oracle_fresnel known-answer entry point both
// call this one function.
// ---------------------------------------------------------------------------
template <class T>
struct Fresnel {
    T Re, Tr;
    V3<T> tdir;
};
template <class T>
Fresnel<T> fresnelReflectance(const V3<T>& inc, const V3<T>& nrm, T ni) {
    Fresnel<T> f{T(1), T(0), V3<T>{0, 0, 0}};
    V3<T> n = nrm;
    T cosi = dot(inc, nrm);
    T etai = T(1), etat = ni;
    if (cosi < T(0)) {
        cosi = -cosi;
    } else {
        std::swap(etai, etat);
        n = nrm * T(-1);
    }
    T eta = etai / etat;
    T k = T(1) - eta * eta * (T(1) - cosi * cosi);
    if (k < T(0)) return f;
    f.tdir = inc * eta + n * (eta * cosi - std::sqrt(k));
    T R0 = (etai - etat) * (etai - etat) / (etat + etai) * (etat + etai);
    T c = T(1) - dot(f.tdir, nrm);
    f.Re = R0 + (T(1) - R0) * c * c * c * c * c;
    f.Tr = T(1) - f.Re;
    return f;
}

// Direction of sampleDiffuseBSDF (renderer.cu:375-393): frame w = n,
// u = normalize(cross(|w.x| > 0.1 ? (0,1,0) : (1,0,0), w)), v = w x u;
// r1 = 2 pi u1 (formed in double from M_PI, then rounded), r2 = u2;
// d = normalize(u cos r1 sqrt r2 + v sin r1 sqrt r2 + w sqrt(1 - r2)).
// samplePixel and the oracle_diffuse_dir known-answer entry point both call it.
template <class T>
V3<T> diffuseDirection(const V3<T>& w, float u1f, float u2f) {
    const T u1 = T(u1f), r2 = T(u2f);
    T r1 = T(2 * M_PI * double(u1));
    T r2sq = std::sqrt(r2);
    V3<T> a = std::fabs(w.x) > T(0.1f) ? V3<T>{0, 1, 0} : V3<T>{1, 0, 0};
    V3<T> uu = normalize(cross(a, w));
    V3<T> vv = cross(w, uu);
    T sr, cr;
    sincosT<T>(r1, &sr, &cr);
    return normalize(uu * cr * r2sq + vv * sr * r2sq + w * std::sqrt(T(1) - r2));
}

// ---------------------------------------------------------------------------
// samplePixel (src/renderer/renderer.cu:119-262).
// ---------------------------------------------------------------------------
struct Counters {
    uint64_t primary = 0, extension = 0, shadow = 0;
};

template <class T>
struct Interaction {
    V3<T> position{0, 0, 0}, normal{0, 0, 0}, inputDirection{0, 0, 0}, outputDirection{0, 0, 0};
    T pdf = 0, u = 0, v = 0;
    uint32_t hitTriIdx = 0;
};

// Path-replay adjoint state of one sample (fp64 whatever the path arithmetic).
struct AdjointSink {
    bool pathwise = false;      // every roulette step as if p >= 1 (no score terms)
    double halfHess = 0;        // (d2 Loss / dI2) / (2 spp^2)
    V3<double> delta{0, 0, 0};  // dLoss / dL of this sample, per channel
    V3<double> rem{0, 0, 0};    // radiance still to be collected on the replayed path
    V3<double> L{0, 0, 0};      // the sample's radiance (primal pass)
    uint32_t prevTies = 0;      // channels at the maximum at the previous roulette step (0: none)
    std::vector<uint32_t> kdVerts;  // materials of the untextured diffuse vertices so far
    double* dkd = nullptr;      // M*3
    double* de = nullptr;       // M
};
template <class T>
inline V3<double> d3(const V3<T>& v) { return {double(v.x), double(v.y), double(v.z)}; }

// Weights of the channels in d log max(mask): 1 for a unique maximum, 1/2 for each
// of several equal maxima (the mean of the one-sided derivatives of max, as a
// central difference measures it).
inline V3<double> rrWeight(uint32_t ties) {
    const int n = int((ties & 1u) != 0) + int((ties & 2u) != 0) + int((ties & 4u) != 0);
    const double w = n == 1 ? 1.0 : 0.5;
    return {(ties & 1u) ? w : 0.0, (ties & 2u) ? w : 0.0, (ties & 4u) ? w : 0.0};
}

template <class T, bool kAdjoint>
V3<T> samplePixel(const oracle_scene& sc, const prt_camera& cam, const oracle_cfg& cfg, int x, int y,
                  uint32_t sample, Counters& cnt, AdjointSink* adj) {
    Sampler smp;
    smp.key[0] = cfg.seed;
    smp.key[1] = cfg.view_id;
    smp.pixel = uint32_t(y) * uint32_t(cam.xpixels) + uint32_t(x);
    smp.sample = sample;
    float c0 = smp.draw(0, 0), c1 = smp.draw(0, 1);
    float c2 = 0.f, c3 = 0.f;
    if (cam.aperture_width > 0.0f) { c2 = smp.draw(0, 2); c3 = smp.draw(0, 3); }
    Ray<T> ray = cameraRay<T>(cam, x, y, c0, c1, c2, c3);

    V3<T> color{0, 0, 0}, mask{1, 1, 1};
    Interaction<T> it;
    int currentBsdf = PRT_DIFFUSE, previousBsdf = PRT_DIFFUSE;
    const T invLA = T(1) / T(sc.lightsArea);
    const V3<T> sky{T(kSky[0]), T(kSky[1]), T(kSky[2])};
    const uint32_t numLights = uint32_t(sc.lights.size());

    for (uint32_t b = 0; b < cfg.max_depth; ++b) {
        if (b == 0) cnt.primary++; else cnt.extension++;
        Hit<T> h;
        if (!intersectScene(sc, ray, h)) {
            color = color + mask * sky;
            break;
        }
        it.hitTriIdx = h.idx;
        it.u = h.u;
        it.v = h.v;
        const Tri& tri = sc.tris[h.idx];
        const uint32_t m = tri.mat;
        const Mat& mat = sc.mats[m];
        bool kdVertex = false;
        if (b == 0 || previousBsdf == PRT_SPECULAR || previousBsdf == PRT_REFRACTIVE) {
            V3<T> e = mask * load3<T>(mat.ka) * invLA;
            color = color + e;
            if (kAdjoint) {
                // ka = (Ka_mtl * 3) * A_L and Ka_mtl = e_m (1,1,1): d/de_m = mask * 3 A_L * invLA.
                const double g = double(kLightsGain) * double(sc.lightsArea) * double(invLA);
                adj->de[m] += dot(adj->delta, d3(mask) * g);
                adj->rem = adj->rem - d3(e);
            }
        }
        it.normal = triNormal(tri, it.u, it.v);
        it.position = ray.o + ray.d * ray.tMax + it.normal * T(kEps2);
        it.outputDirection = normalize(ray.d);
        currentBsdf = mat.bsdf;

        if (currentBsdf == PRT_DIFFSPEC) {
            T p = T(mat.dcoef);
            if (T(smp.draw(blockA(b), 0)) < p) {
                currentBsdf = PRT_DIFFUSE;
                mask = mask * (T(1) / p);
            } else {
                currentBsdf = PRT_SPECULAR;
                mask = mask * (T(1) / (T(1) - p));
            }
        }
        if (currentBsdf == PRT_REFRACTIVE) {
            const Fresnel<T> fr = fresnelReflectance<T>(it.outputDirection, it.normal, T(mat.ni));
            const V3<T> tdir = fr.tdir;
            const T Re = fr.Re, Tr = fr.Tr;
            if (Re == T(1)) {
                currentBsdf = PRT_SPECULAR;
            } else {
                T P = T(0.25) + T(0.5) * Re;
                T RP = Re / P;
                T TP = Tr / (T(1) - P);
                if (T(smp.draw(blockA(b), 0)) > T(0.25)) {
                    it.inputDirection = tdir;
                    mask = mask * load3<T>(mat.ks) * TP;
                } else {
                    mask = mask * RP;
                    currentBsdf = PRT_SPECULAR;
                }
            }
        }
        if (currentBsdf == PRT_DIFFUSE) {
            // sampleDiffuseBSDF (renderer.cu:375-402)
            const float u1 = smp.draw(blockA(b), 1), u2 = smp.draw(blockA(b), 2);
            it.inputDirection = diffuseDirection<T>(it.normal, u1, u2);
            it.pdf = T(0.5);
            T cw = dot(it.inputDirection, it.normal);
            V3<T> kd = load3<T>(mat.kd);
            bool textured = mat.tex >= 0;
            if (textured) {
                T bu = it.u, bv = it.v, bw = T(1) - bu - bv;
                T tu = T(tri.uv1[0]) * bw + T(tri.uv2[0]) * bu + T(tri.uv3[0]) * bv;
                T tv = T(tri.uv1[1]) * bw + T(tri.uv2[1]) * bu + T(tri.uv3[1]) * bv;
                kd = texSample<T>(sc.texs[size_t(mat.tex)], tu, tv);
            }
            V3<T> ds = kd * cw;
            kdVertex = !textured;
            if (kAdjoint && !textured) {
                adj->kdVerts.push_back(m);
                // Everything still to be collected on this path carries kd_c once.
                const double kdc[3] = {mat.kd[0], mat.kd[1], mat.kd[2]};
                const double rc[3] = {adj->rem.x, adj->rem.y, adj->rem.z};
                const double dc[3] = {adj->delta.x, adj->delta.y, adj->delta.z};
                for (int ch = 0; ch < 3; ++ch)
                    if (kdc[ch] != 0.0) adj->dkd[m * 3 + ch] += dc[ch] * rc[ch] / kdc[ch];
            }
            mask = mask * ds / it.pdf;

            // Light choice (renderer.cu:218-221) and estimateDirectLighting (:414-445).
            float pick = smp.draw(blockA(b), 3) * (float(numLights) - .00001f);
            uint32_t li = sc.lights[uint32_t(std::trunc(pick))];
            const Tri& L = sc.tris[li];
            V3<T> lightColor = load3<T>(sc.mats[L.mat].ka);
            V3<T> direct{0, 0, 0};
            if (it.hitTriIdx != li) {
                T lu = T(smp.draw(blockB(b), 0)), lv = T(smp.draw(blockB(b), 1));
                if (lu + lv >= T(1)) {
                    lu = T(1) - lu;
                    lv = T(1) - lv;
                }
                V3<T> pt = load3<T>(L.v1) + load3<T>(L.e1) * lu + load3<T>(L.e2) * lv;
                Ray<T> sray;
                sray.o = it.position + it.normal * T(kEps);
                sray.d = normalize(pt - it.position);
                cnt.shadow++;
                Hit<T> lh;
                bool hit = intersectScene(sc, sray, lh);
                if (hit && lh.idx == li) {
                    T sa = T(L.area);
                    T d2 = sray.tMax * sray.tMax;
                    T cosT = std::fabs(dot(triNormal(L, lh.u, lh.v), sray.d * T(-1)));
                    T wf = sa / (d2 * T(sc.lightsArea)) * cosT;
                    direct = direct + lightColor * wf;
                    if (kAdjoint) {
                        const double g = double(kLightsGain) * double(sc.lightsArea) * double(wf);
                        adj->de[L.mat] += dot(adj->delta, d3(mask) * g);
                    }
                }
            }
            V3<T> nee = mask * direct;
            color = color + nee;
            if (kAdjoint) adj->rem = adj->rem - d3(nee);
        }
        if (currentBsdf == PRT_SPECULAR) {
            // sampleSpecularBSDF / reflect (renderer.cu:404-412)
            V3<T> o = it.outputDirection, n = it.normal;
            it.inputDirection = o - n * dot(o, n) * T(2);
            it.pdf = T(1);
            mask = mask * load3<T>(mat.ks) / it.pdf;
        }
        previousBsdf = currentBsdf;
        ray.o = it.position + it.inputDirection * T(kEps);
        ray.d = it.inputDirection;
        ray.tMin = T(kEps);
        ray.tMax = T(FLT_MAX);
        T p = T(0);
        uint32_t ties = 0;
        if (b >= cfg.rr_start) {
            p = maxComponent(mask);
            ties = (mask.x == p ? 1u : 0u) | (mask.y == p ? 2u : 0u) | (mask.z == p ? 4u : 0u);
            if (T(smp.draw(blockB(b), 2)) > p) break;
            mask = mask / p;
        }
        if (kAdjoint && b >= cfg.rr_start) {
            // Survived roulette step j. R_j = L - C_j (collected after the step)
            // carries 1/p_j; with S_k[c] = sum over diffuse vertices v <= k of
            // d log kd_{m_v,c}, d log p_j = w_j . S_j - w_{j-1} . S_{j-1}, and the step
            // adds -kappa_j d log p_j: kappa_j = delta . R_j where p_j >= 1 (no random
            // decision: mask /= p_j is a deterministic rescale); where p_j < 1 the
            // score term of the survival probability, weighted by the loss change of
            // the decision (delta . R_j - halfHess |R_j|^2), leaves
            // kappa_j = halfHess |R_j|^2.
            const V3<double> R = adj->L - d3(color);
            const double kappa =
                (p >= T(1) || adj->pathwise) ? dot(adj->delta, R) : adj->halfHess * dot(R, R);
            const size_t before = adj->kdVerts.size() - (kdVertex ? 1 : 0);
            if (kappa != 0.0) {
                const V3<double> wj = rrWeight(ties), wp = rrWeight(adj->prevTies);
                for (size_t i = 0; i < adj->kdVerts.size(); ++i) {
                    const uint32_t mv = adj->kdVerts[i];
                    const V3<double> cf = i < before ? wp - wj : wj * -1.0;
                    const double c3[3] = {cf.x, cf.y, cf.z};
                    for (int ch = 0; ch < 3; ++ch) {
                        const double kdc = sc.mats[mv].kd[ch];
                        if (c3[ch] != 0.0 && kdc != 0.0) adj->dkd[mv * 3 + ch] += kappa * c3[ch] / kdc;
                    }
                }
            }
            adj->prevTies = ties;
        }
    }
    return color;
}

template <class T>
int renderImpl(const oracle_scene* s, const prt_camera* cam, uint32_t w, uint32_t h, uint32_t s0,
               uint32_t n, const oracle_cfg* cfg, T* accum, uint64_t counters[3], int threads) {
    if (!s || !cam || !cfg || !accum || w == 0 || h == 0) return PRT_ERR_INVALID_ARG;
    if (s->lights.empty()) return PRT_ERR_NO_EMITTERS;
    if (int64_t(w) != cam->xpixels || int64_t(h) != cam->ypixels) return PRT_ERR_INVALID_ARG;
    uint64_t cp = 0, ce = 0, cs = 0;
    int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt) reduction(+ : cp, ce, cs)
    for (int64_t y = 0; y < int64_t(h); ++y) {
        Counters c;
        for (uint32_t x = 0; x < w; ++x) {
            size_t idx = (size_t(y) * w + x) * 3;
            V3<T> acc{accum[idx], accum[idx + 1], accum[idx + 2]};
            for (uint32_t k = 0; k < n; ++k)
                acc = acc + samplePixel<T, false>(*s, *cam, *cfg, int(x), int(y), s0 + k, c, nullptr);
            accum[idx] = acc.x;
            accum[idx + 1] = acc.y;
            accum[idx + 2] = acc.z;
        }
        cp += c.primary;
        ce += c.extension;
        cs += c.shadow;
    }
    if (counters) {
        counters[0] += cp;
        counters[1] += ce;
        counters[2] += cs;
    }
    return PRT_OK;
}

inline uint8_t toByte(float v) {
    // float3ToUchar4 (renderer.h:149-156) with negative/NaN clamped to 0.
    if (!(v > 0.0f)) return 0;
    return uint8_t((v > 1.0f ? 1.0f : v) * 255.f);
}
inline uint8_t gammaByte(uint8_t b) {
    // gammaCorrectPixel (renderer.cu:495-504)
    float invGamma = 1.f / 2.2f;
    float f = std::pow(float(b) * 1.f / 255.f, invGamma);
    return uint8_t(f * 255.f);
}

template <class T>
void adjointPixels(const oracle_scene* s, const prt_camera* cam, uint32_t w, uint32_t h, uint32_t spp,
                   const oracle_cfg* cfg, const float* image, const float* target, std::vector<double>& gk,
                   std::vector<double>& ge, int nt, double hess, bool pathwise) {
    const size_t M = s->mats.size();
    const double norm = 2.0 / (double(w) * double(h) * 3.0 * double(spp));
#pragma omp parallel num_threads(nt)
    {
        int tid = omp_get_thread_num();
        AdjointSink sink;
        sink.pathwise = pathwise;
        sink.halfHess = hess / (2.0 * double(spp) * double(spp));
        sink.dkd = &gk[size_t(tid) * M * 3];
        sink.de = &ge[size_t(tid) * M];
        Counters c;
#pragma omp for schedule(dynamic, 1)
        for (int64_t y = 0; y < int64_t(h); ++y)
            for (uint32_t x = 0; x < w; ++x) {
                size_t idx = (size_t(y) * w + x) * 3;
                V3<double> delta{norm * (double(image[idx]) - double(target[idx])),
                                 norm * (double(image[idx + 1]) - double(target[idx + 1])),
                                 norm * (double(image[idx + 2]) - double(target[idx + 2]))};
                for (uint32_t k = 0; k < spp; ++k) {
                    // primal pass, then the replay with the same random numbers
                    V3<T> L = samplePixel<T, false>(*s, *cam, *cfg, int(x), int(y), k, c, nullptr);
                    sink.delta = delta;
                    sink.L = d3(L);
                    sink.rem = d3(L);
                    sink.prevTies = 0;
                    sink.kdVerts.clear();
                    samplePixel<T, true>(*s, *cam, *cfg, int(x), int(y), k, c, &sink);
                }
            }
    }
}

}  // namespace

// ===========================================================================
// C entry points
// ===========================================================================
template <class T>
int renderPixelsImpl(const oracle_scene* s, const prt_camera* cam, uint32_t w, uint32_t h, const uint32_t* pixels,
                     uint32_t count, uint32_t s0, uint32_t n, const oracle_cfg* cfg, T* rgb, uint64_t counters[3],
                     int threads) {
    if (!s || !cam || !cfg || !pixels || !rgb) return PRT_ERR_INVALID_ARG;
    if (s->lights.empty()) return PRT_ERR_NO_EMITTERS;
    if (int64_t(w) != cam->xpixels || int64_t(h) != cam->ypixels) return PRT_ERR_INVALID_ARG;
    for (uint32_t i = 0; i < count; ++i)
        if (pixels[i] >= w * h) return PRT_ERR_INVALID_ARG;
    uint64_t cp = 0, ce = 0, cs = 0;
    int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt) reduction(+ : cp, ce, cs)
    for (int64_t i = 0; i < int64_t(count); ++i) {
        Counters c;
        int x = int(pixels[i] % w), y = int(pixels[i] / w);
        V3<T> acc{0, 0, 0};
        for (uint32_t k = 0; k < n; ++k) acc = acc + samplePixel<T, false>(*s, *cam, *cfg, x, y, s0 + k, c, nullptr);
        rgb[i * 3] = acc.x;
        rgb[i * 3 + 1] = acc.y;
        rgb[i * 3 + 2] = acc.z;
        cp += c.primary;
        ce += c.extension;
        cs += c.shadow;
    }
    if (counters) {
        counters[0] += cp;
        counters[1] += ce;
        counters[2] += cs;
    }
    return PRT_OK;
}

extern "C" {

int oracle_scene_create(const prt_scene_desc* d, oracle_scene** out) {
    if (!d || !out || d->num_triangles == 0 || !d->positions || !d->normals || !d->material_ids ||
        d->num_materials == 0 || d->num_materials > PRT_MAX_MATERIALS || !d->materials)
        return PRT_ERR_INVALID_ARG;
    // Non-finite vertex data is refused, as by prt_scene_create.
    for (size_t k = 0; k < size_t(d->num_triangles) * 9; ++k)
        if (!std::isfinite(d->positions[k]) || !std::isfinite(d->normals[k])) return PRT_ERR_INVALID_ARG;
    if (d->uvs)
        for (size_t k = 0; k < size_t(d->num_triangles) * 6; ++k)
            if (!std::isfinite(d->uvs[k])) return PRT_ERR_INVALID_ARG;
    oracle_scene* s = new (std::nothrow) oracle_scene;
    if (!s) return PRT_ERR_OUT_OF_MEMORY;
    const uint32_t n = d->num_triangles;
    // stbi__ldr_to_hdr (src/include/stb_image.h:1682) compiled as C++: pow(float, float).
    for (int i = 0; i < 256; ++i) s->texTable[i] = std::pow(float(i) / 255.0f, 2.2f) * 1.0f;
    // Textures: RGBA8 -> linear float3 (stb ldr->hdr, gamma 2.2).
    s->texs.resize(d->num_textures);
    for (uint32_t t = 0; t < d->num_textures; ++t) {
        const prt_texture& pt = d->textures[t];
        Tex& tx = s->texs[t];
        tx.w = pt.width;
        tx.h = pt.height;
        tx.lin.resize(size_t(tx.w) * tx.h * 3);
        for (size_t p = 0; p < size_t(tx.w) * tx.h; ++p)
            for (int k = 0; k < 3; ++k) tx.lin[p * 3 + k] = s->texTable[pt.rgba[p * 4 + k]];
    }
    // Triangles (Scene::loadTriangles, scene.cpp:74-138).
    std::vector<Tri> tris(n);
    float absMax = 1.f;
    for (uint32_t i = 0; i < n; ++i) {
        Tri& t = tris[i];
        const float* p = d->positions + size_t(i) * 9;
        const float* q = d->normals + size_t(i) * 9;
        for (int k = 0; k < 3; ++k) {
            t.v1[k] = p[k];
            t.e1[k] = p[3 + k] - p[k];
            t.e2[k] = p[6 + k] - p[k];
            t.n1[k] = q[k];
            t.n2[k] = q[3 + k];
            t.n3[k] = q[6 + k];
            for (int v = 0; v < 3; ++v) absMax = std::max(absMax, std::fabs(p[3 * v + k]));
        }
        for (int k = 0; k < 2; ++k) {
            t.uv1[k] = d->uvs ? d->uvs[size_t(i) * 6 + k] : 0.f;
            t.uv2[k] = d->uvs ? d->uvs[size_t(i) * 6 + 2 + k] : 0.f;
            t.uv3[k] = d->uvs ? d->uvs[size_t(i) * 6 + 4 + k] : 0.f;
        }
        V3<float> c = cross(load3<float>(t.e1), load3<float>(t.e2));
        t.area = std::sqrt(c.x * c.x + c.y * c.y + c.z * c.z) / 2.0f;
        t.id = i;
        t.mat = d->material_ids[i];
        if (t.mat >= d->num_materials) { delete s; return PRT_ERR_INVALID_ARG; }
    }
    // Materials: materialFromMtl (scene.cpp:175-190).
    s->mats.resize(d->num_materials);
    for (uint32_t m = 0; m < d->num_materials; ++m) {
        const prt_material& pm = d->materials[m];
        Mat& mt = s->mats[m];
        for (int k = 0; k < 3; ++k) {
            mt.kd[k] = pm.kd[k];
            mt.ka_mtl[k] = pm.ka[k];
            mt.ka[k] = pm.ka[k] * kLightsGain;
            mt.ks[k] = pm.ks[k];
        }
        mt.ns = pm.ns;
        mt.ni = pm.ni;
        mt.dcoef = pm.diffuse_coef;
        mt.bsdf = pm.bsdf;
        mt.tex = pm.tex_kd;
        if (mt.tex >= int(d->num_textures)) { delete s; return PRT_ERR_INVALID_ARG; }
    }
    // Light list and total area in original order (constructLightList, scene.cpp:148-173).
    std::vector<uint32_t> lightIds;
    float area = 0.f;
    for (uint32_t i = 0; i < n; ++i)
        if (s->mats[tris[i].mat].bsdf == PRT_EMISSIVE) {
            lightIds.push_back(i);
            area += tris[i].area;
        }
    s->lightsArea = area;
    for (auto& mt : s->mats)
        for (int k = 0; k < 3; ++k) mt.ka[k] = mt.ka[k] * area;
    // BVH.
    std::vector<BBox> info(n);
    for (uint32_t i = 0; i < n; ++i) {
        const float* p = d->positions + size_t(i) * 9;
        BBox& b = info[i];
        b.tri = i;
        for (int k = 0; k < 3; ++k) {
            b.lo[k] = std::min(p[k], std::min(p[3 + k], p[6 + k]));
            b.hi[k] = std::max(p[k], std::max(p[3 + k], p[6 + k]));
            b.c[k] = (b.lo[k] + b.hi[k]) * 0.5f;
        }
    }
    std::vector<BuildNode> bn;
    bn.reserve(size_t(n) * 2);
    std::vector<uint32_t> order;
    order.reserve(n);
    Builder bld{info, bn, order};
    bld.build(0, n, 0);
    if (bld.maxDepth >= 500) { delete s; return PRT_ERR_LIMIT; }
    s->maxDepth = bld.maxDepth;
    s->nodes.reserve(bn.size());
    flatten(bn, 0, s->nodes, absMax * 0x1p-16f);
    s->tris.resize(n);
    std::vector<uint32_t> where(n);
    for (uint32_t k = 0; k < n; ++k) {
        s->tris[k] = tris[order[k]];
        where[order[k]] = k;
    }
    for (uint32_t id : lightIds) s->lights.push_back(where[id]);
    *out = s;
    return PRT_OK;
}

int oracle_scene_destroy(oracle_scene* s) {
    delete s;
    return PRT_OK;
}

int oracle_scene_info(const oracle_scene* s, uint32_t* num_nodes, uint32_t* max_depth, uint32_t* num_lights,
                      float* lights_area) {
    if (!s) return PRT_ERR_INVALID_ARG;
    if (num_nodes) *num_nodes = uint32_t(s->nodes.size());
    if (max_depth) *max_depth = s->maxDepth;
    if (num_lights) *num_lights = uint32_t(s->lights.size());
    if (lights_area) *lights_area = s->lightsArea;
    return PRT_OK;
}

int oracle_scene_export_bvh(const oracle_scene* s, float* bounds, int32_t* meta, uint32_t* tri_order) {
    if (!s) return PRT_ERR_INVALID_ARG;
    for (size_t i = 0; i < s->nodes.size(); ++i) {
        const Node& nd = s->nodes[i];
        if (bounds)
            for (int k = 0; k < 3; ++k) {
                bounds[i * 6 + k] = nd.lo[k];
                bounds[i * 6 + 3 + k] = nd.hi[k];
            }
        if (meta) {
            meta[i * 3] = nd.offset;
            meta[i * 3 + 1] = nd.ntris;
            meta[i * 3 + 2] = nd.axis;
        }
    }
    if (tri_order)
        for (size_t i = 0; i < s->tris.size(); ++i) tri_order[i] = s->tris[i].id;
    return PRT_OK;
}

int oracle_render_f32(const oracle_scene* s, const prt_camera* cam, uint32_t w, uint32_t h, uint32_t s0,
                      uint32_t n, const oracle_cfg* cfg, float* accum, uint64_t counters[3], int threads) {
    return renderImpl<float>(s, cam, w, h, s0, n, cfg, accum, counters, threads);
}

int oracle_render_f64(const oracle_scene* s, const prt_camera* cam, uint32_t w, uint32_t h, uint32_t s0,
                      uint32_t n, const oracle_cfg* cfg, double* accum, uint64_t counters[3], int threads) {
    return renderImpl<double>(s, cam, w, h, s0, n, cfg, accum, counters, threads);
}

int oracle_render_pixels_f32(const oracle_scene* s, const prt_camera* cam, uint32_t w, uint32_t h,
                             const uint32_t* pixels, uint32_t count, uint32_t s0, uint32_t n,
                             const oracle_cfg* cfg, float* rgb, uint64_t counters[3], int threads) {
    return renderPixelsImpl<float>(s, cam, w, h, pixels, count, s0, n, cfg, rgb, counters, threads);
}

int oracle_render_pixels_f64(const oracle_scene* s, const prt_camera* cam, uint32_t w, uint32_t h,
                             const uint32_t* pixels, uint32_t count, uint32_t s0, uint32_t n,
                             const oracle_cfg* cfg, double* rgb, uint64_t counters[3], int threads) {
    return renderPixelsImpl<double>(s, cam, w, h, pixels, count, s0, n, cfg, rgb, counters, threads);
}


int oracle_first_hit(const oracle_scene* s, const prt_camera* cam, uint32_t w, uint32_t h, uint32_t sample,
                     const oracle_cfg* cfg, int32_t* tri, int32_t* mat, float* t, int brute_force, int threads) {
    if (!s || !cam || !cfg || w == 0 || h == 0) return PRT_ERR_INVALID_ARG;
    if (int64_t(w) != cam->xpixels || int64_t(h) != cam->ypixels) return PRT_ERR_INVALID_ARG;
    int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
    for (int64_t y = 0; y < int64_t(h); ++y)
        for (uint32_t x = 0; x < w; ++x) {
            Sampler smp;
            smp.key[0] = cfg->seed;
            smp.key[1] = cfg->view_id;
            smp.pixel = uint32_t(y) * w + x;
            smp.sample = sample;
            float c0 = smp.draw(0, 0), c1 = smp.draw(0, 1), c2 = 0.f, c3 = 0.f;
            if (cam->aperture_width > 0.0f) { c2 = smp.draw(0, 2); c3 = smp.draw(0, 3); }
            Ray<float> r = cameraRay<float>(*cam, int(x), int(y), c0, c1, c2, c3);
            Hit<float> hh;
            bool hit = brute_force ? intersectBrute(*s, r, hh) : intersectScene(*s, r, hh);
            size_t idx = size_t(y) * w + x;
            if (tri) tri[idx] = hit ? int32_t(s->tris[hh.idx].id) : -1;
            if (mat) mat[idx] = hit ? int32_t(s->tris[hh.idx].mat) : -1;
            if (t) t[idx] = hit ? r.tMax : FLT_MAX;
        }
    return PRT_OK;
}

int oracle_trace(const oracle_scene* s, uint32_t n, const float* orig, const float* dir, int32_t* tri, float* t,
                 float* u, float* v, int brute_force, int threads) {
    if (!s || !orig || !dir) return PRT_ERR_INVALID_ARG;
    int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) num_threads(nt)
    for (int64_t i = 0; i < int64_t(n); ++i) {
        Ray<float> r;
        r.o = load3<float>(orig + i * 3);
        r.d = load3<float>(dir + i * 3);
        Hit<float> hh;
        bool hit = brute_force ? intersectBrute(*s, r, hh) : intersectScene(*s, r, hh);
        if (tri) tri[i] = hit ? int32_t(s->tris[hh.idx].id) : -1;
        if (t) t[i] = hit ? r.tMax : FLT_MAX;
        if (u) u[i] = hit ? hh.u : 0.f;
        if (v) v[i] = hit ? hh.v : 0.f;
    }
    return PRT_OK;
}

int oracle_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    if (!ctr || !key || !out) return PRT_ERR_INVALID_ARG;
    philox(ctr, key, out);
    return PRT_OK;
}

int oracle_sincosf(uint32_t n, const float* x, float* s, float* c) {
    if (!x || !s || !c) return PRT_ERR_INVALID_ARG;
    for (uint32_t i = 0; i < n; ++i) own_sincosf(x[i], &s[i], &c[i]);
    return PRT_OK;
}

int oracle_triangle_intersect(const float v1[3], const float e1[3], const float e2[3], const float o[3],
                              const float d[3], float* t, float* u, float* v) {
    Tri tr{};
    for (int k = 0; k < 3; ++k) {
        tr.v1[k] = v1[k];
        tr.e1[k] = e1[k];
        tr.e2[k] = e2[k];
    }
    float tt = FLT_MAX, uu = 0, vv = 0;
    bool hit = triIntersect<float>(tr, load3<float>(o), load3<float>(d), tt, uu, vv);
    *t = hit ? tt : FLT_MAX;
    *u = uu;
    *v = vv;
    return hit ? 1 : 0;
}

int oracle_box_hit(const float lo[3], const float hi[3], const float o[3], const float d[3]) {
    return boxHit<float>(load3<float>(o), load3<float>(d), lo, hi) ? 1 : 0;
}

int oracle_fresnel(const float wo[3], const float nn[3], float ior, float* re, float* tr, float tdir[3]) {
    if (!wo || !nn || !re || !tr || !tdir) return PRT_ERR_INVALID_ARG;
    const Fresnel<float> f = fresnelReflectance<float>(load3<float>(wo), load3<float>(nn), ior);
    *re = f.Re;
    *tr = f.Tr;
    tdir[0] = f.tdir.x; tdir[1] = f.tdir.y; tdir[2] = f.tdir.z;
    return PRT_OK;
}

int oracle_diffuse_dir(const float nn[3], float u1, float u2, float dir[3], float* cos_weight) {
    if (!nn || !dir || !cos_weight) return PRT_ERR_INVALID_ARG;
    const V3<float> w = load3<float>(nn);
    const V3<float> d = diffuseDirection<float>(w, u1, u2);
    dir[0] = d.x; dir[1] = d.y; dir[2] = d.z;
    *cos_weight = dot(d, w);
    return PRT_OK;
}

int oracle_camera_ray(const prt_camera* cam, int32_t x, int32_t y, float u0, float u1, float u2, float u3,
                      float o[3], float d[3]) {
    if (!cam) return PRT_ERR_INVALID_ARG;
    Ray<float> r = cameraRay<float>(*cam, x, y, u0, u1, u2, u3);
    o[0] = r.o.x; o[1] = r.o.y; o[2] = r.o.z;
    d[0] = r.d.x; d[1] = r.d.y; d[2] = r.d.z;
    return PRT_OK;
}

int oracle_tonemap(uint32_t n, const float* rgb, uint8_t* rgba) {
    if (!rgb || !rgba) return PRT_ERR_INVALID_ARG;
    for (uint32_t i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) rgba[i * 4 + k] = gammaByte(toByte(rgb[i * 3 + k]));
        rgba[i * 4 + 3] = 255;
    }
    return PRT_OK;
}

int oracle_texture_sample(const prt_texture* tex, uint32_t n, const double* uv, double* rgb, int fp64) {
    if (!tex || !tex->rgba || tex->width == 0 || tex->height == 0 || !uv || !rgb) return PRT_ERR_INVALID_ARG;
    Tex tx;
    tx.w = tex->width;
    tx.h = tex->height;
    tx.lin.resize(size_t(tx.w) * tx.h * 3);
    for (size_t p = 0; p < size_t(tx.w) * tx.h; ++p)
        for (int k = 0; k < 3; ++k) tx.lin[p * 3 + k] = std::pow(float(tex->rgba[p * 4 + k]) / 255.0f, 2.2f) * 1.0f;
    for (uint32_t i = 0; i < n; ++i) {
        if (fp64) {
            V3<double> c = texSample<double>(tx, uv[2 * i], uv[2 * i + 1]);
            rgb[3 * i] = c.x, rgb[3 * i + 1] = c.y, rgb[3 * i + 2] = c.z;
        } else {
            V3<float> c = texSample<float>(tx, float(uv[2 * i]), float(uv[2 * i + 1]));
            rgb[3 * i] = c.x, rgb[3 * i + 1] = c.y, rgb[3 * i + 2] = c.z;
        }
    }
    return PRT_OK;
}

int oracle_texture_table(float out[256]) {
    for (int i = 0; i < 256; ++i) out[i] = std::pow(float(i) / 255.0f, 2.2f) * 1.0f;
    return PRT_OK;
}

int oracle_adjoint(const oracle_scene* s, const prt_camera* cam, uint32_t w, uint32_t h, uint32_t spp,
                   const oracle_cfg* cfg, const float* image, const float* target, double* d_albedo,
                   double* d_emission, double hess, int pathwise, int fp64, int threads) {
    if (!s || !cam || !cfg || !image || !target || !d_albedo || !d_emission || spp == 0) return PRT_ERR_INVALID_ARG;
    if (s->lights.empty()) return PRT_ERR_NO_EMITTERS;
    if (int64_t(w) != cam->xpixels || int64_t(h) != cam->ypixels) return PRT_ERR_INVALID_ARG;
    const size_t M = s->mats.size();
    int nt = threads > 0 ? threads : omp_get_max_threads();
    std::vector<double> gk(size_t(nt) * M * 3, 0.0), ge(size_t(nt) * M, 0.0);
    if (fp64) adjointPixels<double>(s, cam, w, h, spp, cfg, image, target, gk, ge, nt, hess, pathwise != 0);
    else adjointPixels<float>(s, cam, w, h, spp, cfg, image, target, gk, ge, nt, hess, pathwise != 0);
    for (size_t i = 0; i < M * 3; ++i) {
        double a = 0;
        for (int t = 0; t < nt; ++t) a += gk[size_t(t) * M * 3 + i];
        d_albedo[i] = a;
    }
    for (size_t i = 0; i < M; ++i) {
        double a = 0;
        for (int t = 0; t < nt; ++t) a += ge[size_t(t) * M + i];
        d_emission[i] = a;
    }
    return PRT_OK;
}

int oracle_scene_set_material_params(oracle_scene* s, uint32_t index, const float kd[3], float emission) {
    if (!s || index >= s->mats.size() || !kd) return PRT_ERR_INVALID_ARG;
    Mat& m = s->mats[index];
    for (int k = 0; k < 3; ++k) {
        m.kd[k] = kd[k];
        m.ka_mtl[k] = emission;
        m.ka[k] = (emission * kLightsGain) * s->lightsArea;
    }
    return PRT_OK;
}

}  // extern "C"
