// Megakernel path tracer, primary-hit kernel and output kernels for gfx950.
//
// One wave64 workgroup renders one 8x8 pixel tile; each lane follows one pixel's
// paths for `nspp` consecutive sample indices and adds them to the running sum in
// sample order, the order SequentialRenderer::renderOneSamplePerPixel
// (src/renderer/sequential_renderer.cu:91-116) and renderKernel
// (src/renderer/parallel_renderer.cu:318-342) accumulate in.
#include "path.hpp"

namespace prt {
namespace dev {

// samplePixel (src/renderer/renderer.cu:119-262) for one pixel-sample.
template <bool COUNT>
__device__ V3 sample_path(const SceneArgs& sc, const CamArgs& cam, const PathArgs& pa, int x, int y, uint32_t sample,
                          Stack& st, Counts& cnt) {
    Rng rng{pa.seed, pa.view, uint32_t(y) * uint32_t(cam.xpixels) + uint32_t(x), sample};
    V3 ro, rd;
    camera_ray(cam, x, y, rng.block(0), ro, rd);
    PathState ps;
    path_init(ps, ro, rd);
    while (true) {
        if (COUNT) {
            if (ps.b == 0) cnt.primary++; else cnt.extension++;
        }
        HitRec h;
        h.t = kFltMax;
        h.id = 0;
        h.rec = 0;
        h.mat = 0;
        h.u = h.v = 0.f;
        if (!traverse<false, COUNT>(sc, ps.ro, ps.rd, kEps, h, st, cnt)) {
            path_miss(ps);
            break;
        }
        bool want_nee;
        NeeSample ns;
        const bool more = shade_vertex<COUNT, false>(sc, pa, rng, ps, h, &want_nee, ns, cnt);
        if (want_nee) {
            ShadowReq sr;
            sr.direct = v3(0.f, 0.f, 0.f);
            bool visible = false;
            if (nee_prepare<COUNT>(sc, ns, sr, cnt)) {
                HitRec occ;
                occ.t = sr.tL;
                occ.id = sr.lid;
                occ.rec = 0;
                occ.mat = 0;
                occ.u = occ.v = 0.f;
                visible = !traverse<true, COUNT>(sc, sr.so, sr.sd, kEps, occ, st, cnt);
            }
            ps.color = add_direct(ps.color, ns.mask, sr.direct, visible);
        }
        if (!more) break;
    }
    return ps.color;
}

template <bool COUNT>
__device__ __forceinline__ void flush_counts(unsigned long long* out, const Counts& c) {
    // Wave-level sums, one atomic per wave and slot.
    unsigned long long v[6] = {c.primary, c.extension, c.shadow, c.nodes, c.tris, c.hits};
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        unsigned long long x = v[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        if ((threadIdx.x & 63) == 0 && x) atomicAdd(out + k, x);
    }
}

template <bool COUNT>
__global__ __launch_bounds__(kBlock, 4) void megakernel(MegaArgs a) {
    __shared__ int s_stack[kStackLds * kBlock];
    const uint32_t tile = xcd_tile(blockIdx.x, a.tiles);
    if (tile >= a.tiles) return;
    const int lane = int(threadIdx.x);
    const int x = int(tile % a.tiles_x) * 8 + (lane & 7);
    const int y = int(tile / a.tiles_x) * 8 + (lane >> 3);
    Counts cnt = {0, 0, 0, 0, 0, 0};
    if (x < int(a.width) && y < int(a.height)) {
        Stack st;
        st.bind(s_stack, a.spill + size_t(blockIdx.x) * kBlock, a.spill_stride);
        const size_t idx = size_t(y) * a.width + size_t(x);
        float4 acc = a.accum[idx];
        V3 sum = v3(acc.x, acc.y, acc.z);
        for (uint32_t k = 0; k < a.nspp; ++k) {
            V3 L = sample_path<COUNT>(a.sc, a.cam, a.path, x, y, a.s0 + k, st, cnt);
            sum = sum + L;
            if (a.samples && a.s0 + k < a.samples_cap)
                a.samples[size_t(a.s0 + k) * a.width * a.height + idx] = make_float4(L.x, L.y, L.z, 0.f);
        }
        a.accum[idx] = make_float4(sum.x, sum.y, sum.z, 0.f);
    }
    if (COUNT) flush_counts<COUNT>(a.counters, cnt);
}

__global__ __launch_bounds__(kBlock) void first_hit_kernel(FirstHitArgs a) {
    __shared__ int s_stack[kStackLds * kBlock];
    const uint32_t tile = xcd_tile(blockIdx.x, a.tiles);
    if (tile >= a.tiles) return;
    const int lane = int(threadIdx.x);
    const int x = int(tile % a.tiles_x) * 8 + (lane & 7);
    const int y = int(tile / a.tiles_x) * 8 + (lane >> 3);
    if (x >= int(a.width) || y >= int(a.height)) return;
    Stack st;
    st.bind(s_stack, a.spill + size_t(blockIdx.x) * kBlock, a.spill_stride);
    Rng rng{a.path.seed, a.path.view, uint32_t(y) * uint32_t(a.cam.xpixels) + uint32_t(x), a.sample};
    V3 ro, rd;
    camera_ray(a.cam, x, y, rng.block(0), ro, rd);
    HitRec h;
    h.t = kFltMax;
    h.id = 0;
    h.rec = 0;
    h.mat = 0;
    h.u = h.v = 0.f;
    Counts cnt;
    bool hit = traverse<false, false>(a.sc, ro, rd, kEps, h, st, cnt);
    const size_t idx = size_t(y) * a.width + size_t(x);
    if (a.tri) a.tri[idx] = hit ? int32_t(h.id) : -1;
    if (a.mat) a.mat[idx] = hit ? int32_t(h.mat) : -1;
    if (a.t) a.t[idx] = hit ? h.t : kFltMax;
}

// float3ToUchar4 then gammaCorrectPixel (renderer.h:149-156, renderer.cu:495-504)
// of accum / samples; negative and NaN map to 0.
__global__ void tonemap_kernel(const float4* accum, uint32_t n, uint32_t samples, const uint8_t* lut, uchar4* out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 a = accum[i];
    float s = float(samples);
    float c[3] = {a.x / s, a.y / s, a.z / s};
    uint8_t b[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float v = c[k];
        uint8_t q = 0;
        if (v > 0.0f) q = uint8_t((v > 1.0f ? 1.0f : v) * 255.f);
        b[k] = lut[q];
    }
    out[i] = make_uchar4(b[0], b[1], b[2], 255);
}

__global__ void mean_kernel(const float4* accum, uint32_t n, uint32_t samples, float* out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 a = accum[i];
    float s = float(samples);
    out[size_t(i) * 3] = a.x / s;
    out[size_t(i) * 3 + 1] = a.y / s;
    out[size_t(i) * 3 + 2] = a.z / s;
}

}  // namespace dev

hipError_t launch_megakernel(const MegaArgs& a, bool count, hipStream_t s) {
    const uint32_t grid = ((a.tiles + 7) / 8) * 8;
    if (count) hipLaunchKernelGGL(dev::megakernel<true>, dim3(grid), dim3(kBlock), 0, s, a);
    else hipLaunchKernelGGL(dev::megakernel<false>, dim3(grid), dim3(kBlock), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_first_hit(const FirstHitArgs& a, hipStream_t s) {
    const uint32_t grid = ((a.tiles + 7) / 8) * 8;
    hipLaunchKernelGGL(dev::first_hit_kernel, dim3(grid), dim3(kBlock), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_tonemap(const float4* accum, uint32_t n, float, uint32_t samples, const uint8_t* lut, uint8_t* out,
                          hipStream_t s) {
    hipLaunchKernelGGL(dev::tonemap_kernel, dim3((n + 255) / 256), dim3(256), 0, s, accum, n, samples, lut,
                       reinterpret_cast<uchar4*>(out));
    return hipGetLastError();
}

hipError_t launch_mean(const float4* accum, uint32_t n, uint32_t samples, float* out, hipStream_t s) {
    hipLaunchKernelGGL(dev::mean_kernel, dim3((n + 255) / 256), dim3(256), 0, s, accum, n, samples, out);
    return hipGetLastError();
}

}  // namespace prt
