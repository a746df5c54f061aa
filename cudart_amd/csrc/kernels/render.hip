// Megakernel path tracer, primary-hit kernel and output kernels for gfx950.
//
// One wave64 workgroup renders one 8x8 pixel tile; each lane follows one pixel's
// paths for `nspp` consecutive sample indices and adds them to the running sum in
// sample order, the order SequentialRenderer::renderOneSamplePerPixel
// (src/renderer/sequential_renderer.cu:91-116) and renderKernel
// (src/renderer/parallel_renderer.cu:318-342) accumulate in.
#include "device_common.hpp"

namespace prt {
namespace dev {

// samplePixel (src/renderer/renderer.cu:119-262) for one pixel-sample.
template <bool COUNT>
__device__ V3 sample_path(const SceneArgs& sc, const CamArgs& cam, const PathArgs& pa, int x, int y, uint32_t sample,
                          Stack& st, Counts& cnt) {
    Rng rng{pa.seed, pa.view, uint32_t(y) * uint32_t(cam.xpixels) + uint32_t(x), sample};
    V3 ro, rd;
    camera_ray(cam, x, y, rng.block(0), ro, rd);

    V3 color = v3(0.f, 0.f, 0.f), mask = v3(1.f, 1.f, 1.f);
    V3 inputDirection = v3(0.f, 0.f, 0.f);  // SurfaceInteraction() is zeroed; stale across bounces (Q4)
    int previousBsdf = 0;
    const V3 sky = v3(0.0f, 0.025f, 0.05f);
    const float invLA = sc.inv_lights_area;

    for (uint32_t b = 0; b < pa.max_depth; ++b) {
        if (COUNT) {
            if (b == 0) cnt.primary++; else cnt.extension++;
        }
        HitRec h;
        h.t = kFltMax;
        h.id = 0;
        h.rec = 0;
        h.u = h.v = 0.f;
        if (!traverse<false, COUNT>(sc, ro, rd, kEps, h, st, cnt)) {
            color = color + mask * sky;
            break;
        }
        if (COUNT) cnt.hits++;
        const float4* tp = sc.tris + size_t(h.rec) * 3;
        const uint32_t m = __float_as_uint(tp[1].w);
        const float4* mp = sc.mats + size_t(m) * 4;
        const float4 mkd = mp[0], mka = mp[1], mks = mp[2], maux = mp[3];
        const int bsdf0 = __float_as_int(mkd.w);
        if (b == 0 || previousBsdf == 1 /*SPECULAR*/ || previousBsdf == 4 /*REFRACTIVE*/)
            color = color + mask * v3(mka) * invLA;

        const float4* shp = sc.shade + size_t(h.rec) * 4;
        const float4 s0 = shp[0], s1 = shp[1], s2 = shp[2];
        const V3 normal = shading_normal(s0, s1, s2, h.u, h.v);
        const V3 position = ro + rd * h.t + normal * kEps2;
        const V3 outputDirection = normalize(rd);
        int bsdf = bsdf0;
        uint4 ra = make_uint4(0, 0, 0, 0);
        const bool needA = bsdf == 0 || bsdf == 2 || bsdf == 4;
        if (needA) ra = rng.block(1u + 2u * b);

        if (bsdf == 2 /*DIFFSPEC*/) {
            float p = maux.y;
            if (u01(ra.x) < p) {
                bsdf = 0;
                mask = mask * (1.0f / p);
            } else {
                bsdf = 1;
                mask = mask * (1.0f / (1.0f - p));
            }
        }
        if (bsdf == 4 /*REFRACTIVE*/) {
            // getFresnelReflectance (renderer.cu:464-493) with its quirks (Q5).
            V3 tdir = v3(0.f, 0.f, 0.f);
            V3 n = normal;
            float cosi = dot(outputDirection, normal);
            float etai = 1.f, etat = maux.x;
            if (cosi < 0.0f) {
                cosi = -cosi;
            } else {
                float tmp = etai;
                etai = etat;
                etat = tmp;
                n = normal * -1.0f;
            }
            float eta = etai / etat;
            float k = 1.f - eta * eta * (1.f - cosi * cosi);
            float Re, Tr;
            if (k < 0.f) {
                Re = 1.0f;
                Tr = 0.0f;
            } else {
                tdir = outputDirection * eta + n * (eta * cosi - sqrtf(k));
                float R0 = (etai - etat) * (etai - etat) / (etat + etai) * (etat + etai);
                float c = 1.f - dot(tdir, normal);
                Re = R0 + (1.f - R0) * c * c * c * c * c;
                Tr = 1.f - Re;
            }
            if (Re == 1.0f) {
                bsdf = 1;
            } else {
                float P = 0.25f + .5f * Re;
                float RP = Re / P;
                float TP = Tr / (1.f - P);
                if (u01(ra.x) > 0.25f) {
                    inputDirection = tdir;
                    mask = mask * v3(mks) * TP;
                } else {
                    mask = mask * RP;
                    bsdf = 1;
                }
            }
        }
        uint4 rb = make_uint4(0, 0, 0, 0);
        const bool needB = bsdf == 0 || b >= pa.rr_start;
        if (needB) rb = rng.block(2u + 2u * b);
        if (bsdf == 0 /*DIFFUSE*/) {
            // sampleDiffuseBSDF (renderer.cu:375-402): pdf 0.5, weight kd * cos (Q1).
            float r1 = float(2 * M_PI * double(u01(ra.y)));
            float r2 = u01(ra.z);
            float r2sq = sqrtf(r2);
            V3 w = normal;
            V3 a = fabsf(w.x) > 0.1f ? v3(0.f, 1.f, 0.f) : v3(1.f, 0.f, 0.f);
            V3 uu = normalize(cross(a, w));
            V3 vv = cross(w, uu);
            float sr, cr;
            own_sincosf(r1, &sr, &cr);
            inputDirection = normalize(uu * cr * r2sq + vv * sr * r2sq + w * sqrtf(1.f - r2));
            float cw = dot(inputDirection, normal);
            V3 kd = v3(mkd);
            const int tex = __float_as_int(mka.w);
            if (tex >= 0) {
                const float4 s3 = shp[3];
                float bu = h.u, bv = h.v, bw = 1.f - bu - bv;
                // uv1 = (s2.y, s2.z), uv2 = (s2.w, s3.x), uv3 = (s3.y, s3.z)
                float tu = s2.y * bw + s2.w * bu + s3.y * bv;
                float tv = s2.z * bw + s3.x * bu + s3.z * bv;
                kd = tex_sample(sc, tex, tu, tv);
            }
            mask = mask * (kd * cw) / 0.5f;

            // Light choice (renderer.cu:218-221) and estimateDirectLighting (:414-445):
            // visible iff the closest hit of the shadow ray is the chosen light, tested as
            // "the light is hit at t_L and nothing is hit before it (ties to lower id)".
            float pick = u01(ra.w) * (float(sc.num_lights) - .00001f);
            const uint32_t li = sc.lights[uint32_t(truncf(pick))];
            V3 direct = v3(0.f, 0.f, 0.f);
            if (h.rec != li) {
                const float4* lp = sc.tris + size_t(li) * 3;
                const float4 la = lp[0], lb = lp[1], lc = lp[2];
                float lu = u01(rb.x), lv = u01(rb.y);
                if (lu + lv >= 1.0f) {
                    lu = 1.0f - lu;
                    lv = 1.0f - lv;
                }
                V3 pt = v3(la) + v3(lb) * lu + v3(lc) * lv;
                V3 so = position + normal * kEps;
                V3 sd = normalize(pt - position);
                if (COUNT) cnt.shadow++;
                float tL, uL, vL;
                if (COUNT) cnt.tris++;
                if (mt_intersect(la, lb, lc, so, sd, tL, uL, vL)) {
                    HitRec occ;
                    occ.t = tL;
                    occ.id = __float_as_uint(la.w);
                    occ.rec = li;
                    occ.u = uL;
                    occ.v = vL;
                    if (!traverse<true, COUNT>(sc, so, sd, kEps, occ, st, cnt)) {
                        const float4* lsp = sc.shade + size_t(li) * 4;
                        const float4 l0 = lsp[0], l1 = lsp[1], l2 = lsp[2], l3 = lsp[3];
                        if (COUNT) cnt.hits++;
                        float sa = l3.w;
                        float d2 = tL * tL;
                        float cosT = fabsf(dot(shading_normal(l0, l1, l2, uL, vL), sd * -1.0f));
                        float wf = sa / (d2 * sc.lights_area) * cosT;
                        const float4 lka = sc.mats[size_t(__float_as_uint(lb.w)) * 4 + 1];
                        direct = direct + v3(lka) * wf;
                    }
                }
            }
            color = color + mask * direct;
        }
        if (bsdf == 1 /*SPECULAR*/) {
            // sampleSpecularBSDF / reflect (renderer.cu:404-412)
            inputDirection = outputDirection - normal * dot(outputDirection, normal) * 2.f;
            mask = mask * v3(mks) / 1.0f;
        }
        previousBsdf = bsdf;
        ro = position + inputDirection * kEps;
        rd = inputDirection;
        if (b >= pa.rr_start) {
            float p = max_component(mask);
            if (u01(rb.z) > p) break;
            mask = mask / p;
        }
    }
    return color;
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
__global__ __launch_bounds__(kBlock) void megakernel(MegaArgs a) {
    __shared__ int s_stack[kStackLds * kBlock];
    const uint32_t tile = xcd_tile(blockIdx.x, a.tiles);
    if (tile >= a.tiles) return;
    const int lane = int(threadIdx.x);
    const int x = int(tile % a.tiles_x) * 8 + (lane & 7);
    const int y = int(tile / a.tiles_x) * 8 + (lane >> 3);
    Counts cnt = {0, 0, 0, 0, 0, 0};
    if (x < int(a.width) && y < int(a.height)) {
        Stack st;
        st.lds = s_stack + lane;
        st.spill_stride = a.spill_stride;
        st.spill = a.spill + size_t(blockIdx.x) * kBlock + lane;
        st.sp = 0;
        const size_t idx = size_t(y) * a.width + size_t(x);
        float4 acc = a.accum[idx];
        V3 sum = v3(acc.x, acc.y, acc.z);
        for (uint32_t k = 0; k < a.nspp; ++k) {
            V3 L = sample_path<COUNT>(a.sc, a.cam, a.path, x, y, a.s0 + k, st, cnt);
            sum = sum + L;
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
    st.lds = s_stack + lane;
    st.spill_stride = a.spill_stride;
    st.spill = a.spill + size_t(blockIdx.x) * kBlock + lane;
    st.sp = 0;
    Rng rng{a.path.seed, a.path.view, uint32_t(y) * uint32_t(a.cam.xpixels) + uint32_t(x), a.sample};
    V3 ro, rd;
    camera_ray(a.cam, x, y, rng.block(0), ro, rd);
    HitRec h;
    h.t = kFltMax;
    h.id = 0;
    h.rec = 0;
    h.u = h.v = 0.f;
    Counts cnt;
    bool hit = traverse<false, false>(a.sc, ro, rd, kEps, h, st, cnt);
    const size_t idx = size_t(y) * a.width + size_t(x);
    if (a.tri) a.tri[idx] = hit ? int32_t(h.id) : -1;
    if (a.mat) a.mat[idx] = hit ? int32_t(__float_as_uint(a.sc.tris[size_t(h.rec) * 3 + 1].w)) : -1;
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
