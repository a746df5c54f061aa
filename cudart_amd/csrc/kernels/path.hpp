// One path vertex of samplePixel (src/renderer/renderer.cu:119-262), shared by
// the megakernel and the wavefront kernels so both integrators perform the same
// operations in the same order.
//
// Per vertex the caller
//   1. traces the ray (closest hit); on a miss calls path_miss and stops;
//   2. calls shade_vertex, which adds emission, samples the BSDF, updates the
//      throughput, picks the light sample and decides Russian roulette;
//   3. when shade_vertex asks for next-event estimation, calls nee_prepare (the
//      light-sample geometry of estimateDirectLighting, renderer.cu:414-445) and,
//      if the light sample is reachable, tests the shadow ray for occlusion; then
//      adds mask * direct with direct = 0 when occluded (renderer.cu:235).
#pragma once

#include "device_common.hpp"

namespace prt {
namespace dev {

struct PathState {
    V3 ro, rd;
    V3 mask, color;
    V3 input;      // SurfaceInteraction::inputDirection, stale across EMISSIVE vertices (Q4)
    int prev;      // previousBsdf
    uint32_t b;    // bounce index
};

// Light sample chosen at a diffuse vertex.
struct NeeSample {
    V3 position, normal;
    V3 mask;         // throughput after this vertex's diffuse weight
    uint32_t li;     // light record
    float lu, lv;    // point on the light (already folded when lu + lv >= 1)
};

struct ShadowReq {
    V3 so, sd;       // shadow ray
    float tL;        // distance to the light sample (the occlusion bound)
    uint32_t lid;    // original id of the light triangle (ties)
    V3 direct;       // 0 + lightColor * wf, the value when the light is visible
    float wf;        // geometric weight A_l cos / (d^2 A_L)
    uint32_t lmat;   // material of the light triangle
};

__device__ __forceinline__ void path_init(PathState& ps, V3 ro, V3 rd) {
    ps.ro = ro;
    ps.rd = rd;
    ps.mask = v3(1.f, 1.f, 1.f);
    ps.color = v3(0.f, 0.f, 0.f);
    ps.input = v3(0.f, 0.f, 0.f);
    ps.prev = 0;
    ps.b = 0;
}

__device__ __forceinline__ void path_miss(PathState& ps) {
    ps.color = ps.color + ps.mask * v3(0.0f, 0.025f, 0.05f);  // skybox, renderer.cu:64-70,147-149
}

// color + mask * direct (visible) or color + mask * 0 (occluded or unreachable).
__device__ __forceinline__ V3 add_direct(V3 color, V3 mask, V3 direct, bool visible) {
    return color + mask * (visible ? direct : v3(0.f, 0.f, 0.f));
}

// Russian-roulette step of one vertex, reported to the adjoint: p = max(mask)
// before the division, and the channels that attain it (bit c set).
struct RrInfo {
    float p;       // < 0: no roulette at this vertex
    uint32_t ties;
};

// Shades the hit `h` of the current ray. Returns true when the path continues
// (next ray in ps.ro / ps.rd, bounce incremented); sets *want_nee and fills `ns`
// at a diffuse vertex whose chosen light is another triangle.
// EAGER evaluates both Philox blocks of the bounce up front, independent of any
// load (latency hiding in the shade kernel); otherwise only the blocks the
// material needs are evaluated (fewer registers in the megakernel). The values
// are the same either way: a block depends only on (pixel, sample, bounce).
template <bool COUNT, bool EAGER, class RngType>
__device__ __forceinline__ bool shade_vertex(const SceneArgs& sc, const PathArgs& pa, const RngType& rng, PathState& ps,
                                             const HitRec& h, bool* want_nee, NeeSample& ns, Counts& cnt,
                                             RrInfo* rr = nullptr) {
    *want_nee = false;
    if (rr) rr->p = -1.f;
    const uint32_t b = ps.b;
    uint4 ra = make_uint4(0, 0, 0, 0), rb = make_uint4(0, 0, 0, 0);
    if (EAGER) {
        ra = rng.block(1u + 2u * b);
        rb = rng.block(2u + 2u * b);
    }
    if (COUNT) cnt.hits++;
    const float4* mp = sc.mats + size_t(h.mat) * 4;
    const float4* shp = sc.shade + size_t(h.rec) * 4;
    const float4 mkd = mp[0], mka = mp[1], mks = mp[2], maux = mp[3];
    const float4 s0 = shp[0], s1 = shp[1], s2 = shp[2];
    if (!EAGER) {
        const int k = __float_as_int(mkd.w);
        if (k == 0 || k == 2 || k == 4) ra = rng.block(1u + 2u * b);
    }
    if (b == 0 || ps.prev == 1 /*SPECULAR*/ || ps.prev == 4 /*REFRACTIVE*/)
        ps.color = ps.color + ps.mask * v3(mka) * sc.inv_lights_area;

    const V3 normal = shading_normal(s0, s1, s2, h.u, h.v);
    const V3 position = ps.ro + ps.rd * h.t + normal * kEps2;
    const V3 outputDirection = normalize(ps.rd);
    int bsdf = __float_as_int(mkd.w);

    if (bsdf == 2 /*DIFFSPEC*/) {
        float p = maux.y;
        if (u01(ra.x) < p) {
            bsdf = 0;
            ps.mask = ps.mask * (1.0f / p);
        } else {
            bsdf = 1;
            ps.mask = ps.mask * (1.0f / (1.0f - p));
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
                ps.input = tdir;
                ps.mask = ps.mask * v3(mks) * TP;
            } else {
                ps.mask = ps.mask * RP;
                bsdf = 1;
            }
        }
    }
    if (!EAGER && (bsdf == 0 || b >= pa.rr_start)) rb = rng.block(2u + 2u * b);
    if (bsdf == 0 /*DIFFUSE*/) {
        // sampleDiffuseBSDF (renderer.cu:375-402): pdf 0.5, weight kd * cos (Q1).
        float r1 = float(2 * M_PI * double(u01(ra.y)));
        float r2 = u01(ra.z);
        float r2sq = sqrtf(r2);
        V3 w = normal;
        V3 a = fabsf(w.x) > 0.1f ? v3(0.f, 1.f, 0.f) : v3(1.f, 0.f, 0.f);
        V3 uu = normalize(cross(a, w));
        V3 vv = cross(w, uu);
        float sn, cs;
        own_sincosf(r1, &sn, &cs);
        ps.input = normalize(uu * cs * r2sq + vv * sn * r2sq + w * sqrtf(1.f - r2));
        float cw = dot(ps.input, normal);
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
        ps.mask = ps.mask * (kd * cw) / 0.5f;

        // Light choice (renderer.cu:218-221); estimateDirectLighting returns 0 when the
        // shading triangle is the chosen light (:421-423).
        float pick = u01(ra.w) * (float(sc.num_lights) - .00001f);
        const uint32_t li = sc.lights[uint32_t(truncf(pick))];
        if (h.rec != li) {
            float lu = u01(rb.x), lv = u01(rb.y);
            if (lu + lv >= 1.0f) {
                lu = 1.0f - lu;
                lv = 1.0f - lv;
            }
            ns.position = position;
            ns.normal = normal;
            ns.mask = ps.mask;
            ns.li = li;
            ns.lu = lu;
            ns.lv = lv;
            *want_nee = true;
        } else {
            ps.color = ps.color + ps.mask * v3(0.f, 0.f, 0.f);
        }
    }
    if (bsdf == 1 /*SPECULAR*/) {
        // sampleSpecularBSDF / reflect (renderer.cu:404-412)
        ps.input = outputDirection - normal * dot(outputDirection, normal) * 2.f;
        ps.mask = ps.mask * v3(mks) / 1.0f;
    }
    ps.prev = bsdf;
    ps.ro = position + ps.input * kEps;
    ps.rd = ps.input;
    if (b >= pa.rr_start) {
        float p = max_component(ps.mask);
        if (rr) {
            rr->p = p;
            rr->ties = (ps.mask.x == p ? 1u : 0u) | (ps.mask.y == p ? 2u : 0u) | (ps.mask.z == p ? 4u : 0u);
        }
        if (u01(rb.z) > p) return false;
        ps.mask = ps.mask / p;
    }
    ps.b = b + 1;
    return ps.b < pa.max_depth;
}

// Shadow ray toward the light sample (estimateDirectLighting, renderer.cu:414-445).
// Returns false when the ray misses the light triangle itself: its closest hit
// cannot be the light, so the sample contributes mask * 0. Otherwise the light is
// visible iff nothing has (t, id) below (tL, light id) along the ray.
template <bool COUNT>
__device__ __forceinline__ bool nee_prepare(const SceneArgs& sc, const NeeSample& ns, ShadowReq& sr, Counts& cnt) {
    const float4* lp = sc.tris + size_t(ns.li) * 3;
    const float4 la = lp[0], lb = lp[1], lc = lp[2];
    V3 pt = v3(la) + v3(lb) * ns.lu + v3(lc) * ns.lv;
    sr.so = ns.position + ns.normal * kEps;
    sr.sd = normalize(pt - ns.position);
    if (COUNT) {
        cnt.shadow++;
        cnt.tris++;
    }
    float tL, uL, vL;
    if (!mt_intersect(la, lb, lc, sr.so, sr.sd, tL, uL, vL)) return false;
    const float4* lsp = sc.shade + size_t(ns.li) * 4;
    const float4 l0 = lsp[0], l1 = lsp[1], l2 = lsp[2], l3 = lsp[3];
    if (COUNT) cnt.hits++;
    float sa = l3.w;
    float d2 = tL * tL;
    float cosT = fabsf(dot(shading_normal(l0, l1, l2, uL, vL), sr.sd * -1.0f));
    float wf = sa / (d2 * sc.lights_area) * cosT;
    const float4 lka = sc.mats[size_t(__float_as_uint(lb.w)) * 4 + 1];
    sr.tL = tL;
    sr.lid = __float_as_uint(la.w);
    sr.direct = v3(0.f, 0.f, 0.f) + v3(lka) * wf;
    sr.wf = wf;
    sr.lmat = __float_as_uint(lb.w);
    return true;
}

}  // namespace dev
}  // namespace prt
