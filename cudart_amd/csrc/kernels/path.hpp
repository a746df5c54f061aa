// One path vertex of samplePixel (src/renderer/renderer.cu:119-262), shared by
// the megakernel and the wavefront shade kernel so both integrators perform the
// same operations in the same order.
//
// The caller traces the ray (closest hit), adds the sky on a miss, calls
// shade_vertex on a hit, then — when it returns a shadow request — tests the
// shadow ray and adds mask * direct to the colour, with direct = 0 when the light
// is occluded, exactly as estimateDirectLighting's caller does (renderer.cu:235).
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

struct ShadowReq {
    V3 so, sd;       // shadow ray
    float tL;        // distance to the light sample (the occlusion bound)
    uint32_t lid;    // original id of the light triangle (ties)
    V3 mask;         // path throughput after this vertex's diffuse weight
    V3 direct;       // 0 + lightColor * wf, the value when the light is visible
};

// color + mask * direct (visible) or color + mask * 0 (occluded), renderer.cu:235.
__device__ __forceinline__ V3 add_direct(V3 color, const ShadowReq& sr, bool visible) {
    return color + sr.mask * (visible ? sr.direct : v3(0.f, 0.f, 0.f));
}

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

// Shades the hit `h` of the current ray. Returns true when the path continues
// (next ray in ps.ro / ps.rd, bounce incremented); sets *want_shadow and fills
// `sr` when a shadow ray must be tested.
template <bool COUNT>
__device__ __forceinline__ bool shade_vertex(const SceneArgs& sc, const PathArgs& pa, const Rng& rng, PathState& ps,
                                             const HitRec& h, bool* want_shadow, ShadowReq& sr, Counts& cnt) {
    *want_shadow = false;
    const uint32_t b = ps.b;
    if (COUNT) cnt.hits++;
    const float4* tp = sc.tris + size_t(h.rec) * 3;
    const uint32_t m = __float_as_uint(tp[1].w);
    const float4* mp = sc.mats + size_t(m) * 4;
    const float4 mkd = mp[0], mka = mp[1], mks = mp[2], maux = mp[3];
    if (b == 0 || ps.prev == 1 /*SPECULAR*/ || ps.prev == 4 /*REFRACTIVE*/)
        ps.color = ps.color + ps.mask * v3(mka) * sc.inv_lights_area;

    const float4* shp = sc.shade + size_t(h.rec) * 4;
    const float4 s0 = shp[0], s1 = shp[1], s2 = shp[2];
    const V3 normal = shading_normal(s0, s1, s2, h.u, h.v);
    const V3 position = ps.ro + ps.rd * h.t + normal * kEps2;
    const V3 outputDirection = normalize(ps.rd);
    int bsdf = __float_as_int(mkd.w);
    uint4 ra = make_uint4(0, 0, 0, 0);
    if (bsdf == 0 || bsdf == 2 || bsdf == 4) ra = rng.block(1u + 2u * b);

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
    uint4 rb = make_uint4(0, 0, 0, 0);
    if (bsdf == 0 || b >= pa.rr_start) rb = rng.block(2u + 2u * b);
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

        // Light choice (renderer.cu:218-221) and estimateDirectLighting (:414-445).
        // Visible iff the closest hit of the shadow ray is the chosen light: the light
        // is hit at tL and nothing has (t, id) lexicographically below (tL, light id).
        float pick = u01(ra.w) * (float(sc.num_lights) - .00001f);
        const uint32_t li = sc.lights[uint32_t(truncf(pick))];
        bool shadow = false;
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
            if (COUNT) {
                cnt.shadow++;
                cnt.tris++;
            }
            float tL, uL, vL;
            if (mt_intersect(la, lb, lc, so, sd, tL, uL, vL)) {
                const float4* lsp = sc.shade + size_t(li) * 4;
                const float4 l0 = lsp[0], l1 = lsp[1], l2 = lsp[2], l3 = lsp[3];
                if (COUNT) cnt.hits++;
                float sa = l3.w;
                float d2 = tL * tL;
                float cosT = fabsf(dot(shading_normal(l0, l1, l2, uL, vL), sd * -1.0f));
                float wf = sa / (d2 * sc.lights_area) * cosT;
                const float4 lka = sc.mats[size_t(__float_as_uint(lb.w)) * 4 + 1];
                sr.so = so;
                sr.sd = sd;
                sr.tL = tL;
                sr.lid = __float_as_uint(la.w);
                sr.mask = ps.mask;
                sr.direct = v3(0.f, 0.f, 0.f) + v3(lka) * wf;
                shadow = true;
            }
        }
        if (shadow) *want_shadow = true;  // the caller adds mask * direct after the visibility test
        else ps.color = ps.color + ps.mask * v3(0.f, 0.f, 0.f);  // light missed or is this triangle
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
        if (u01(rb.z) > p) return false;
        ps.mask = ps.mask / p;
    }
    ps.b = b + 1;
    return ps.b < pa.max_depth;
}

}  // namespace dev
}  // namespace prt
