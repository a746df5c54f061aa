// prt_scene: light list, BVH and material packing on the host.
//
// Scene::constructLightList (src/geometry/scene.cpp:148-173) collects emissive
// triangles and multiplies every material's ka by the total light area; this
// does the same, with the light list and the area sum in original triangle order
// so the result does not depend on the BVH. Materials are packed like
// ParallelRenderer::createMaterialsData (src/renderer/parallel_renderer.cu:252-279)
// but as four float4 per material so one 64-byte load fetches a material.
#include <omp.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <memory>

#include "common.hpp"
#include "scene.hpp"

namespace prt {

void pack_material(const prt_material& m, float lights_area, float out[16]) {
    std::memset(out, 0, 16 * sizeof(float));
    for (int k = 0; k < 3; ++k) {
        out[k] = m.kd[k];
        float ka = m.ka[k] * kLightsGain;  // materialFromMtl, scene.cpp:177
        out[4 + k] = ka * lights_area;     // constructLightList, scene.cpp:161-164
        out[8 + k] = m.ks[k];
    }
    std::memcpy(&out[3], &m.bsdf, 4);
    std::memcpy(&out[7], &m.tex_kd, 4);
    out[11] = m.ns;
    out[12] = m.ni;
    out[13] = m.diffuse_coef;
}

}  // namespace prt

using namespace prt;

extern "C" int prt_scene_create(const prt_scene_desc* d, prt_scene** out) {
    clear_error();
    if (!d || !out) return fail(PRT_ERR_INVALID_ARG, "null argument");
    if (d->num_triangles == 0 || !d->positions || !d->normals || !d->material_ids)
        return fail(PRT_ERR_INVALID_ARG, "scene has no triangles");
    if (d->num_materials == 0 || !d->materials) return fail(PRT_ERR_INVALID_ARG, "scene has no materials");
    if (d->num_materials > PRT_MAX_MATERIALS) return fail(PRT_ERR_LIMIT, "more than 128 materials");
    if (d->num_triangles >= kMaxRecords) return fail(PRT_ERR_LIMIT, "more than 2^28 triangles");
    PRT_TRY_BEGIN
    std::unique_ptr<prt_scene> s(new prt_scene);
    const uint32_t n = d->num_triangles;
    s->num_triangles = n;
    s->num_materials = d->num_materials;
    s->materials.assign(d->materials, d->materials + d->num_materials);
    s->camera = d->camera;
    for (uint32_t i = 0; i < n; ++i)
        if (d->material_ids[i] >= d->num_materials) return fail(PRT_ERR_INVALID_ARG, "material id out of range");
    for (size_t k = 0; k < size_t(n) * 9; ++k) {
        if (!std::isfinite(d->positions[k])) return fail(PRT_ERR_INVALID_ARG, "vertex position not finite");
        if (!std::isfinite(d->normals[k])) return fail(PRT_ERR_INVALID_ARG, "vertex normal not finite");
    }
    if (d->uvs)
        for (size_t k = 0; k < size_t(n) * 6; ++k)
            if (!std::isfinite(d->uvs[k])) return fail(PRT_ERR_INVALID_ARG, "texture coordinate not finite");
    for (uint32_t m = 0; m < d->num_materials; ++m) {
        int t = d->materials[m].tex_kd;
        if (t != PRT_NO_TEXTURE && (t < 0 || uint32_t(t) >= d->num_textures))
            return fail(PRT_ERR_INVALID_ARG, "texture index out of range");
        if (d->materials[m].bsdf < PRT_DIFFUSE || d->materials[m].bsdf > PRT_EMISSIVE)
            return fail(PRT_ERR_INVALID_ARG, "unknown bsdf");
    }
    // Surface areas, lights and total light area in original order.
    std::vector<float> area(n);
    for (uint32_t i = 0; i < n; ++i) {
        const float* p = d->positions + size_t(i) * 9;
        F3 e1 = f3(p + 3) - f3(p), e2 = f3(p + 6) - f3(p);
        area[i] = length(cross(e1, e2)) / 2.0f;
    }
    std::vector<uint32_t> light_ids;
    float la = 0.f;
    for (uint32_t i = 0; i < n; ++i)
        if (d->materials[d->material_ids[i]].bsdf == PRT_EMISSIVE) {
            light_ids.push_back(i);
            la += area[i];
        }
    s->lights_area = la;

    BvhParams bp;
    build_bvh(d->positions, n, bp, s->bvh);
    // A visit pushes at most three siblings; the traversal stack holds 128 entries.
    if (3 * (s->bvh.max_depth + 1) > 128) return fail(PRT_ERR_LIMIT, "BVH deeper than the traversal stack");
    const std::vector<uint32_t>& order = s->bvh.order;
    std::vector<uint32_t> where(n);
    for (uint32_t r = 0; r < n; ++r) where[order[r]] = r;
    for (uint32_t id : light_ids) s->lights.push_back(where[id]);

    s->tris.resize(size_t(n) * 12);
    s->shade.resize(size_t(n) * 16);
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < int64_t(n); ++r) {
        const uint32_t id = order[size_t(r)];
        const float* p = d->positions + size_t(id) * 9;
        const float* q = d->normals + size_t(id) * 9;
        float* t = &s->tris[size_t(r) * 12];
        for (int k = 0; k < 3; ++k) {
            t[k] = p[k];
            t[4 + k] = p[3 + k] - p[k];
            t[8 + k] = p[6 + k] - p[k];
        }
        uint32_t mat = d->material_ids[id];
        std::memcpy(&t[3], &id, 4);
        std::memcpy(&t[7], &mat, 4);
        t[11] = 0.f;
        float* sh = &s->shade[size_t(r) * 16];
        for (int k = 0; k < 9; ++k) sh[k] = q[k];
        for (int k = 0; k < 6; ++k) sh[9 + k] = d->uvs ? d->uvs[size_t(id) * 6 + k] : 0.f;
        sh[15] = area[id];
    }
    s->mats.resize(size_t(d->num_materials) * 16);
    for (uint32_t m = 0; m < d->num_materials; ++m) pack_material(d->materials[m], la, &s->mats[size_t(m) * 16]);

    // Textures: RGBA8 kept as bytes; texel -> linear through a 256-entry table
    // (stbi__ldr_to_hdr, src/include/stb_image.h:1682, pow(b/255, 2.2)).
    for (int i = 0; i < 256; ++i) s->lin[i] = std::pow(float(i) / 255.0f, 2.2f) * 1.0f;
    uint64_t texel = 0;
    for (uint32_t t = 0; t < d->num_textures; ++t) {
        const prt_texture& tx = d->textures[t];
        if (tx.width == 0 || tx.height == 0 || !tx.rgba) return fail(PRT_ERR_INVALID_ARG, "empty texture");
        s->texinfo.insert(s->texinfo.end(), {uint32_t(texel), tx.width, tx.height, 0u});
        s->texels.insert(s->texels.end(), tx.rgba, tx.rgba + size_t(tx.width) * tx.height * 4);
        texel += uint64_t(tx.width) * tx.height;
        if (texel >= (1ull << 32)) return fail(PRT_ERR_LIMIT, "textures larger than 2^32 texels");
    }
    *out = s.release();
    return PRT_OK;
    PRT_TRY_END
}

extern "C" int prt_scene_bvh_info(const prt_scene* s, prt_bvh_info* info) {
    clear_error();
    if (!s || !info) return fail(PRT_ERR_INVALID_ARG, "null argument");
    info->num_nodes = s->bvh.num_nodes;
    info->num_leaves = s->bvh.num_leaves;
    info->max_depth = s->bvh.max_depth;
    info->max_leaf_size = s->bvh.max_leaf;
    info->node_bytes = uint64_t(s->bvh.num_nodes) * kNodeFloats * sizeof(float);
    info->node_read_bytes = kNodeReadBytes;
    info->triangle_bytes = uint64_t(s->num_triangles) * 48;
    info->shading_bytes = uint64_t(s->num_triangles) * 64;
    info->sah_cost = s->bvh.sah_cost;
    return PRT_OK;
}

extern "C" int prt_scene_export_bvh(const prt_scene* s, float* nodes, float* tris, uint32_t* tri_order) {
    clear_error();
    if (!s) return fail(PRT_ERR_INVALID_ARG, "null argument");
    if (nodes) std::memcpy(nodes, s->bvh.nodes.data(), s->bvh.nodes.size() * sizeof(float));
    if (tris) std::memcpy(tris, s->tris.data(), s->tris.size() * sizeof(float));
    if (tri_order) std::memcpy(tri_order, s->bvh.order.data(), s->bvh.order.size() * sizeof(uint32_t));
    return PRT_OK;
}

extern "C" int prt_scene_num_lights(const prt_scene* s, uint32_t* num_lights, float* lights_area) {
    clear_error();
    if (!s) return fail(PRT_ERR_INVALID_ARG, "null argument");
    if (num_lights) *num_lights = uint32_t(s->lights.size());
    if (lights_area) *lights_area = s->lights_area;
    return PRT_OK;
}

extern "C" int prt_scene_set_material_params(prt_scene* s, uint32_t index, const float kd[3], float emission) {
    clear_error();
    if (!s || !kd) return fail(PRT_ERR_INVALID_ARG, "null argument");
    if (index >= s->num_materials) return fail(PRT_ERR_INVALID_ARG, "material index out of range");
    std::lock_guard<std::mutex> g(s->dev_mutex);
    prt_material& m = s->materials[index];
    for (int k = 0; k < 3; ++k) {
        m.kd[k] = kd[k];
        m.ka[k] = emission;
    }
    pack_material(m, s->lights_area, &s->mats[size_t(index) * 16]);
    s->mats_version++;
    return PRT_OK;
}

extern "C" int prt_scene_destroy(prt_scene* s) {
    if (!s) return PRT_OK;
    for (auto& d : s->devices) release_device_scene(d);
    delete s;
    return PRT_OK;
}
