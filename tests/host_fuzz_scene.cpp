// Host-only driver for the scene API (prt_scene_create and the calls on a built
// scene), built by tests/test_host_sanitizers.py with AddressSanitizer and UBSan
// (g++, no GPU code): random valid scenes of every BSDF and with textures, and
// descriptions that must be refused (material, texture or BSDF out of range,
// empty textures, too many materials, no triangles, non-finite positions,
// normals or texture coordinates), and
// degenerate triangles and coordinates near the float range, which must build.
#include "scene.hpp"

#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

// A host-only scene never gets device copies, so there is nothing to release.
void prt::release_device_scene(DeviceScene&) {}

namespace {

struct Owned {
    std::vector<float> pos, nrm, uv;
    std::vector<uint32_t> mid;
    std::vector<prt_material> mats;
    std::vector<std::vector<uint8_t>> texels;
    std::vector<prt_texture> tex;
    prt_scene_desc desc() const {
        prt_scene_desc d{};
        d.num_triangles = uint32_t(mid.size());
        d.positions = pos.data();
        d.normals = nrm.data();
        d.uvs = uv.empty() ? nullptr : uv.data();
        d.material_ids = mid.data();
        d.num_materials = uint32_t(mats.size());
        d.materials = mats.data();
        d.num_textures = uint32_t(tex.size());
        d.textures = tex.data();
        return d;
    }
};

Owned random_scene(std::mt19937& g, uint32_t n, uint32_t nm, uint32_t nt, int geometry) {
    std::normal_distribution<float> N(0.f, 1.f);
    Owned o;
    o.pos.resize(size_t(n) * 9);
    for (auto& x : o.pos) x = N(g);
    if (geometry == 1 && n > 2) {  // degenerate triangles, coordinates whose extents overflow
        for (int k = 0; k < 9; ++k) o.pos[18 + k] = 1.f;
        o.pos[0] = 3e38f;
        o.pos[10] = -3e38f;
    }
    o.nrm.assign(size_t(n) * 9, 0.f);
    for (uint32_t i = 0; i < n * 3; ++i) o.nrm[size_t(i) * 3 + 2] = 1.f;
    if (g() & 1) {
        o.uv.resize(size_t(n) * 6);
        for (auto& x : o.uv) x = N(g);
    }
    for (uint32_t i = 0; i < n; ++i) o.mid.push_back(g() % nm);
    for (uint32_t t = 0; t < nt; ++t) {
        const uint32_t w = 1 + g() % 33, h = 1 + g() % 17;
        o.texels.emplace_back(size_t(w) * h * 4, uint8_t(t * 37));
        o.tex.push_back(prt_texture{w, h, o.texels.back().data()});
    }
    for (uint32_t m = 0; m < nm; ++m) {
        prt_material x{};
        x.bsdf = int32_t(g() % 6);
        for (int k = 0; k < 3; ++k) {
            x.kd[k] = float(g() % 100) / 100.f;
            x.ka[k] = x.bsdf == PRT_EMISSIVE ? 3.f : 0.f;
            x.ks[k] = 0.5f;
        }
        x.ni = 1.5f;
        x.diffuse_coef = 0.5f;
        x.tex_kd = (nt && (g() & 1)) ? int32_t(g() % nt) : PRT_NO_TEXTURE;
        o.mats.push_back(x);
    }
    return o;
}

int exercise(const Owned& o) {
    const prt_scene_desc d = o.desc();
    prt_scene* s = nullptr;
    const int rc = prt_scene_create(&d, &s);
    if (rc != PRT_OK) return rc;
    prt_bvh_info info{};
    prt_scene_bvh_info(s, &info);
    std::vector<float> nodes(size_t(info.num_nodes) * 16), tris(size_t(d.num_triangles) * 12);
    std::vector<uint32_t> order(d.num_triangles);
    prt_scene_export_bvh(s, nodes.data(), tris.data(), order.data());
    uint32_t nl = 0;
    float la = 0.f;
    prt_scene_num_lights(s, &nl, &la);
    const float kd[3] = {0.2f, 0.3f, 0.4f};
    prt_scene_set_material_params(s, 0, kd, 2.f);
    const int bad = prt_scene_set_material_params(s, d.num_materials, kd, 2.f);
    prt_scene_destroy(s);
    return bad == PRT_OK ? 100 : PRT_OK;
}

}  // namespace

int main() {
    std::mt19937 g(9);
    int built = 0, refused = 0, wrong = 0;
    for (int k = 0; k < 60; ++k) {
        const uint32_t n = 1 + g() % 3000, nm = 1 + g() % 128, nt = g() % 4;
        Owned o = random_scene(g, n, nm, nt, k % 2);
        const int rc = exercise(o);
        if (rc == PRT_OK) ++built; else ++wrong;
        // The same scene made invalid in one way at a time: each must be refused.
        Owned a = o; a.mid[g() % n] = nm;                        // material id out of range
        Owned b = o; b.mats[0].tex_kd = int32_t(nt);             // texture index out of range
        Owned c = o; c.mats[g() % nm].bsdf = 6 + int32_t(g() % 5);  // unknown bsdf
        Owned e = o; e.mats.resize(129, o.mats[0]);              // too many materials
        Owned f = o; f.mid.clear();                              // no triangles
        Owned h = o;
        if (nt) h.tex[0].width = 0; else h.mats[0].bsdf = -1;    // empty texture (or negative bsdf)
        Owned i = o; i.pos[g() % i.pos.size()] = (k & 1) ? NAN : -INFINITY;  // non-finite vertex
        Owned j = o; j.nrm[g() % j.nrm.size()] = NAN;                      // non-finite normal
        Owned l = o;                                                       // non-finite uv
        if (l.uv.empty()) l.uv.assign(size_t(n) * 6, 0.5f);
        l.uv[g() % l.uv.size()] = INFINITY;
        for (const Owned* x : {&a, &b, &c, &e, &f, &h, &i, &j, &l}) {
            if (exercise(*x) != PRT_OK) ++refused; else ++wrong;
        }
    }
    std::printf("built %d, refused %d, wrong %d\n", built, refused, wrong);
    return wrong == 0 ? 0 : 1;
}
