// Procedural scenes from a seed (the synthetic workloads the path is measured on).
// Every random number comes from Philox4x32-10 keyed by (seed, 0x5CE7E5) with
// counter (stream, index, 0, 0), so a scene depends only on its kind, its
// parameters and the seed.
//
//   spheres          displaced icosphere instances over a ground quad, lit by
//                    emissive quads (the 1M-triangle scene at the defaults)
//   cornell-spheres  closed Cornell-like box with a ceiling light and spheres
//                    (the differentiable-render scene)
//   mixed            spheres scene sized to a target triangle count (batches)
//   terrain-spheres  fBm height field plus UV-sphere instances with RGBA8
//                    textures (the large textured scene)
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <sstream>
#include <stdexcept>

#include "common.hpp"
#include "mesh.hpp"

namespace prt {
namespace {

void philox_host(const uint32_t c_in[4], const uint32_t k_in[2], uint32_t out[4]) {
    uint32_t c0 = c_in[0], c1 = c_in[1], c2 = c_in[2], c3 = c_in[3];
    uint32_t k0 = k_in[0], k1 = k_in[1];
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint64_t p0 = uint64_t(0xD2511F53u) * c0, p1 = uint64_t(0xCD9E8D57u) * c2;
        uint32_t n0 = uint32_t(p1 >> 32) ^ c1 ^ k0, n2 = uint32_t(p0 >> 32) ^ c3 ^ k1;
        c1 = uint32_t(p1);
        c3 = uint32_t(p0);
        c0 = n0;
        c2 = n2;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

struct Rng {
    uint32_t key[2];
    explicit Rng(uint32_t seed) : key{seed, 0x5CE7E5u} {}
    float u(uint32_t stream, uint32_t index) const {
        uint32_t c[4] = {stream, index, 0u, 0u}, o[4];
        philox_host(c, key, o);
        return float(o[0] >> 8) * 0x1p-24f;
    }
    uint32_t bits(uint32_t stream, uint32_t index) const {
        uint32_t c[4] = {stream, index, 1u, 0u}, o[4];
        philox_host(c, key, o);
        return o[0];
    }
};

enum Stream : uint32_t {
    kCenters = 1,
    kDisplace = 2,
    kAlbedo = 3,
    kKind = 4,
    kTerrain = 5,
    kTexels = 6,
    kInstMat = 7,
    kRadius = 8,
};

struct Params {
    std::map<std::string, std::string> kv;
    explicit Params(const std::string& s) {
        std::stringstream ss(s);
        std::string item;
        while (std::getline(ss, item, ';')) {
            size_t eq = item.find('=');
            if (eq == std::string::npos) continue;
            auto trim = [](std::string x) {
                size_t a = x.find_first_not_of(" \t"), b = x.find_last_not_of(" \t");
                return a == std::string::npos ? std::string() : x.substr(a, b - a + 1);
            };
            kv[trim(item.substr(0, eq))] = trim(item.substr(eq + 1));
        }
    }
    // The value of key k, or def. Not a number, not finite or beyond 1e9 in magnitude
    // is an error (the callers convert to int).
    double num(const char* k, double def) const {
        auto it = kv.find(k);
        if (it == kv.end()) return def;
        const double v = std::stod(it->second);
        if (!std::isfinite(v) || std::fabs(v) > 1e9)
            throw std::invalid_argument(std::string("parameter ") + k + " out of range");
        return v;
    }
    // A count or size: num(), which must not be negative, truncated.
    uint32_t count(const char* k, double def) const {
        const double v = num(k, def);
        if (v < 0.0) throw std::invalid_argument(std::string("parameter ") + k + " is negative");
        return uint32_t(v);
    }
};

struct Sphere {
    std::vector<F3> verts;       // unit positions (= normals)
    std::vector<uint32_t> faces; // triples
    std::vector<float> uv;       // per vertex (u, v) for UV spheres; empty for icospheres
};

Sphere icosphere(int level) {
    Sphere s;
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    const double base[12][3] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                                {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    std::vector<std::array<double, 3>> v;
    for (auto& b : base) {
        double n = std::sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
        v.push_back({b[0] / n, b[1] / n, b[2] / n});
    }
    std::vector<uint32_t> f = {0, 11, 5, 0, 5, 1,  0, 1, 7,  0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4,
                               11, 10, 2, 10, 7, 6, 7, 1, 8, 3, 9, 4,  3, 4, 2,  3, 2, 6, 3, 6, 8,
                               3, 8, 9, 4, 9, 5,  2, 4, 11, 6, 2, 10, 8, 6, 7,  9, 8, 1};
    for (int l = 0; l < level; ++l) {
        std::map<std::pair<uint32_t, uint32_t>, uint32_t> mid;
        auto midpoint = [&](uint32_t a, uint32_t b) {
            auto key = std::make_pair(std::min(a, b), std::max(a, b));
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            std::array<double, 3> m = {(v[a][0] + v[b][0]) * 0.5, (v[a][1] + v[b][1]) * 0.5, (v[a][2] + v[b][2]) * 0.5};
            double n = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
            v.push_back({m[0] / n, m[1] / n, m[2] / n});
            uint32_t id = uint32_t(v.size() - 1);
            mid[key] = id;
            return id;
        };
        std::vector<uint32_t> nf;
        nf.reserve(f.size() * 4);
        for (size_t i = 0; i < f.size(); i += 3) {
            uint32_t a = f[i], b = f[i + 1], c = f[i + 2];
            uint32_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
            uint32_t q[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
            nf.insert(nf.end(), q, q + 12);
        }
        f.swap(nf);
    }
    for (auto& p : v) s.verts.push_back(F3{float(p[0]), float(p[1]), float(p[2])});
    s.faces = f;
    return s;
}

// Latitude-longitude sphere with `slices` x `stacks` cells: 2*slices*(stacks-1) triangles.
Sphere uvsphere(int slices, int stacks) {
    Sphere s;
    for (int j = 0; j <= stacks; ++j) {
        double th = M_PI * double(j) / stacks;
        for (int i = 0; i <= slices; ++i) {
            double ph = 2.0 * M_PI * double(i) / slices;
            s.verts.push_back(F3{float(std::sin(th) * std::cos(ph)), float(std::cos(th)), float(std::sin(th) * std::sin(ph))});
            s.uv.push_back(float(double(i) / slices));
            s.uv.push_back(float(1.0 - double(j) / stacks));
        }
    }
    auto id = [&](int i, int j) { return uint32_t(j * (slices + 1) + i); };
    for (int j = 0; j < stacks; ++j)
        for (int i = 0; i < slices; ++i) {
            uint32_t a = id(i, j), b = id(i + 1, j), c = id(i, j + 1), d = id(i + 1, j + 1);
            if (j != 0) s.faces.insert(s.faces.end(), {a, b, d});
            if (j != stacks - 1) s.faces.insert(s.faces.end(), {a, d, c});
        }
    return s;
}

prt_material mat_of(int bsdf, F3 albedo) {
    prt_material m = default_material();
    m.bsdf = bsdf;
    m.kd[0] = albedo.x; m.kd[1] = albedo.y; m.kd[2] = albedo.z;
    m.ks[0] = albedo.x; m.ks[1] = albedo.y; m.ks[2] = albedo.z;
    m.ni = 1.5f;
    m.diffuse_coef = 0.5f;
    return m;
}
prt_material emitter(float radiance) {
    prt_material m = default_material();
    m.bsdf = PRT_EMISSIVE;
    for (int k = 0; k < 3; ++k) {
        m.kd[k] = 1.f;
        m.ka[k] = radiance / kLightsGain;  // emitted radiance = LIGHTS_GAIN * Ka
    }
    m.ni = 1.f;
    return m;
}

void add_quad(prt_mesh& m, F3 a, F3 b, F3 c, F3 d, F3 n, uint32_t mat) {
    // a b c d counter-clockwise; triangles (a,b,c) and (a,c,d).
    float nn[9] = {n.x, n.y, n.z, n.x, n.y, n.z, n.x, n.y, n.z};
    float p1[9] = {a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z};
    float t1[6] = {0, 0, 1, 0, 1, 1};
    float p2[9] = {a.x, a.y, a.z, c.x, c.y, c.z, d.x, d.y, d.z};
    float t2[6] = {0, 0, 1, 1, 0, 1};
    m.add_triangle(p1, nn, t1, mat);
    m.add_triangle(p2, nn, t2, mat);
}

void add_sphere(prt_mesh& m, const Sphere& s, F3 c, float r, float disp, const Rng& rng, uint32_t inst,
                uint32_t mat) {
    std::vector<F3> p(s.verts.size());
    for (size_t j = 0; j < s.verts.size(); ++j) {
        float d = disp > 0.f ? (rng.u(kDisplace, inst * 4096u + uint32_t(j)) - 0.5f) * 2.f * disp : 0.f;
        p[j] = c + s.verts[j] * (r + d);
    }
    for (size_t i = 0; i < s.faces.size(); i += 3) {
        float pp[9], nn[9], tt[6] = {0, 0, 0, 0, 0, 0};
        for (int k = 0; k < 3; ++k) {
            uint32_t vi = s.faces[i + size_t(k)];
            pp[k * 3] = p[vi].x; pp[k * 3 + 1] = p[vi].y; pp[k * 3 + 2] = p[vi].z;
            nn[k * 3] = s.verts[vi].x; nn[k * 3 + 1] = s.verts[vi].y; nn[k * 3 + 2] = s.verts[vi].z;
            if (!s.uv.empty()) {
                tt[k * 2] = s.uv[vi * 2];
                tt[k * 2 + 1] = s.uv[vi * 2 + 1];
            }
        }
        m.add_triangle(pp, nn, tt, mat);
    }
}

// Displaced icosphere instances over a ground quad, lit from above.
int gen_spheres(const Params& P, uint32_t seed, prt_mesh& m) {
    Rng rng(seed);
    const int instances = int(P.num("instances", 200));
    const int level = int(P.num("level", 4));
    const float radius = float(P.num("radius", 1.0));
    const float disp = float(P.num("disp", 0.05));
    const int nmat = int(P.num("materials", 32));
    const float half = float(P.num("ground", 100.0));
    const float emission = float(P.num("emission", 15.0));
    const uint32_t w = P.count("width", 1920), h = P.count("height", 1080);
    const float spread = float(P.num("spread", 10.0));
    if (instances < 0 || level < 0 || level > 7 || nmat < 1 || nmat > PRT_MAX_MATERIALS - 2)
        return fail(PRT_ERR_INVALID_ARG, "spheres: parameter out of range");
    // Materials: 60% diffuse, 25% specular, the rest refractive (IOR 1.5), albedo U[0.1, 0.8]^3.
    const int ndiff = int(std::lround(nmat * 0.6)), nspec = int(std::lround(nmat * 0.25));
    for (int i = 0; i < nmat; ++i) {
        F3 a{0.1f + 0.7f * rng.u(kAlbedo, uint32_t(i) * 3), 0.1f + 0.7f * rng.u(kAlbedo, uint32_t(i) * 3 + 1),
             0.1f + 0.7f * rng.u(kAlbedo, uint32_t(i) * 3 + 2)};
        int bsdf = i < ndiff ? PRT_DIFFUSE : (i < ndiff + nspec ? PRT_SPECULAR : PRT_REFRACTIVE);
        m.materials.push_back(mat_of(bsdf, a));
        m.material_names.push_back("sphere" + std::to_string(i));
    }
    const uint32_t groundMat = uint32_t(m.materials.size());
    m.materials.push_back(mat_of(PRT_DIFFUSE, F3{0.5f, 0.5f, 0.5f}));
    m.material_names.push_back("ground");
    const uint32_t lightMat = uint32_t(m.materials.size());
    m.materials.push_back(emitter(emission));
    m.material_names.push_back("light");

    Sphere s = icosphere(level);
    for (int i = 0; i < instances; ++i) {
        uint32_t k = uint32_t(i) * 3;
        F3 c{(rng.u(kCenters, k) * 2.f - 1.f) * spread, 1.f + rng.u(kCenters, k + 1) * 2.f * spread,
             (rng.u(kCenters, k + 2) * 2.f - 1.f) * spread};
        add_sphere(m, s, c, radius, disp, rng, uint32_t(i), uint32_t(i % nmat));
    }
    add_quad(m, F3{-half, 0, -half}, F3{-half, 0, half}, F3{half, 0, half}, F3{half, 0, -half}, F3{0, 1, 0}, groundMat);
    const float ly = 1.f + 2.f * spread + 9.f;
    for (float lx : {-5.f, 5.f})
        add_quad(m, F3{lx - 2, ly, -2}, F3{lx + 2, ly, -2}, F3{lx + 2, ly, 2}, F3{lx - 2, ly, 2}, F3{0, -1, 0}, lightMat);
    const float eye[3] = {0.f, 12.f, 35.f}, tgt[3] = {0.f, 8.f, 0.f}, up[3] = {0.f, 1.f, 0.f};
    camera_look_at(eye, tgt, up, 45.f, w, h, m.camera);
    m.has_camera = true;
    return PRT_OK;
}

// Cornell-like closed box (10 units) with a ceiling light and displaced spheres.
int gen_cornell_spheres(const Params& P, uint32_t seed, prt_mesh& m) {
    Rng rng(seed);
    const int instances = int(P.num("instances", 19));
    const int level = int(P.num("level", 4));
    const int nmat = int(P.num("materials", 16));
    const float emission = float(P.num("emission", 15.0));
    const uint32_t w = P.count("width", 512), h = P.count("height", 512);
    if (nmat < 5 || nmat > PRT_MAX_MATERIALS || level < 0 || level > 6)
        return fail(PRT_ERR_INVALID_ARG, "cornell-spheres: parameter out of range");
    m.materials.push_back(mat_of(PRT_DIFFUSE, F3{0.6f, 0.6f, 0.6f}));
    m.material_names.push_back("white");
    m.materials.push_back(mat_of(PRT_DIFFUSE, F3{0.7f, 0.15f, 0.12f}));
    m.material_names.push_back("red");
    m.materials.push_back(mat_of(PRT_DIFFUSE, F3{0.15f, 0.6f, 0.15f}));
    m.material_names.push_back("green");
    m.materials.push_back(emitter(emission));
    m.material_names.push_back("light");
    for (int i = 4; i < nmat; ++i) {
        F3 a{0.1f + 0.7f * rng.u(kAlbedo, uint32_t(i) * 3), 0.1f + 0.7f * rng.u(kAlbedo, uint32_t(i) * 3 + 1),
             0.1f + 0.7f * rng.u(kAlbedo, uint32_t(i) * 3 + 2)};
        int bsdf = (i % 6 == 4) ? PRT_SPECULAR : (i % 6 == 5 ? PRT_DIFFSPEC : PRT_DIFFUSE);
        m.materials.push_back(mat_of(bsdf, a));
        m.material_names.push_back("sphere" + std::to_string(i));
    }
    const float S = 5.f;
    add_quad(m, F3{-S, 0, -S}, F3{-S, 0, S}, F3{S, 0, S}, F3{S, 0, -S}, F3{0, 1, 0}, 0);          // floor
    add_quad(m, F3{-S, 2 * S, -S}, F3{S, 2 * S, -S}, F3{S, 2 * S, S}, F3{-S, 2 * S, S}, F3{0, -1, 0}, 0);  // ceiling
    add_quad(m, F3{-S, 0, -S}, F3{S, 0, -S}, F3{S, 2 * S, -S}, F3{-S, 2 * S, -S}, F3{0, 0, 1}, 0);  // back
    add_quad(m, F3{-S, 0, -S}, F3{-S, 2 * S, -S}, F3{-S, 2 * S, S}, F3{-S, 0, S}, F3{1, 0, 0}, 1);  // left (red)
    add_quad(m, F3{S, 0, -S}, F3{S, 0, S}, F3{S, 2 * S, S}, F3{S, 2 * S, -S}, F3{-1, 0, 0}, 2);    // right (green)
    const float ly = 2 * S - 0.5f, L = 1.25f;  // half a unit below the ceiling: no near-coplanar light
    add_quad(m, F3{-L, ly, -L}, F3{L, ly, -L}, F3{L, ly, L}, F3{-L, ly, L}, F3{0, -1, 0}, 3);
    Sphere s = icosphere(level);
    for (int i = 0; i < instances; ++i) {
        uint32_t k = uint32_t(i) * 3;
        float r = 0.6f + 0.6f * rng.u(kRadius, uint32_t(i));
        F3 c{(rng.u(kCenters, k) * 2.f - 1.f) * (S - 1.3f), 1.3f + rng.u(kCenters, k + 1) * (2 * S - 3.2f),
             (rng.u(kCenters, k + 2) * 2.f - 1.f) * (S - 1.3f)};
        add_sphere(m, s, c, r, 0.05f * r, rng, uint32_t(i), uint32_t(4 + i % (nmat - 4)));
    }
    const float eye[3] = {0.f, S, 3.f * S + 3.f}, tgt[3] = {0.f, S, 0.f}, up[3] = {0.f, 1.f, 0.f};
    camera_look_at(eye, tgt, up, 40.f, w, h, m.camera);
    m.has_camera = true;
    return PRT_OK;
}

// Spheres scene sized to about `tris` triangles.
int gen_mixed(const Params& P, uint32_t seed, prt_mesh& m) {
    const double tris = P.num("tris", 100000);
    const int level = tris >= 50000 ? 4 : 3;
    const double per = 20.0 * std::pow(4.0, level);
    int inst = int(std::max(1.0, std::round((tris - 6) / per)));
    std::ostringstream ss;
    ss << "instances=" << inst << ";level=" << level << ";width=" << P.num("width", 512)
       << ";height=" << P.num("height", 512) << ";materials=" << P.num("materials", 32)
       << ";spread=" << P.num("spread", 10.0);
    return gen_spheres(Params(ss.str()), seed, m);
}

// Value noise on an integer lattice, smoothly interpolated; fBm of `oct` octaves.
float lattice(const Rng& rng, int x, int y) {
    uint32_t h = uint32_t(x) * 73856093u ^ uint32_t(y) * 19349663u;
    return rng.u(kTerrain, h);
}
float vnoise(const Rng& rng, float x, float y) {
    int xi = int(std::floor(x)), yi = int(std::floor(y));
    float fx = x - float(xi), fy = y - float(yi);
    float sx = fx * fx * (3.f - 2.f * fx), sy = fy * fy * (3.f - 2.f * fy);
    float a = lattice(rng, xi, yi), b = lattice(rng, xi + 1, yi), c = lattice(rng, xi, yi + 1), d = lattice(rng, xi + 1, yi + 1);
    return (a * (1 - sx) + b * sx) * (1 - sy) + (c * (1 - sx) + d * sx) * sy;
}
float fbm(const Rng& rng, float x, float y, int oct) {
    float v = 0, amp = 0.5f, f = 1.f;
    for (int o = 0; o < oct; ++o) {
        v += amp * (vnoise(rng, x * f, y * f) * 2.f - 1.f);
        amp *= 0.5f;
        f *= 2.f;
    }
    return v;
}

// fBm terrain plus textured UV spheres.
int gen_terrain_spheres(const Params& P, uint32_t seed, prt_mesh& m) {
    Rng rng(seed);
    const int grid = int(P.num("grid", 1415));  // (grid-1)^2 * 2 terrain triangles
    const int instances = int(P.num("instances", 2000));
    const int slices = int(P.num("slices", 50)), stacks = int(P.num("stacks", 31));
    const int nmat = int(P.num("materials", 64));
    const int ntex = int(P.num("textures", 16));
    const int texSize = int(P.num("texsize", 1024));
    const float extent = float(P.num("extent", 200.0));
    const float emission = float(P.num("emission", 15.0));
    const uint32_t w = P.count("width", 3840), h = P.count("height", 2160);
    if (grid < 2 || nmat < 4 || nmat > PRT_MAX_MATERIALS - 1 || ntex < 0 || texSize < 1 || texSize > 8192)
        return fail(PRT_ERR_INVALID_ARG, "terrain-spheres: parameter out of range");
    // Textures: smooth random colour fields, RGBA8.
    m.texture_data.resize(size_t(ntex));
    for (int t = 0; t < ntex; ++t) {
        auto& px = m.texture_data[size_t(t)];
        px.resize(size_t(texSize) * texSize * 4);
        Rng tr(seed ^ (0x9E3779B9u * uint32_t(t + 1)));
        for (int y = 0; y < texSize; ++y)
            for (int x = 0; x < texSize; ++x) {
                float fx = float(x) * 16.f / texSize, fy = float(y) * 16.f / texSize;
                for (int k = 0; k < 3; ++k) {
                    float v = 0.5f + 0.5f * fbm(tr, fx + 37.f * k, fy, 4);
                    v = std::min(1.f, std::max(0.f, v));
                    px[(size_t(y) * texSize + x) * 4 + k] = uint8_t(v * 255.f);
                }
                px[(size_t(y) * texSize + x) * 4 + 3] = 255;
            }
    }
    for (int t = 0; t < ntex; ++t)
        m.textures.push_back(prt_texture{uint32_t(texSize), uint32_t(texSize), m.texture_data[size_t(t)].data()});
    for (int i = 0; i < nmat; ++i) {
        F3 a{0.1f + 0.7f * rng.u(kAlbedo, uint32_t(i) * 3), 0.1f + 0.7f * rng.u(kAlbedo, uint32_t(i) * 3 + 1),
             0.1f + 0.7f * rng.u(kAlbedo, uint32_t(i) * 3 + 2)};
        int bsdf = i < nmat * 6 / 10 ? PRT_DIFFUSE : (i < nmat * 85 / 100 ? PRT_SPECULAR : PRT_REFRACTIVE);
        prt_material pm = mat_of(bsdf, a);
        if (bsdf == PRT_DIFFUSE && ntex > 0) pm.tex_kd = i % ntex;
        m.materials.push_back(pm);
        m.material_names.push_back("mat" + std::to_string(i));
    }
    const uint32_t lightMat = uint32_t(m.materials.size());
    m.materials.push_back(emitter(emission));
    m.material_names.push_back("light");
    // Terrain: height field over [-extent/2, extent/2]^2, material 0 (textured).
    std::vector<F3> tv(size_t(grid) * grid), tn(size_t(grid) * grid);
    const float cell = extent / float(grid - 1);
    for (int j = 0; j < grid; ++j)
        for (int i = 0; i < grid; ++i) {
            float x = -extent / 2 + cell * i, z = -extent / 2 + cell * j;
            float y = 8.f * fbm(rng, x * 0.02f, z * 0.02f, 6);
            tv[size_t(j) * grid + i] = F3{x, y, z};
        }
    for (int j = 0; j < grid; ++j)
        for (int i = 0; i < grid; ++i) {
            const F3& l = tv[size_t(j) * grid + std::max(0, i - 1)];
            const F3& r = tv[size_t(j) * grid + std::min(grid - 1, i + 1)];
            const F3& d = tv[size_t(std::max(0, j - 1)) * grid + i];
            const F3& u = tv[size_t(std::min(grid - 1, j + 1)) * grid + i];
            tn[size_t(j) * grid + i] = normalize(cross(u - d, r - l));
        }
    for (int j = 0; j + 1 < grid; ++j)
        for (int i = 0; i + 1 < grid; ++i) {
            uint32_t a = uint32_t(j) * grid + i, b = a + 1, c = a + grid, d = c + 1;
            const uint32_t tri[2][3] = {{a, c, b}, {b, c, d}};
            for (auto& t : tri) {
                float pp[9], nn[9], uv[6];
                for (int k = 0; k < 3; ++k) {
                    const F3& p = tv[t[k]];
                    const F3& n = tn[t[k]];
                    pp[k * 3] = p.x; pp[k * 3 + 1] = p.y; pp[k * 3 + 2] = p.z;
                    nn[k * 3] = n.x; nn[k * 3 + 1] = n.y; nn[k * 3 + 2] = n.z;
                    uv[k * 2] = (p.x / extent + 0.5f) * 8.f;
                    uv[k * 2 + 1] = (p.z / extent + 0.5f) * 8.f;
                }
                m.add_triangle(pp, nn, uv, 0);
            }
        }
    Sphere s = uvsphere(slices, stacks);
    for (int i = 0; i < instances; ++i) {
        uint32_t k = uint32_t(i) * 3;
        float x = (rng.u(kCenters, k) - 0.5f) * extent * 0.9f, z = (rng.u(kCenters, k + 2) - 0.5f) * extent * 0.9f;
        float r = 0.8f + 1.2f * rng.u(kRadius, uint32_t(i));
        F3 c{x, 8.f * fbm(rng, x * 0.02f, z * 0.02f, 6) + r + 0.5f * rng.u(kCenters, k + 1), z};
        add_sphere(m, s, c, r, 0.05f * r, rng, uint32_t(i), uint32_t(1 + i % (nmat - 1)));
    }
    for (float lx : {-30.f, 30.f})
        add_quad(m, F3{lx - 5, 60, -5}, F3{lx + 5, 60, -5}, F3{lx + 5, 60, 5}, F3{lx - 5, 60, 5}, F3{0, -1, 0}, lightMat);
    const float eye[3] = {0.f, 35.f, 110.f}, tgt[3] = {0.f, 0.f, 0.f}, up[3] = {0.f, 1.f, 0.f};
    camera_look_at(eye, tgt, up, 50.f, w, h, m.camera);
    m.has_camera = true;
    return PRT_OK;
}

}  // namespace

int generate_scene(const std::string& kind, const std::string& params, uint32_t seed, prt_mesh& m) {
    m = prt_mesh();
    Params P(params);
    int rc;
    if (kind == "spheres") rc = gen_spheres(P, seed, m);
    else if (kind == "cornell-spheres") rc = gen_cornell_spheres(P, seed, m);
    else if (kind == "mixed") rc = gen_mixed(P, seed, m);
    else if (kind == "terrain-spheres") rc = gen_terrain_spheres(P, seed, m);
    else return fail(PRT_ERR_INVALID_ARG, "unknown scene kind: " + kind);
    if (rc != PRT_OK) return rc;
    m.mesh_names.push_back(kind);
    m.mesh_first_tri.push_back(0);
    if (m.materials.size() > PRT_MAX_MATERIALS) return fail(PRT_ERR_LIMIT, "more than 128 materials");
    return PRT_OK;
}

}  // namespace prt
