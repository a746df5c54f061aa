// Scene input: OBJ/MTL, camera JSON, .env, and the prt_mesh C entry points.
//
// The OBJ/MTL reader reproduces the format semantics of the vendored objl loader
// (src/include/obj_load.h) that CudaRT feeds into Scene::loadTriangles
// (src/geometry/scene.cpp:74-138):
//   - a new mesh starts at "o"/"g" (or any line starting with 'g') once the current
//     one has faces, and at "usemtl" inside a mesh (obj_load.h:483-530, 600-627);
//   - the i-th "usemtl" name is given to the i-th mesh (obj_load.h:670-684);
//   - faces accept v, v/t, v//n, v/t/n with negative (relative) indices
//     (obj_load.h:737-844, getElement :208-217); a face with any vertex lacking a
//     normal gets cross(p0-p1, p2-p1) on every vertex (:832-844);
//   - quads become (0,1,3),(1,2,3) as VertexTriangluation's 4-vertex case does
//     (:920-951); larger polygons are fanned from vertex 0;
//   - MTL keys Ka Kd Ks Ns Ni d illum map_Kd and CudaRT's "type" and "diffuse"
//     (:1016-1186), with objl::Material() defaults.
// Fixes: triangles are counted from the face indices (F17), one material per
// newmtl name instead of the value dedupe of scene.cpp:82-97 (F5), unknown indices
// are an error instead of undefined behaviour.
#include <cfloat>
#include <cmath>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>

#include "common.hpp"
#include "mesh.hpp"

namespace prt {
namespace {

std::string first_token(const std::string& in) {
    size_t s = in.find_first_not_of(" \t\r");
    if (s == std::string::npos) return "";
    size_t e = in.find_first_of(" \t\r", s);
    return in.substr(s, e == std::string::npos ? std::string::npos : e - s);
}
std::string tail(const std::string& in) {
    size_t s = in.find_first_not_of(" \t\r");
    if (s == std::string::npos) return "";
    size_t sp = in.find_first_of(" \t\r", s);
    if (sp == std::string::npos) return "";
    size_t t = in.find_first_not_of(" \t\r", sp);
    if (t == std::string::npos) return "";
    size_t e = in.find_last_not_of(" \t\r");
    return in.substr(t, e - t + 1);
}
std::vector<std::string> words(const std::string& in) {
    std::vector<std::string> out;
    std::istringstream ss(in);
    std::string w;
    while (ss >> w) out.push_back(w);
    return out;
}
std::vector<std::string> split_keep_empty(const std::string& in, char sep) {
    std::vector<std::string> out;
    std::string cur;
    for (char ch : in) {
        if (ch == sep) {
            out.push_back(cur);
            cur.clear();
        } else {
            cur.push_back(ch);
        }
    }
    out.push_back(cur);
    return out;
}
float to_float(const std::string& s) {
    size_t used = 0;
    float v = std::stof(s, &used);
    return v;
}
template <class V>
const V& element(const std::vector<V>& v, const std::string& s, int line) {
    long i = std::stol(s);
    long k = i < 0 ? long(v.size()) + i : i - 1;
    if (k < 0 || k >= long(v.size()))
        throw std::runtime_error("OBJ line " + std::to_string(line) + ": index " + s + " out of range");
    return v[size_t(k)];
}

struct V3 {
    float x = 0, y = 0, z = 0;
};
struct V2 {
    float x = 0, y = 0;
};
struct Vertex {
    V3 p, n;
    V2 t;
};
struct MtlRec {
    std::string name;
    float ka[3] = {0, 0, 0}, kd[3] = {0, 0, 0}, ks[3] = {0, 0, 0};
    float ns = 0, ni = 0, diffuse = 0;
    std::string type = "LAMBERT";
    std::string map_kd;
};
struct MeshRec {
    std::string name;
    std::vector<Vertex> verts;
    std::vector<uint32_t> idx;
};

// LoadMaterials (obj_load.h:1016-1186).
void load_mtl(const std::string& path, std::vector<MtlRec>& out) {
    std::ifstream f(path);
    if (!f.is_open()) return;  // objl returns false and the OBJ load goes on
    MtlRec cur;
    bool listening = false;
    std::string line;
    while (std::getline(f, line)) {
        std::string tok = first_token(line);
        std::string rest = tail(line);
        if (tok == "newmtl") {
            if (listening) out.push_back(cur);
            cur = MtlRec();
            listening = true;
            cur.name = line.size() > 7 ? rest : "none";
        } else if (tok == "Ka" || tok == "Kd" || tok == "Ks") {
            auto w = words(rest);
            if (w.size() != 3) continue;
            float* dst = tok == "Ka" ? cur.ka : (tok == "Kd" ? cur.kd : cur.ks);
            for (int k = 0; k < 3; ++k) dst[k] = to_float(w[size_t(k)]);
        } else if (tok == "Ns") {
            cur.ns = to_float(rest);
        } else if (tok == "Ni") {
            cur.ni = to_float(rest);
        } else if (tok == "map_Kd") {
            cur.map_kd = rest;
        } else if (tok == "type") {
            cur.type = rest;
        } else if (tok == "diffuse") {
            cur.diffuse = to_float(rest);
        }
    }
    out.push_back(cur);
}

int bsdf_from_type(const std::string& t) {
    // reflDict, src/geometry/scene.cpp:27-34; unknown names stay DIFFUSE (:185-188).
    static const std::map<std::string, int> dict = {{"LAMBERT", PRT_DIFFUSE},     {"SPECULAR", PRT_SPECULAR},
                                                    {"DIFFSPEC", PRT_DIFFSPEC},   {"MICROFACET", PRT_MICROFACET},
                                                    {"REFRACTIVE", PRT_REFRACTIVE}, {"EMISSIVE", PRT_EMISSIVE}};
    auto it = dict.find(t);
    return it == dict.end() ? PRT_DIFFUSE : it->second;
}

}  // namespace

prt_material default_material() {
    prt_material m;
    std::memset(&m, 0, sizeof m);
    m.bsdf = PRT_DIFFUSE;
    m.ni = 0.0f;
    m.diffuse_coef = 0.0f;
    m.tex_kd = PRT_NO_TEXTURE;
    return m;
}

int load_obj(const std::string& path, prt_mesh& out) {
    if (path.size() < 4 || path.substr(path.size() - 4) != ".obj")
        return fail(PRT_ERR_INVALID_ARG, "not an .obj path: " + path);
    std::ifstream f(path);
    if (!f.is_open()) return fail(PRT_ERR_IO, "cannot open " + path);
    std::vector<V3> pos, nrm;
    std::vector<V2> tex;
    std::vector<MeshRec> meshes;
    std::vector<std::string> matNames;
    std::vector<MtlRec> mtls;
    MeshRec cur;
    bool listening = false;
    std::string meshname;
    std::string line;
    int lineNo = 0;
    try {
        while (std::getline(f, line)) {
            ++lineNo;
            std::string tok = first_token(line);
            if (tok == "o" || tok == "g" || (!line.empty() && line[0] == 'g')) {
                std::string nm = (tok == "o" || tok == "g") ? tail(line) : std::string("unnamed");
                if (!listening) {
                    listening = true;
                    meshname = nm;
                } else if (!cur.idx.empty() && !cur.verts.empty()) {
                    cur.name = meshname;
                    meshes.push_back(std::move(cur));
                    cur = MeshRec();
                    meshname = tail(line);
                } else {
                    meshname = nm;
                }
            }
            if (tok == "v" || tok == "vn") {
                auto w = words(tail(line));
                if (w.size() < 3) throw std::runtime_error("OBJ line " + std::to_string(lineNo) + ": short vector");
                V3 v{to_float(w[0]), to_float(w[1]), to_float(w[2])};
                (tok == "v" ? pos : nrm).push_back(v);
            } else if (tok == "vt") {
                auto w = words(tail(line));
                if (w.size() < 2) throw std::runtime_error("OBJ line " + std::to_string(lineNo) + ": short vt");
                tex.push_back(V2{to_float(w[0]), to_float(w[1])});
            } else if (tok == "f") {
                auto fw = words(tail(line));
                std::vector<Vertex> vv;
                bool noNormal = false;
                for (const auto& s : fw) {
                    auto parts = split_keep_empty(s, '/');
                    Vertex v;
                    v.p = element(pos, parts[0], lineNo);
                    if (parts.size() >= 2 && !parts[1].empty()) v.t = element(tex, parts[1], lineNo);
                    if (parts.size() >= 3 && !parts[2].empty()) v.n = element(nrm, parts[2], lineNo);
                    else noNormal = true;
                    vv.push_back(v);
                }
                if (vv.size() < 3) continue;
                if (noNormal) {
                    V3 a{vv[0].p.x - vv[1].p.x, vv[0].p.y - vv[1].p.y, vv[0].p.z - vv[1].p.z};
                    V3 b{vv[2].p.x - vv[1].p.x, vv[2].p.y - vv[1].p.y, vv[2].p.z - vv[1].p.z};
                    V3 c{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
                    for (auto& v : vv) v.n = c;
                }
                uint32_t base = uint32_t(cur.verts.size());
                cur.verts.insert(cur.verts.end(), vv.begin(), vv.end());
                if (vv.size() == 3) {
                    for (uint32_t k : {0u, 1u, 2u}) cur.idx.push_back(base + k);
                } else if (vv.size() == 4) {
                    for (uint32_t k : {0u, 1u, 3u, 1u, 2u, 3u}) cur.idx.push_back(base + k);
                } else {
                    for (uint32_t k = 1; k + 1 < vv.size(); ++k) {
                        cur.idx.push_back(base);
                        cur.idx.push_back(base + k);
                        cur.idx.push_back(base + k + 1);
                    }
                }
            } else if (tok == "usemtl") {
                matNames.push_back(tail(line));
                if (!cur.idx.empty() && !cur.verts.empty()) {
                    cur.name = meshname + "_2";
                    meshes.push_back(std::move(cur));
                    cur = MeshRec();
                }
            } else if (tok == "mtllib") {
                size_t slash = path.find_last_of('/');
                std::string dir = slash == std::string::npos ? "" : path.substr(0, slash + 1);
                load_mtl(dir + tail(line), mtls);
            }
        }
    } catch (const std::exception& e) {
        return fail(PRT_ERR_PARSE, std::string("parse error in ") + path + ": " + e.what());
    }
    if (!cur.idx.empty() && !cur.verts.empty()) {
        cur.name = meshname;
        meshes.push_back(std::move(cur));
    }
    if (meshes.empty()) return fail(PRT_ERR_PARSE, "no faces in " + path);

    out = prt_mesh();
    // One material per newmtl, in MTL order, then objl's default material for
    // meshes that get no name.
    std::map<std::string, uint32_t> byName;
    for (const auto& m : mtls) {
        if (byName.count(m.name)) continue;
        prt_material pm = default_material();
        pm.bsdf = bsdf_from_type(m.type);
        for (int k = 0; k < 3; ++k) {
            pm.kd[k] = m.kd[k];
            pm.ka[k] = m.ka[k];
            pm.ks[k] = m.ks[k];
        }
        pm.ns = m.ns;
        pm.ni = m.ni;
        pm.diffuse_coef = m.diffuse;
        pm.tex_kd = PRT_NO_TEXTURE;  // texture files are not decoded; textures come through prt_scene_desc
        byName[m.name] = uint32_t(out.materials.size());
        out.materials.push_back(pm);
        out.material_names.push_back(m.name);
    }
    int32_t defaultId = -1;
    for (size_t i = 0; i < meshes.size(); ++i) {
        uint32_t mat;
        auto it = i < matNames.size() ? byName.find(matNames[i]) : byName.end();
        if (it != byName.end()) {
            mat = it->second;
        } else {
            if (defaultId < 0) {
                defaultId = int32_t(out.materials.size());
                out.materials.push_back(default_material());
                out.material_names.push_back("not-initialized");
            }
            mat = uint32_t(defaultId);
        }
        const MeshRec& m = meshes[i];
        out.mesh_names.push_back(m.name);
        out.mesh_first_tri.push_back(out.num_triangles());
        for (size_t k = 0; k + 2 < m.idx.size(); k += 3) {
            float p[9], n[9], t[6];
            for (int j = 0; j < 3; ++j) {
                const Vertex& v = m.verts[m.idx[k + size_t(j)]];
                p[j * 3] = v.p.x; p[j * 3 + 1] = v.p.y; p[j * 3 + 2] = v.p.z;
                n[j * 3] = v.n.x; n[j * 3 + 1] = v.n.y; n[j * 3 + 2] = v.n.z;
                t[j * 2] = v.t.x; t[j * 2 + 1] = v.t.y;
            }
            out.add_triangle(p, n, t, mat);
        }
    }
    if (out.materials.size() > PRT_MAX_MATERIALS) return fail(PRT_ERR_LIMIT, "more than 128 materials");
    return PRT_OK;
}

// ---------------------------------------------------------------------------
// A small JSON reader: enough for the Maya camera files (objects, arrays,
// numbers, strings, literals).
// ---------------------------------------------------------------------------
namespace {
struct JVal {
    enum Kind { Null, Num, Str, Arr, Obj, Bool } kind = Null;
    double num = 0;
    std::string str;
    std::vector<JVal> arr;
    std::map<std::string, JVal> obj;
};
struct JParser {
    const std::string& s;
    size_t i = 0;
    void ws() {
        while (i < s.size() && (s[i] == ' ' || s[i] == '\t' || s[i] == '\n' || s[i] == '\r')) ++i;
    }
    [[noreturn]] void bad() { throw std::runtime_error("malformed JSON at offset " + std::to_string(i)); }
    std::string str() {
        if (s[i] != '"') bad();
        ++i;
        std::string out;
        while (i < s.size() && s[i] != '"') {
            if (s[i] == '\\' && i + 1 < s.size()) ++i;
            out.push_back(s[i++]);
        }
        if (i >= s.size()) bad();
        ++i;
        return out;
    }
    JVal val() {
        ws();
        if (i >= s.size()) bad();
        JVal v;
        char c = s[i];
        if (c == '{') {
            v.kind = JVal::Obj;
            ++i;
            ws();
            if (i < s.size() && s[i] == '}') { ++i; return v; }
            while (true) {
                ws();
                std::string k = str();
                ws();
                if (i >= s.size() || s[i] != ':') bad();
                ++i;
                v.obj[k] = val();
                ws();
                if (i < s.size() && s[i] == ',') { ++i; continue; }
                if (i < s.size() && s[i] == '}') { ++i; return v; }
                bad();
            }
        } else if (c == '[') {
            v.kind = JVal::Arr;
            ++i;
            ws();
            if (i < s.size() && s[i] == ']') { ++i; return v; }
            while (true) {
                v.arr.push_back(val());
                ws();
                if (i < s.size() && s[i] == ',') { ++i; continue; }
                if (i < s.size() && s[i] == ']') { ++i; return v; }
                bad();
            }
        } else if (c == '"') {
            v.kind = JVal::Str;
            v.str = str();
        } else if (s.compare(i, 4, "true") == 0) {
            v.kind = JVal::Bool; v.num = 1; i += 4;
        } else if (s.compare(i, 5, "false") == 0) {
            v.kind = JVal::Bool; v.num = 0; i += 5;
        } else if (s.compare(i, 4, "null") == 0) {
            i += 4;
        } else {
            size_t used = 0;
            v.kind = JVal::Num;
            try {
                v.num = std::stod(s.substr(i), &used);
            } catch (...) {
                bad();
            }
            if (used == 0) bad();
            i += used;
        }
        return v;
    }
};
double jnum(const JVal& o, const char* key) {
    auto it = o.obj.find(key);
    if (it == o.obj.end() || it->second.kind != JVal::Num) throw std::runtime_error(std::string("missing number ") + key);
    return it->second.num;
}
F3 jvec(const JVal& o, const char* key) {
    auto it = o.obj.find(key);
    if (it == o.obj.end() || it->second.kind != JVal::Arr || it->second.arr.size() < 3)
        throw std::runtime_error(std::string("missing 3-vector ") + key);
    F3 r;
    r.x = float(it->second.arr[0].num);
    r.y = float(it->second.arr[1].num);
    r.z = float(it->second.arr[2].num);
    return r;
}
}  // namespace

int camera_from_json(const std::string& path, uint32_t w, uint32_t h, prt_camera& cam) {
    if (w == 0 || h == 0) return fail(PRT_ERR_INVALID_ARG, "zero image size");
    std::ifstream f(path);
    if (!f.is_open()) return fail(PRT_ERR_IO, "File " + path + " not found.");
    std::string text((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    try {
        JParser p{text};
        JVal root = p.val();
        if (root.kind != JVal::Obj) throw std::runtime_error("camera JSON is not an object");
        // CameraJsonLoader::getCamera (src/context/json_loader.cpp:26-48).
        float fovDeg = float(jnum(root, "fieldOfView"));
        float focal = float(jnum(root, "focalLength"));
        float fstop = float(jnum(root, "fStop"));
        F3 e = jvec(root, "eye"), d = jvec(root, "viewDirection"), u = jvec(root, "upDirection");
        std::memset(&cam, 0, sizeof cam);
        cam.xpixels = int32_t(w);
        cam.ypixels = int32_t(h);
        cam.fov = std::tan(float(double(fovDeg * 0.5f) * M_PI / 180.0f));
        cam.eye[0] = e.x; cam.eye[1] = e.y; cam.eye[2] = e.z;
        cam.focus_distance = length(d);
        F3 dn = normalize(d), un = normalize(u), rn = normalize(cross(dn, un));
        cam.dir[0] = dn.x; cam.dir[1] = dn.y; cam.dir[2] = dn.z;
        cam.up[0] = un.x; cam.up[1] = un.y; cam.up[2] = un.z;
        cam.right[0] = rn.x; cam.right[1] = rn.y; cam.right[2] = rn.z;
        cam.aperture_width = focal / fstop;
        cam.aspect = float(w) / float(h);
    } catch (const std::exception& ex) {
        return fail(PRT_ERR_PARSE, std::string(ex.what()) + " loading camera!");
    }
    return PRT_OK;
}

void camera_look_at(const float eye[3], const float target[3], const float up[3], float vfov_deg, uint32_t w,
                    uint32_t h, prt_camera& cam) {
    std::memset(&cam, 0, sizeof cam);
    F3 e = f3(eye), t = f3(target), u = f3(up);
    F3 d = t - e;
    F3 dn = normalize(d);
    F3 rn = normalize(cross(dn, u));
    F3 un = normalize(cross(rn, dn));
    for (int k = 0; k < 3; ++k) cam.eye[k] = eye[k];
    cam.dir[0] = dn.x; cam.dir[1] = dn.y; cam.dir[2] = dn.z;
    cam.up[0] = un.x; cam.up[1] = un.y; cam.up[2] = un.z;
    cam.right[0] = rn.x; cam.right[1] = rn.y; cam.right[2] = rn.z;
    // computeCameraRay scales `fov` by a normalised coordinate in [-0.5, 0.5]
    // (camera.cu:36-44), so the image spans the full vertical angle vfov when
    // fov = 2 tan(vfov / 2). (The JSON path keeps the reference's tan(f / 2).)
    cam.fov = 2.0f * std::tan(float(double(vfov_deg * 0.5f) * M_PI / 180.0f));
    cam.aspect = float(w) / float(h);
    cam.aperture_width = 0.0f;
    cam.focus_distance = length(d);
    cam.xpixels = int32_t(w);
    cam.ypixels = int32_t(h);
}

}  // namespace prt

// ===========================================================================
// C entry points
// ===========================================================================
using namespace prt;

extern "C" int prt_mesh_load_obj(const char* obj_path, prt_mesh** out) {
    clear_error();
    if (!obj_path || !out) return fail(PRT_ERR_INVALID_ARG, "null argument");
    PRT_TRY_BEGIN
    std::unique_ptr<prt_mesh> m(new prt_mesh);
    int rc = load_obj(obj_path, *m);
    if (rc != PRT_OK) return rc;
    *out = m.release();
    return PRT_OK;
    PRT_TRY_END
}

extern "C" int prt_mesh_generate(const char* kind, const char* params, uint32_t seed, prt_mesh** out) {
    clear_error();
    if (!kind || !out) return fail(PRT_ERR_INVALID_ARG, "null argument");
    PRT_TRY_BEGIN
    std::unique_ptr<prt_mesh> m(new prt_mesh);
    int rc = generate_scene(kind, params ? params : "", seed, *m);
    if (rc != PRT_OK) return rc;
    *out = m.release();
    return PRT_OK;
    PRT_TRY_END
}

extern "C" int prt_mesh_get_desc(const prt_mesh* m, prt_scene_desc* d) {
    clear_error();
    if (!m || !d) return fail(PRT_ERR_INVALID_ARG, "null argument");
    std::memset(d, 0, sizeof *d);
    d->num_triangles = m->num_triangles();
    d->positions = m->positions.data();
    d->normals = m->normals.data();
    d->uvs = m->uvs.data();
    d->material_ids = m->material_ids.data();
    d->num_materials = uint32_t(m->materials.size());
    d->materials = m->materials.data();
    d->num_textures = uint32_t(m->textures.size());
    d->textures = m->textures.empty() ? nullptr : m->textures.data();
    d->camera = m->camera;
    return PRT_OK;
}

extern "C" int prt_mesh_set_material(prt_mesh* m, uint32_t index, const prt_material* mat) {
    clear_error();
    if (!m || !mat) return fail(PRT_ERR_INVALID_ARG, "null argument");
    if (index >= m->materials.size()) return fail(PRT_ERR_INVALID_ARG, "material index out of range");
    m->materials[index] = *mat;
    return PRT_OK;
}

extern "C" int prt_mesh_material_name(const prt_mesh* m, uint32_t index, char* buf, size_t cap) {
    clear_error();
    if (!m || !buf) return fail(PRT_ERR_INVALID_ARG, "null argument");
    if (index >= m->materials.size()) return fail(PRT_ERR_INVALID_ARG, "material index out of range");
    const std::string& nm = index < m->material_names.size() ? m->material_names[index] : std::string();
    if (nm.size() + 1 > cap) return fail(PRT_ERR_LIMIT, "name longer than the buffer");
    std::memcpy(buf, nm.c_str(), nm.size() + 1);
    return PRT_OK;
}

extern "C" int prt_mesh_set_camera(prt_mesh* m, const prt_camera* cam) {
    clear_error();
    if (!m || !cam) return fail(PRT_ERR_INVALID_ARG, "null argument");
    m->camera = *cam;
    m->has_camera = true;
    return PRT_OK;
}

extern "C" int prt_mesh_destroy(prt_mesh* m) {
    delete m;
    return PRT_OK;
}

extern "C" int prt_camera_from_json(const char* path, uint32_t w, uint32_t h, prt_camera* out) {
    clear_error();
    if (!path || !out) return fail(PRT_ERR_INVALID_ARG, "null argument");
    PRT_TRY_BEGIN
    return camera_from_json(path, w, h, *out);
    PRT_TRY_END
}

extern "C" int prt_camera_look_at(const float eye[3], const float target[3], const float up[3], float vfov_deg,
                                  uint32_t w, uint32_t h, prt_camera* out) {
    clear_error();
    if (!eye || !target || !up || !out || w == 0 || h == 0) return fail(PRT_ERR_INVALID_ARG, "bad argument");
    camera_look_at(eye, target, up, vfov_deg, w, h, *out);
    return PRT_OK;
}

// EnvLoader (src/context/env_loader.cpp:14-37): KEY=value lines, '#' comments.
extern "C" int prt_env_load(const char* env_path, char* meshes, char* cameras, char* textures, size_t cap) {
    clear_error();
    if (!env_path) return fail(PRT_ERR_INVALID_ARG, "null argument");
    std::ifstream f(env_path);
    if (!f.is_open()) return fail(PRT_ERR_IO, std::string("Error opening ") + env_path);
    const char* keys[3] = {"MESHES_PATH", "CAMERA_PATH", "TEXTURES_PATH"};
    char* dst[3] = {meshes, cameras, textures};
    std::string vals[3];
    std::string line;
    while (std::getline(f, line)) {
        if (!line.empty() && (line[0] == '#' || line[0] == '\n')) continue;
        for (int k = 0; k < 3; ++k) {
            std::string key = keys[k];
            if (line.find(key) != std::string::npos) {
                vals[k] = line.size() > key.size() ? line.substr(key.size() + 1) : "";
                break;
            }
        }
    }
    for (int k = 0; k < 3; ++k) {
        if (!dst[k]) continue;
        if (vals[k].size() + 1 > cap) return fail(PRT_ERR_LIMIT, "path longer than the buffer");
        std::memcpy(dst[k], vals[k].c_str(), vals[k].size() + 1);
    }
    return PRT_OK;
}
