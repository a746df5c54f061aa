// Command-line front end: src/main.cpp's option handling and scene set-up, and
// TerminalLauncher's render / save loop (src/context/terminal_launcher.cpp:10-17,
// src/context/launcher.cpp:13-21), over the C-ABI.
//
// Differences from main.cpp, each a fix of a reference defect:
//   - `-o` is parsed (the usage lists it but main.cpp never reads it);
//   - a scene without a camera JSON gets a default camera framing its bounds
//     (main.cpp throws uncaught when `<scene>-camera.json` is missing);
//   - a scene without emitters is refused with a message naming --emit (the
//     reference divides by a zero light area);
//   - no implicit device switch (main.cpp:160-162 picks device 1 when there are
//     several): --device selects it, default 0.
#include "cli.hpp"

#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace prt::cli {
namespace {

void usage() {
    std::fprintf(stderr,
                 "Usage: prt <options>\n"
                 "-o \t<output file>\tdefault: <scene>.png\n"
                 "-s \t<number of samples>\tdefault: 10\n"
                 "-w \t<width>\tdefault: 640px\n"
                 "-h \t<height>\tdefault: 480px\n"
                 "-f \t<scene name, or path to an .obj>\tdefault: cornell\n"
                 "--cpu \tthe CPU renderer (prt_cpu)\n"
                 "--x \trewrite the output image after every sample pass (progressive preview)\n"
                 "--emit <material>=<Ka>\tmake an MTL material an emitter with Ka = (Ka, Ka, Ka)\n"
                 "--pfm <file>\talso write the mean radiance as a float PFM\n"
                 "--integrator megakernel|wavefront\tdefault: megakernel\n"
                 "--depth <n>\tpath vertices, default 6\n"
                 "--seed <n>\tdefault 24301 (0x5EED)\n"
                 "--device <n>\tHIP device, default 0\n"
                 "--env <path>\t.env with MESHES_PATH / CAMERA_PATH / TEXTURES_PATH, default ../.env\n");
}

bool file_exists(const std::string& p) {
    struct stat st;
    return ::stat(p.c_str(), &st) == 0 && S_ISREG(st.st_mode);
}

bool ends_with(const std::string& s, const std::string& t) {
    return s.size() >= t.size() && s.compare(s.size() - t.size(), t.size(), t) == 0;
}

bool to_int(const char* s, long lo, long hi, long* out) {
    char* end = nullptr;
    long v = std::strtol(s, &end, 0);
    if (!s[0] || *end || v < lo || v > hi) return false;
    *out = v;
    return true;
}

// Eye on +z of the bounds' centre, far enough for the larger half-extent to fit
// a 40 degree vertical field of view.
int default_camera(const prt_scene_desc& d, uint32_t w, uint32_t h, prt_camera* cam) {
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (size_t i = 0; i < size_t(d.num_triangles) * 3; ++i)
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], d.positions[i * 3 + k]);
            hi[k] = std::max(hi[k], d.positions[i * 3 + k]);
        }
    const float c[3] = {(lo[0] + hi[0]) / 2, (lo[1] + hi[1]) / 2, (lo[2] + hi[2]) / 2};
    const float ext = std::max({hi[0] - lo[0], hi[1] - lo[1]}) / 2;
    const float vfov = 40.f;
    const float dist = ext / std::tan(vfov * 0.5f * float(M_PI) / 180.f) + (hi[2] - lo[2]) / 2;
    const float eye[3] = {c[0], c[1], c[2] + dist}, up[3] = {0.f, 1.f, 0.f};
    return prt_camera_look_at(eye, c, up, vfov, w, h, cam);
}

}  // namespace

int parse(int argc, char** argv, Options& o) {
    if (argc < 2) {
        usage();
        return 1;
    }
    std::vector<std::string> a(argv + 1, argv + argc);
    for (size_t i = 0; i < a.size(); ++i) {
        const std::string& k = a[i];
        const bool has = i + 1 < a.size();
        long v = 0;
        if (k == "--cpu") {
            o.cpu = true;
        } else if (k == "--x") {
            o.progressive = true;
        } else if (k == "--help") {
            usage();
            return 1;
        } else if (!has && (k == "-s" || k == "-w" || k == "-h" || k == "-f" || k == "-o" || k == "--emit" ||
                            k == "--pfm" || k == "--integrator" || k == "--depth" || k == "--seed" ||
                            k == "--device" || k == "--env")) {
            std::fprintf(stderr, "Missing argument to %s!\n", k.c_str());
            return 2;
        } else if (k == "-s" || k == "-w" || k == "-h" || k == "--depth" || k == "--device") {
            const long hi = k == "--depth" ? 64 : k == "--device" ? 1024 : 1L << 20;
            if (!to_int(a[i + 1].c_str(), k == "--device" ? 0 : 1, hi, &v)) {
                std::fprintf(stderr, "Invalid argument to %s!\n", k.c_str());
                return 2;
            }
            if (k == "-s") o.samples = int(v);
            else if (k == "-w") o.width = int(v);
            else if (k == "-h") o.height = int(v);
            else if (k == "--depth") o.max_depth = uint32_t(v);
            else o.device = int(v);
            ++i;
        } else if (k == "--seed") {
            if (!to_int(a[i + 1].c_str(), 0, 0xFFFFFFFFL, &v)) {
                std::fprintf(stderr, "Invalid argument to --seed!\n");
                return 2;
            }
            o.seed = uint32_t(v);
            ++i;
        } else if (k == "-f") {
            o.scene = a[++i];
        } else if (k == "-o") {
            o.out = a[++i];
        } else if (k == "--pfm") {
            o.pfm = a[++i];
        } else if (k == "--env") {
            o.env = a[++i];
        } else if (k == "--integrator") {
            o.integrator = a[++i];
            if (o.integrator != "megakernel" && o.integrator != "wavefront") {
                std::fprintf(stderr, "Invalid argument to --integrator!\n");
                return 2;
            }
        } else if (k == "--emit") {
            const std::string s = a[++i];
            const size_t eq = s.find('=');
            char* end = nullptr;
            const float ka = eq == std::string::npos ? 0.f : std::strtof(s.c_str() + eq + 1, &end);
            if (eq == std::string::npos || eq == 0 || !end || *end || !(ka > 0.f)) {
                std::fprintf(stderr, "Invalid argument to --emit (expected <material>=<Ka>)!\n");
                return 2;
            }
            o.emit.emplace_back(s.substr(0, eq), ka);
        } else {
            std::fprintf(stderr, "Unknown option %s\n", k.c_str());
            usage();
            return 2;
        }
    }
    // Paths: .env (EnvLoader) or the defaults of main.cpp:49-53.
    char meshes[4096], cameras[4096], textures[4096];
    std::string mdir = "../meshes/", cdir = "../settings/", tdir = "../textures/";
    if (prt_env_load(o.env.c_str(), meshes, cameras, textures, sizeof meshes) == PRT_OK) {
        mdir = std::string(meshes) + "/";
        cdir = std::string(cameras) + "/";
        tdir = std::string(textures) + "/";
    }
    std::string name = o.scene;
    if (ends_with(name, ".obj")) {
        o.obj_path = name;
        const size_t slash = name.find_last_of('/');
        std::string base = name.substr(slash == std::string::npos ? 0 : slash + 1);
        base = base.substr(0, base.size() - 4);
        o.camera_path = name.substr(0, name.size() - 4) + "-camera.json";
        if (!file_exists(o.camera_path)) o.camera_path = cdir + base + "-camera.json";
        name = base;
    } else {
        o.obj_path = mdir + name + ".obj";
        o.camera_path = cdir + name + "-camera.json";
    }
    o.textures_path = tdir;
    if (o.out.empty()) o.out = name + ".png";
    return 0;
}

int run(int argc, char** argv, Backend& be) {
    Options o;
    int prc = parse(argc, argv, o);
    if (prc != 0) return prc;
    if (o.cpu && std::strcmp(be.name(), "cpu") != 0) {
        std::fprintf(stderr,
                     "--cpu: this binary renders on a HIP device only. The CPU renderer (the reference's\n"
                     "SequentialRenderer, restated) is the test oracle's `prt_cpu`, with the same options.\n");
        return 2;
    }
    std::printf("Samples: %d\nWidth: %d\nHeight: %d\nObj path: %s\nCamera path: %s\nOutput: %s\n", o.samples,
                o.width, o.height, o.obj_path.c_str(), o.camera_path.c_str(), o.out.c_str());

    prt_mesh* mesh = nullptr;
    if (prt_mesh_load_obj(o.obj_path.c_str(), &mesh) != PRT_OK) {
        std::fprintf(stderr, "error: %s\n", prt_last_error());
        return 3;
    }
    int rc = 0;
    prt_scene_desc desc;
    prt_camera cam;
    bool any_emitter = false;
    for (const auto& e : o.emit) {
        char nm[256];
        bool found = false;
        for (uint32_t i = 0; prt_mesh_material_name(mesh, i, nm, sizeof nm) == PRT_OK; ++i) {
            if (e.first != nm) continue;
            prt_mesh_get_desc(mesh, &desc);
            prt_material m = desc.materials[i];
            m.bsdf = PRT_EMISSIVE;
            m.ka[0] = m.ka[1] = m.ka[2] = e.second;
            prt_mesh_set_material(mesh, i, &m);
            found = true;
        }
        if (!found) {
            std::fprintf(stderr, "error: --emit: no material named `%s`\n", e.first.c_str());
            prt_mesh_destroy(mesh);
            return 2;
        }
    }
    prt_mesh_get_desc(mesh, &desc);
    // Lights are the triangles of EMISSIVE materials (Scene::constructLightList,
    // scene.cpp:148-165); the MTL `type` key sets it, --emit overrides it.
    for (uint32_t i = 0; i < desc.num_triangles && !any_emitter; ++i)
        any_emitter = desc.materials[desc.material_ids[i]].bsdf == PRT_EMISSIVE;
    if (!any_emitter) {
        std::fflush(stdout);
        std::fprintf(stderr,
                     "error: the scene has no emitter (no material of type EMISSIVE); name one with "
                     "--emit <material>=<Ka>\n");
        prt_mesh_destroy(mesh);
        return 3;
    }
    if (file_exists(o.camera_path)) {
        if (prt_camera_from_json(o.camera_path.c_str(), uint32_t(o.width), uint32_t(o.height), &cam) != PRT_OK) {
            std::fprintf(stderr, "error: %s\n", prt_last_error());
            prt_mesh_destroy(mesh);
            return 3;
        }
    } else {
        std::printf("No camera file; using a default camera framing the scene\n");
        default_camera(desc, uint32_t(o.width), uint32_t(o.height), &cam);
    }
    std::printf("Total number of triangles:\t%u\nMaterials:\t%u\n", desc.num_triangles, desc.num_materials);
    std::printf("\nLaunching prt (%s)\n", be.name());

    std::fflush(stdout);
    if (be.create(desc, cam, o) != PRT_OK) {
        std::fprintf(stderr, "error: %s\n", prt_last_error());
        prt_mesh_destroy(mesh);
        return 4;
    }
    std::vector<uint8_t> img(size_t(o.width) * o.height * 4);
    const auto t0 = std::chrono::high_resolution_clock::now();
    for (int s = 0; s < o.samples && rc == 0; ++s) {
        if (be.render_one() != PRT_OK) rc = 4;
        if (rc == 0 && o.progressive) {
            // The preview: the 8-bit image after this pass, replaced atomically.
            const std::string tmp = o.out + ".tmp";
            if (be.image_bytes(img.data()) != PRT_OK ||
                prt_write_png(tmp.c_str(), uint32_t(o.width), uint32_t(o.height), img.data()) != PRT_OK ||
                std::rename(tmp.c_str(), o.out.c_str()) != 0)
                rc = 5;
        }
    }
    const auto t1 = std::chrono::high_resolution_clock::now();
    if (rc == 0 && be.image_bytes(img.data()) != PRT_OK) rc = 4;
    if (rc == 0 && prt_write_png(o.out.c_str(), uint32_t(o.width), uint32_t(o.height), img.data()) != PRT_OK) rc = 5;
    if (rc == 0 && !o.pfm.empty()) {
        std::vector<float> rgb(size_t(o.width) * o.height * 3);
        if (be.radiance(rgb.data()) != PRT_OK ||
            prt_write_pfm(o.pfm.c_str(), uint32_t(o.width), uint32_t(o.height), rgb.data()) != PRT_OK)
            rc = 5;
    }
    if (rc != 0) {
        std::fprintf(stderr, "error: %s\n", prt_last_error());
    } else {
        const double secs = std::chrono::duration<double>(t1 - t0).count();
        const uint32_t n = be.samples_rendered();
        std::printf("Rendered to %s for %u samples per pixel. \n", o.out.c_str(), n);
        std::printf("Elapsed time in seconds = %g\n", secs);
        std::printf("Average samples per pixel = %g\n", double(n) / secs);
    }
    prt_mesh_destroy(mesh);
    return rc;
}

}  // namespace prt::cli
