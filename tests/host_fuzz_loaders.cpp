// Host-only fuzz driver for the OBJ / MTL / camera-JSON / env loaders, built by
// tests/test_host_sanitizers.py with AddressSanitizer and UBSan (g++, no GPU code):
// hand-written malformed inputs plus 300 random ones; every call must return an
// error code or a mesh, never touch memory it does not own.
#include "prt.h"
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <random>
#include <string>
#include <vector>
static void put(const std::string& p, const std::string& s) { std::ofstream(p, std::ios::binary) << s; }
int main(int argc, char** argv) {
    const std::string d = argc > 1 ? std::string(argv[1]) + "/" : std::string("./");
    std::vector<std::string> objs = {
        "", "v\n", "f 1 2 3\n", "v 0 0 0\nv 1 0 0\nf 1 2 3\n", "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 9\n",
        "v 0 0 0\nv 1 0 0\nv 0 1 0\nf -1 -2 -3\n", "v 0 0 0\nv 1 0 0\nv 0 1 0\nf -4 -2 -1\n",
        "v 0 0 0\nv 1 0 0\nv 0 1 0\nvn 0 0 1\nvt 0 0\nf 1/1/1 2/1/1 3/9/1\n",
        "v 0 0 0\nv 1 0 0\nv 0 1 0\nv 1 1 0\nf 1 2 3 4 1 2 3 4\n", "v nan inf -inf\nv 1e40 0 0\nv 0 1 0\nf 1 2 3\n",
        "mtllib missing.mtl\nusemtl nothing\nv 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n",
        "mtllib bad.mtl\nusemtl m\nv 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n",
        "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1//1 2//1 3//1\n", "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1/ 2/ 3/\n",
        "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 0 0 0\n", "v 0 0\nv 1\nv\nf 1 2 3\n", "f\nf 1\nf 1 2\n",
        "v 0 0 0\nv 1 0 0\nv 0 1 0\ng\no\ns off\nusemtl\nf 1 2 3\n"};
    put(d + "bad.mtl", "newmtl m\nKd\nKa 1\nKs 1 2\nNi\nillum x\nmap_Kd\nnewmtl\nKd 1 1 1\n");
    std::mt19937 g(11);
    const char alphabet[] = "vfntlmuse/ -0123456789.#\n\r\tgo";
    for (int k = 0; k < 300; ++k) {
        std::string s;
        const int len = int(g() % 400);
        for (int i = 0; i < len; ++i) s += alphabet[g() % (sizeof alphabet - 1)];
        objs.push_back(s);
    }
    int ok = 0, err = 0;
    for (size_t i = 0; i < objs.size(); ++i) {
        const std::string p = d + "t" + std::to_string(i) + ".obj";
        put(p, objs[i]);
        prt_mesh* m = nullptr;
        if (prt_mesh_load_obj(p.c_str(), &m) == PRT_OK) {
            prt_scene_desc desc;
            prt_mesh_get_desc(m, &desc);
            prt_mesh_destroy(m);
            ++ok;
        } else {
            ++err;
        }
    }
    const std::vector<std::string> cams = {"", "{", "{}", "{\"camera\": {}}", "[1,2,3]", "{\"a\": [1, 2, {\"b\": \"c\"}]}",
                                           "{\"eye\": [1, 2], \"fov\": \"x\"}", "nul", "{\"x\": 1e999}", "\"unterminated"};
    for (size_t i = 0; i < cams.size(); ++i) {
        const std::string p = d + "c" + std::to_string(i) + ".json";
        put(p, cams[i]);
        prt_camera c;
        (void)prt_camera_from_json(p.c_str(), 64, 48, &c);
    }
    const std::vector<std::string> envs = {"", "MESHES=\n", "=x\n", "MESHES_PATH=/tmp\nCAMERAS_PATH\n", std::string(5000, 'A') + "=B\n"};
    for (size_t i = 0; i < envs.size(); ++i) {
        const std::string p = d + "e" + std::to_string(i) + ".env";
        put(p, envs[i]);
        char a[64], b[64], c[64];
        (void)prt_env_load(p.c_str(), a, b, c, sizeof a);
    }
    // Procedural scenes: small sizes, then one malformed or out-of-range value each.
    const std::vector<std::pair<std::string, std::string>> gens = {
        {"spheres", "instances=3;level=1;width=32;height=24"}, {"cornell-spheres", "instances=2;level=1"},
        {"mixed", "tris=300;width=40"}, {"terrain-spheres", "grid=9;instances=3;stacks=4;textures=2;texsize=8"}};
    const std::vector<std::string> bad = {"", ";;==;", "level=nan", "instances=1e300", "materials=-4", "tris=abc",
                                          "ground=inf", "level=99", "materials=0", "width=-1e20", "texsize=x",
                                          "emission=1e400", "spread", "=5", "radius=-1", "width=-5", "height=0"};
    int gen_ok = 0, gen_err = 0;
    for (const auto& [kind, base] : gens)
        for (const auto& extra : bad) {
            prt_mesh* m = nullptr;
            const std::string params = base + ";" + extra;
            if (prt_mesh_generate(kind.c_str(), params.c_str(), 7, &m) == PRT_OK) {
                prt_scene_desc desc;
                prt_mesh_get_desc(m, &desc);
                prt_mesh_destroy(m);
                ++gen_ok;
            } else {
                ++gen_err;
            }
        }
    prt_mesh* m = nullptr;
    if (prt_mesh_generate("no-such-kind", "", 1, &m) == PRT_OK) return 1;
    std::printf("loaders under asan: %d loaded, %d refused; generators: %d built, %d refused\n", ok, err, gen_ok,
                gen_err);
    return 0;
}
