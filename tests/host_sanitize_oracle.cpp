// Host-only driver for the oracle's render paths, built by
// tests/test_host_sanitizers.py with AddressSanitizer and UBSan (g++, no GPU code):
// small images of generated scenes with every BSDF, textures and area lights, in
// fp32 and fp64, first hits, and the adjoint (expected-loss and pathwise).
#include "oracle_api.h"
#include "prt.h"

#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

int main() {
    const char* kinds[][2] = {{"spheres", "instances=6;level=2;width=24;height=16;materials=8"},
                              {"cornell-spheres", "instances=3;level=2;width=16;height=16"},
                              {"mixed", "tris=400;width=24;height=16"},
                              {"terrain-spheres", "grid=17;instances=4;stacks=6;textures=3;texsize=16;width=24;height=16"}};
    int checked = 0;
    for (const auto& k : kinds) {
        prt_mesh* m = nullptr;
        if (prt_mesh_generate(k[0], k[1], 11, &m) != PRT_OK) {
            std::printf("generate %s failed: %s\n", k[0], prt_last_error());
            return 1;
        }
        prt_scene_desc d;
        prt_mesh_get_desc(m, &d);
        oracle_scene* s = nullptr;
        auto failed = [&](const char* what) {
            std::printf("%s: %s failed\n", k[0], what);
            return 1;
        };
        if (oracle_scene_create(&d, &s) != PRT_OK) return failed("oracle_scene_create");
        const uint32_t w = d.camera.xpixels, h = d.camera.ypixels;  // the generator's image size
        if (w == 0 || h == 0 || w * h > 4096) return failed("small camera");
        oracle_cfg c{6, 3, 0x5EED, 0};
        std::vector<float> acc(size_t(w) * h * 3, 0.f);
        std::vector<double> acc64(size_t(w) * h * 3, 0.0);
        uint64_t cnt[3] = {0, 0, 0};
        if (oracle_render_f32(s, &d.camera, w, h, 0, 4, &c, acc.data(), cnt, 2) != PRT_OK) return failed("render_f32");
        if (oracle_render_f64(s, &d.camera, w, h, 0, 2, &c, acc64.data(), cnt, 2) != PRT_OK) return failed("render_f64");
        std::vector<int32_t> tri(size_t(w) * h), mat(size_t(w) * h);
        std::vector<float> t(size_t(w) * h);
        for (int brute = 0; brute < 2; ++brute)
            if (oracle_first_hit(s, &d.camera, w, h, 1, &c, tri.data(), mat.data(), t.data(), brute, 2) != PRT_OK)
                return failed("first_hit");
        std::vector<float> img(acc.size()), target(acc.size());
        for (size_t i = 0; i < acc.size(); ++i) {
            img[i] = acc[i] / 4.f;
            target[i] = img[i] * 0.9f + 0.01f;
        }
        std::vector<double> da(size_t(d.num_materials) * 3), de(d.num_materials);
        for (int pathwise = 0; pathwise < 2; ++pathwise)
            if (oracle_adjoint(s, &d.camera, w, h, 2, &c, img.data(), target.data(), da.data(), de.data(), 0.01,
                               pathwise, pathwise, 2) != PRT_OK)
                return failed("adjoint");
        for (double x : da)
            if (!std::isfinite(x)) return failed("finite gradients");
        oracle_scene_destroy(s);
        prt_mesh_destroy(m);
        ++checked;
    }
    std::printf("oracle ok: %d scenes\n", checked);
    return 0;
}
