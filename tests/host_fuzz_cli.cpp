// Host-only driver for the command-line front end, built by
// tests/test_host_sanitizers.py with AddressSanitizer and UBSan (g++, no GPU code):
// random argument lists through parse(), then run()'s whole load / render / save
// sequence over a backend that returns a fixed image instead of rendering.
#include "cli.hpp"

#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

namespace {

class FixedImage : public prt::cli::Backend {
public:
    const char* name() const override { return "fixed"; }
    int create(const prt_scene_desc& desc, const prt_camera&, const prt::cli::Options& o) override {
        w_ = uint32_t(o.width);
        h_ = uint32_t(o.height);
        return desc.num_triangles > 0 ? PRT_OK : PRT_ERR_INVALID_ARG;
    }
    int render_one() override {
        ++n_;
        return PRT_OK;
    }
    int image_bytes(uint8_t* rgba) override {
        for (size_t i = 0; i < size_t(w_) * h_; ++i) {
            rgba[4 * i] = uint8_t(i);
            rgba[4 * i + 1] = uint8_t(i >> 8);
            rgba[4 * i + 2] = uint8_t(n_);
            rgba[4 * i + 3] = 255;
        }
        return PRT_OK;
    }
    int radiance(float* rgb) override {
        for (size_t i = 0; i < size_t(w_) * h_ * 3; ++i) rgb[i] = float(i % 7) * 0.125f;
        return PRT_OK;
    }
    uint32_t samples_rendered() const override { return n_; }

private:
    uint32_t w_ = 0, h_ = 0, n_ = 0;
};

int call_run(const std::vector<std::string>& args) {
    std::vector<std::string> a = {"prt"};
    a.insert(a.end(), args.begin(), args.end());
    std::vector<char*> argv;
    for (auto& s : a) argv.push_back(s.data());
    argv.push_back(nullptr);
    FixedImage be;
    return prt::cli::run(int(a.size()), argv.data(), be);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    const std::string dir = argv[1], obj = argv[2];
    const std::vector<std::string> pool = {"-s", "-w", "-h", "-f", "-o", "--env", "--pfm", "--emit", "--integrator",
                                           "--depth", "--seed", "--device", "--cpu", "--x", "--help", "abc", "-1",
                                           "0", "3", "99999999999", "1e3", "=", "name=", "=3", "x=y", "", "-",
                                           "--", "wavefront", "megakernel", "lambert4SG=nan"};
    std::mt19937 g(5);
    int parsed = 0;
    for (int k = 0; k < 3000; ++k) {
        std::vector<std::string> a = {"prt"};
        const int n = int(g() % 9);
        for (int i = 0; i < n; ++i) a.push_back(pool[g() % pool.size()]);
        std::vector<char*> av;
        for (auto& s : a) av.push_back(s.data());
        av.push_back(nullptr);
        prt::cli::Options o;
        parsed += prt::cli::parse(int(a.size()), av.data(), o) == 0;
    }
    // The whole sequence on the bundled Cornell box: PNG, PFM, progressive preview.
    const int rc = call_run({"-f", obj, "-s", "3", "-w", "17", "-h", "9", "--emit", "lambert4SG=5", "-o",
                             dir + "/out.png", "--pfm", dir + "/out.pfm", "--x"});
    // Refusals: no emitter, an unknown material, a missing file.
    const int no_emitter = call_run({"-f", obj, "-s", "1", "-w", "8", "-h", "8", "-o", dir + "/b.png"});
    const int bad_emit = call_run({"-f", obj, "--emit", "nosuch=1", "-o", dir + "/c.png"});
    const int missing = call_run({"-f", dir + "/missing.obj", "-o", dir + "/d.png"});
    std::printf("parsed %d of 3000; run %d, no emitter %d, unknown material %d, missing file %d\n", parsed, rc,
                no_emitter, bad_emit, missing);
    return (rc == 0 && no_emitter != 0 && bad_emit != 0 && missing != 0) ? 0 : 1;
}
