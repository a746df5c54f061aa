// `prt`: the command-line renderer on a HIP device (ParallelRenderer path of
// src/main.cpp:157-171). There is no CPU fallback: without a device it exits
// with an error.
#include <cstdio>

#include "cli.hpp"

namespace {

class GpuBackend : public prt::cli::Backend {
public:
    ~GpuBackend() override {
        if (r_) prt_renderer_destroy(r_);
        if (s_) prt_scene_destroy(s_);
    }
    const char* name() const override { return "hip"; }
    int create(const prt_scene_desc& desc, const prt_camera& cam, const prt::cli::Options& o) override {
        int n = 0;
        if (prt_device_count(&n) != PRT_OK || n == 0) {
            std::fprintf(stderr, "no HIP device is visible\n");
            return PRT_ERR_NO_DEVICE;
        }
        std::printf("Got %d HIP devices; using device %d\n", n, o.device);
        int rc = prt_scene_create(&desc, &s_);
        if (rc != PRT_OK) return rc;
        prt_render_config cfg;
        prt_render_config_default(&cfg);
        cfg.max_depth = o.max_depth;
        cfg.seed = o.seed;
        cfg.device = o.device;
        cfg.integrator = o.integrator == "wavefront" ? PRT_INTEGRATOR_WAVEFRONT : PRT_INTEGRATOR_MEGAKERNEL;
        return prt_renderer_create(s_, &cam, uint32_t(o.width), uint32_t(o.height), uint32_t(o.samples), &cfg, &r_);
    }
    int render_one() override { return prt_render_spp(r_, 1); }
    int image_bytes(uint8_t* rgba) override { return prt_copy_image_bytes(r_, rgba); }
    int radiance(float* rgb) override { return prt_get_radiance(r_, rgb); }
    uint32_t samples_rendered() const override {
        uint32_t n = 0;
        prt_renderer_samples_rendered(r_, &n);
        return n;
    }

private:
    prt_scene* s_ = nullptr;
    prt_renderer* r_ = nullptr;
};

}  // namespace

int main(int argc, char** argv) {
    GpuBackend be;
    return prt::cli::run(argc, argv, be);
}
