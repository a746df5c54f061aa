/*
 * ORACLE -- test infrastructure only.
 *
 * `prt_cpu`: the command-line renderer on the host, the reference's --cpu
 * path (SequentialRenderer, src/renderer/sequential_renderer.cu:91-125, chosen at
 * src/main.cpp:157-158) restated by the oracle (fp32, OpenMP over all host threads).
 * Scene loading, options and output are the product CLI's; only the rendering
 * comes from liboracle.
 */
#include <cstdio>
#include <cstring>
#include <vector>

#include "../cudart_amd/csrc/host/cli.hpp"
#include "oracle_api.h"

namespace {

class CpuBackend : public prt::cli::Backend {
public:
    ~CpuBackend() override {
        if (s_) oracle_scene_destroy(s_);
    }
    const char* name() const override { return "cpu"; }
    int create(const prt_scene_desc& desc, const prt_camera& cam, const prt::cli::Options& o) override {
        int rc = oracle_scene_create(&desc, &s_);
        if (rc != PRT_OK) {
            std::fprintf(stderr, "oracle_scene_create failed (%d)\n", rc);
            return rc;
        }
        cam_ = cam;
        w_ = uint32_t(o.width);
        h_ = uint32_t(o.height);
        cfg_.max_depth = o.max_depth;
        cfg_.rr_start = 3;
        cfg_.seed = o.seed;
        cfg_.view_id = 0;
        accum_.assign(size_t(w_) * h_ * 3, 0.f);
        return PRT_OK;
    }
    int render_one() override {
        uint64_t cnt[3] = {0, 0, 0};
        int rc = oracle_render_f32(s_, &cam_, w_, h_, n_, 1, &cfg_, accum_.data(), cnt, 0);
        if (rc == PRT_OK) ++n_;
        return rc;
    }
    int radiance(float* rgb) override {
        if (n_ == 0) return PRT_ERR_INVALID_ARG;
        for (size_t i = 0; i < accum_.size(); ++i) rgb[i] = accum_[i] / float(n_);
        return PRT_OK;
    }
    int image_bytes(uint8_t* rgba) override {
        std::vector<float> mean(accum_.size());
        int rc = radiance(mean.data());
        if (rc != PRT_OK) return rc;
        return oracle_tonemap(w_ * h_, mean.data(), rgba);
    }
    uint32_t samples_rendered() const override { return n_; }

private:
    oracle_scene* s_ = nullptr;
    prt_camera cam_{};
    oracle_cfg cfg_{};
    uint32_t w_ = 0, h_ = 0, n_ = 0;
    std::vector<float> accum_;
};

}  // namespace

int main(int argc, char** argv) {
    CpuBackend be;
    return prt::cli::run(argc, argv, be);
}
