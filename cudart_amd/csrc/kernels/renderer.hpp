// prt_renderer: one image on one device.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../host/scene.hpp"
#include "kernel_api.hpp"

namespace prt {
struct WavefrontState;
}

struct prt_renderer {
    prt_scene* scene = nullptr;
    prt_camera cam{};
    uint32_t w = 0, h = 0, spp = 0;
    prt_render_config cfg{};
    prt::SceneArgs sc{};
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float4* accum = nullptr;               // w*h running sums
    unsigned long long* counters = nullptr;
    uint8_t* gamma_lut = nullptr;
    int* spill = nullptr;
    uint32_t spill_threads = 0;
    uint32_t tiles_x = 0, tiles = 0;
    uint32_t samples_done = 0;
    uint64_t samples_total = 0;
    double last_ms = 0.0;
    double kernel_ms[3] = {0, 0, 0};  // profile_kernels: extend, shade, shadow of the last wavefront render
    uint32_t kernel_launches[3] = {0, 0, 0};
    prt::WavefrontState* wf = nullptr;
    // keep_samples: radiance of pixel-sample (s, p) at samples[s * w * h + p], s < spp.
    // Valid for the adjoint while every sample since the last reset was rendered with
    // the scene's current materials (samples_version).
    float4* samples = nullptr;
    bool samples_ok = false;
    uint64_t samples_version = 0;
};

namespace prt {
int hip_fail(hipError_t e, const char* what);
// Copies and fills ordered on the caller's stream and waited for. The library never
// uses the legacy (null) stream: HIP fails a legacy-stream operation issued while
// another thread captures a stream into a graph, and ends that capture, so renderers
// driven from several host threads at once would break each other's launches.
inline hipError_t copy_wait(void* dst, const void* src, size_t bytes, hipMemcpyKind kind, hipStream_t s) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, s);
    return e == hipSuccess ? hipStreamSynchronize(s) : e;
}
inline hipError_t fill_wait(void* dst, int value, size_t bytes, hipStream_t s) {
    hipError_t e = hipMemsetAsync(dst, value, bytes, s);
    return e == hipSuccess ? hipStreamSynchronize(s) : e;
}
// Uploads the scene to `device` on first use (copies ordered on `stream`).
int device_scene(prt_scene* s, int device, SceneArgs& out, hipStream_t stream);
CamArgs cam_args(const prt_camera& c);
int set_device(int device);
int ensure_spill(prt_renderer* r, uint32_t threads);
int render_wavefront(prt_renderer* r, uint32_t n);  // wavefront.hip
// Samples 0..spp-1 of the renderer's view into `accum` (zeroed) and each sample's
// radiance into `samples` ([sample * pixels + pixel]), without touching the
// renderer's own accumulation or counters.
int wavefront_samples(prt_renderer* r, uint32_t spp, float4* accum, float4* samples);  // wavefront.hip
void release_wavefront(prt_renderer* r);            // wavefront.hip
}  // namespace prt
