// Persistent wavefront path tracer and batched rendering (placeholder until the
// queue kernels land): every entry point reports PRT_ERR_UNSUPPORTED.
#include "../host/common.hpp"
#include "renderer.hpp"

namespace prt {

int render_wavefront(prt_renderer*, uint32_t) { return fail(PRT_ERR_UNSUPPORTED, "wavefront integrator not built"); }
void release_wavefront(prt_renderer*) {}

}  // namespace prt

extern "C" int prt_render_batch(prt_job*, uint32_t, const prt_render_config*, const prt_tune*, prt_stats*) {
    return prt::fail(PRT_ERR_UNSUPPORTED, "batched wavefront not built");
}

extern "C" int prt_autotune(prt_job*, uint32_t, const prt_render_config*, prt_tune*, double*) {
    return prt::fail(PRT_ERR_UNSUPPORTED, "autotuner not built");
}

extern "C" int prt_adjoint(prt_renderer*, const float*, const float*, uint32_t, float*, float*, float*) {
    return prt::fail(PRT_ERR_UNSUPPORTED, "adjoint not built");
}
