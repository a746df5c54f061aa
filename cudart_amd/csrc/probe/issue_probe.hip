// Calibration kernel for instruction-issue measurements: the VALU issue rate a
// kernel can reach on this GPU when nothing but independent fp32 FMAs are issued.
// bench.py --issue-probe launches it; rocprofv3 counts its SQ_INSTS_VALU, and the
// traversal kernels' VALU instructions per second are quoted as a fraction of
// this kernel's. Not part of the renderer (built into its own libprt_probe.so).
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

constexpr int kChains = 8;

// Every thread runs kChains independent multiply-add chains for `iters` steps:
// no memory traffic in the loop, no divergence, full occupancy at 256 threads.
__global__ __launch_bounds__(256) void valu_fma_probe(float* out, uint32_t iters, float a, float b) {
    float x[kChains];
#pragma unroll
    for (int k = 0; k < kChains; ++k) x[k] = float(threadIdx.x + k) * 1e-3f;
    for (uint32_t i = 0; i < iters; ++i) {
#pragma unroll
        for (int k = 0; k < kChains; ++k) x[k] = __builtin_fmaf(x[k], a, b);
    }
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < kChains; ++k) s += x[k];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

}  // namespace

// Launches the probe `launches` times over `blocks` workgroups of 256 threads with
// `iters` loop steps (kChains FMAs each), into an output buffer of its own. Returns
// the device time of all launches in ms through *ms; 0, or a negative hipError_t.
extern "C" int prt_probe_valu(uint32_t blocks, uint32_t iters, uint32_t launches, float* ms) {
    float* out = nullptr;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&out), size_t(blocks) * 256 * sizeof(float));
    if (e != hipSuccess) return -int(e);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipEventRecord(e0, nullptr);
    for (uint32_t l = 0; l < launches && e == hipSuccess; ++l) {
        hipLaunchKernelGGL(valu_fma_probe, dim3(blocks), dim3(256), 0, nullptr, out, iters, 0.999f, 1e-4f);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(out);
    return e == hipSuccess ? 0 : -int(e);
}
