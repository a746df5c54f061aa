// Calibration kernels, not part of the renderer (built into their own
// libprt_probe.so):
//   valu_fma_probe   the VALU issue rate a kernel reaches on this GPU when nothing but
//                    independent fp32 FMAs are issued (bench.py --issue-probe;
//                    rocprofv3 counts its SQ_INSTS_VALU, and the traversal kernels'
//                    VALU rates are quoted as a fraction of it);
//   l1_gather_probe  the rate of 16-byte vector loads from 64-byte lines of a
//                    buffer that fits the L1, the L2 or neither, with the 64 lanes of
//                    each load on 64 different lines, on one line, 16 to a line, four
//                    to a line (each lane one 16-byte part), or with every other lane
//                    masked off (bench.py --l1-probe): the ceiling for the traversal
//                    kernels' per-lane node and triangle fetches.
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

// Every thread makes four independent 16-byte loads per step, `iters` steps, from
// pseudo-random 64-byte lines of buf (line index masked to the buffer: always in
// bounds). MODE 0: the four parts of a line of the lane's own; 1: of the wave's
// first lane's line in every lane; 2: 16 lanes to a line (4 lines per load); 3: each
// load 16 lines, the four lanes of a quad reading the four parts of one line (a
// quad's lanes draw the same lines); 4: as 0 with every other lane masked off.
template <int MODE>
__global__ __launch_bounds__(256) void l1_gather_probe(const float4* buf, uint32_t mask_lines, uint32_t iters,
                                                       float* out) {
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t s = (blockIdx.x * 256u + (MODE == 3 ? threadIdx.x & ~3u : threadIdx.x)) * 2654435761u + 12345u;
    float ax = 0.f, ay = 0.f;
    for (uint32_t i = 0; i < iters; ++i) {
        s = s * 1664525u + 1013904223u;
        if (MODE == 4 && (lane & 1u)) continue;
        float4 a, b, c, d;
        if (MODE == 3) {
            uint32_t l[4] = {s, s * 747796405u + 2891336453u, s * 1103515245u + 12345u, s * 22695477u + 1u};
            const uint32_t part = lane & 3u;
            a = buf[size_t((l[0] >> 8) & mask_lines) * 4 + part];
            b = buf[size_t((l[1] >> 8) & mask_lines) * 4 + part];
            c = buf[size_t((l[2] >> 8) & mask_lines) * 4 + part];
            d = buf[size_t((l[3] >> 8) & mask_lines) * 4 + part];
        } else {
            // The first lane's draw through a lane permute: the compiler cannot tell that
            // it is uniform, so the loads stay vector loads (not scalar ones).
            const uint32_t u = (MODE == 1 || MODE == 2) ? uint32_t(__shfl(int(s), 0)) : s;
            const uint32_t line = ((u >> 8) + (MODE == 2 ? (lane >> 4) : 0u)) & mask_lines;
            const float4* p = buf + size_t(line) * 4;
            a = p[0];
            b = p[1];
            c = p[2];
            d = p[3];
        }
        // Every component is used, so each load stays one 16-byte load (with one
        // component per part the compiler narrows them to 4-byte loads).
        ax += (a.x + a.y) + (a.z + a.w) + (b.x + b.y) + (b.z + b.w);
        ay += (c.x + c.y) + (c.z + c.w) + (d.x + d.y) + (d.z + d.w);
    }
    out[blockIdx.x * blockDim.x + threadIdx.x] = ax + ay;
}

}  // namespace

// Runs l1_gather_probe (mode 0 to 4) `launches` times over `blocks` workgroups of
// 256 threads on a zeroed buffer of `lines` 64-byte lines (a power of two) and an
// output of its own. *ms = the device time of all launches; 0, or a negative
// hipError_t (-1 for a bad argument).
extern "C" int prt_probe_l1(uint32_t lines, uint32_t mode, uint32_t blocks, uint32_t iters, uint32_t launches,
                            float* ms) {
    if (lines == 0 || (lines & (lines - 1)) != 0 || mode > 4 || !ms) return -1;
    float4* buf = nullptr;
    float* out = nullptr;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&buf), size_t(lines) * 64);
    if (e == hipSuccess) e = hipMemset(buf, 0, size_t(lines) * 64);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&out), size_t(blocks) * 256 * sizeof(float));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipEventRecord(e0, nullptr);
    for (uint32_t l = 0; l < launches && e == hipSuccess; ++l) {
        if (mode == 0) hipLaunchKernelGGL(l1_gather_probe<0>, dim3(blocks), dim3(256), 0, nullptr, buf, lines - 1, iters, out);
        if (mode == 1) hipLaunchKernelGGL(l1_gather_probe<1>, dim3(blocks), dim3(256), 0, nullptr, buf, lines - 1, iters, out);
        if (mode == 2) hipLaunchKernelGGL(l1_gather_probe<2>, dim3(blocks), dim3(256), 0, nullptr, buf, lines - 1, iters, out);
        if (mode == 3) hipLaunchKernelGGL(l1_gather_probe<3>, dim3(blocks), dim3(256), 0, nullptr, buf, lines - 1, iters, out);
        if (mode == 4) hipLaunchKernelGGL(l1_gather_probe<4>, dim3(blocks), dim3(256), 0, nullptr, buf, lines - 1, iters, out);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (buf) (void)hipFree(buf);
    if (out) (void)hipFree(out);
    return e == hipSuccess ? 0 : -int(e);
}

// Compute units of the current device, or a negative hipError_t.
extern "C" int prt_probe_cu_count(void) {
    int dev = 0, n = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev);
    return e == hipSuccess ? n : -int(e);
}

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
