// Throughput probe of the FP32 exponential (v_exp_f32), the instruction that bounds
// the mixed-precision inward integration: every lane runs 8 independent chains of
// x <- exp2(-x) (values stay in (0.5, 1]), so the issue rate, not the latency, is
// measured.  One value per lane is stored so that nothing is optimised away.
#include "common.h"

namespace portal {

constexpr int PROBE_CHAINS = 8;

__global__ void __launch_bounds__(256) exp_probe_kernel(float *out, int iters) {
    float x[PROBE_CHAINS];
#pragma unroll
    for (int c = 0; c < PROBE_CHAINS; c++) x[c] = 0.5f + 0.01f * (float)((threadIdx.x + c) & 31);
    for (int i = 0; i < iters; i++) {
#pragma unroll
        for (int c = 0; c < PROBE_CHAINS; c++) x[c] = __builtin_amdgcn_exp2f(-x[c]);
    }
    float s = 0.0f;
#pragma unroll
    for (int c = 0; c < PROBE_CHAINS; c++) s += x[c];
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

}  // namespace portal

extern "C" {

// Exponentials per second of the whole device (v_exp_f32), measured with HIP events
// over `iters` x 8 exponentials per thread on 8192 workgroups of 256 threads.
int portal_debug_exp_peak(portal_ctx *x, int iters, double *exp_per_s) {
    if (!x || iters <= 0 || !exp_per_s) return PORTAL_ERR_ARG;
    portal::Ctx &c = x->c;
    hipSetDevice(c.device);
    const int blocks = 8192, threads = 256;
    float *d = nullptr;
    if (hipMalloc((void **)&d, sizeof(float) * (size_t)blocks * threads) != hipSuccess) return PORTAL_ERR_NOMEM;
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    hipLaunchKernelGGL(portal::exp_probe_kernel, dim3(blocks), dim3(threads), 0, c.stream, d, 16);  // warm-up
    hipEventRecord(e0, c.stream);
    hipLaunchKernelGGL(portal::exp_probe_kernel, dim3(blocks), dim3(threads), 0, c.stream, d, iters);
    hipEventRecord(e1, c.stream);
    hipError_t err = hipEventSynchronize(e1);
    float ms = 0.0f;
    hipEventElapsedTime(&ms, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(d);
    if (err != hipSuccess || hipGetLastError() != hipSuccess) return PORTAL_ERR_DEVICE;
    *exp_per_s = (double)blocks * threads * portal::PROBE_CHAINS * (double)iters / (ms * 1e-3);
    return PORTAL_OK;
}

}  // extern "C"
