// Issue rates of the vector ALU on this device, measured: FP32 FMA (not
// packed: the library is built without packed FP32) and the hardware
// transcendentals the assembly kernels use (v_sin_f32 / v_cos_f32 on a phase
// in revolutions, v_rsq_f32).  Every thread runs 8 independent chains so the
// latency of one chain is hidden by the others and the other waves; the
// chains feed back into themselves, so nothing is folded away, and one
// thread's results are stored at the end.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mba_common.hpp"

namespace {

constexpr int RATE_CHAINS = 8, RATE_ITERS = 4096, RATE_THREADS = 256;

template <int OP>
__global__ __launch_bounds__(RATE_THREADS) void k_rate(float seed, float* __restrict__ out)
{
    float x[RATE_CHAINS];
#pragma unroll
    for (int c = 0; c < RATE_CHAINS; ++c) x[c] = seed + 0.001f * (float)(threadIdx.x + c);
    const float a = 0.999f, b = 0.0001f;
    for (int it = 0; it < RATE_ITERS; ++it) {
#pragma unroll
        for (int c = 0; c < RATE_CHAINS; ++c) {
            if (OP == 0) x[c] = fmaf(x[c], a, b);                        // one FMA
            if (OP == 1) x[c] = __builtin_amdgcn_sinf(x[c]);             // one v_sin_f32
            if (OP == 2) x[c] = __builtin_amdgcn_cosf(x[c]);             // one v_cos_f32
            if (OP == 3) x[c] = __builtin_amdgcn_rsqf(x[c]);             // one v_rsq_f32 (x stays > 0)
        }
    }
    float s = 0.f;
#pragma unroll
    for (int c = 0; c < RATE_CHAINS; ++c) s += x[c];
    if (s == 12345.678f) out[blockIdx.x * RATE_THREADS + threadIdx.x] = s;  // keeps the chains live
    if (blockIdx.x == 0 && threadIdx.x == 0) out[0] = s;
}

}  // namespace

extern "C" {

// out[4] = operations per second: FP32 FMA (one FMA = one operation, two
// flops), v_sin_f32, v_cos_f32, v_rsq_f32, each over 256 threads x 8 chains x
// 4096 iterations per workgroup and `blocks_per_cu` workgroups per CU (0: 8).
// Synchronises.  No reference counterpart (a measurement of the device).
int mba_measure_valu_rates(double* out, int blocks_per_cu, mba_stream_t stream)
{
    if (!out) return MBA_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        return MBA_ERR_LAUNCH;
    const int per = blocks_per_cu > 0 ? blocks_per_cu : 8;
    const unsigned grid = (unsigned)(cus * per);
    float* buf = nullptr;
    if (hipMalloc((void**)&buf, (size_t)grid * RATE_THREADS * sizeof(float)) != hipSuccess) return MBA_ERR_ALLOC;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = MBA_OK;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = MBA_ERR_LAUNCH;
    const double ops = (double)grid * RATE_THREADS * RATE_CHAINS * RATE_ITERS;
    for (int op = 0; op < 4 && rc == MBA_OK; ++op) {
        auto launch = [&](float seed) {
            if (op == 0) hipLaunchKernelGGL(k_rate<0>, dim3(grid), dim3(RATE_THREADS), 0, st, seed, buf);
            if (op == 1) hipLaunchKernelGGL(k_rate<1>, dim3(grid), dim3(RATE_THREADS), 0, st, seed, buf);
            if (op == 2) hipLaunchKernelGGL(k_rate<2>, dim3(grid), dim3(RATE_THREADS), 0, st, seed, buf);
            if (op == 3) hipLaunchKernelGGL(k_rate<3>, dim3(grid), dim3(RATE_THREADS), 0, st, seed, buf);
        };
        launch(0.25f);  // warm-up
        float ms = 0.f;
        if (hipEventRecord(e0, st) != hipSuccess) rc = MBA_ERR_LAUNCH;
        for (int r = 0; r < 3; ++r) launch(0.3f + 0.01f * r);
        if (rc == MBA_OK && (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                             hipEventElapsedTime(&ms, e0, e1) != hipSuccess || hipGetLastError() != hipSuccess))
            rc = MBA_ERR_LAUNCH;
        if (rc == MBA_OK) out[op] = 3.0 * ops / (ms * 1e-3);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(buf);
    return rc;
}

}  // extern "C"
