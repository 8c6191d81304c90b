// probe.hip -- VALU FMA throughput probe (FP32 and FP64), for the measured peak
// that decides the external product's arithmetic: the library is built without
// packed FP32, so a 32-bit FMA issues at one lane-op per cycle, the FP64 rate.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace tfhe {

template <typename T>
__global__ __launch_bounds__(256) void k_fma_chain(T* __restrict__ out, int iters, T a, T b) {
    // 8 independent dependency chains per thread hide the FMA latency
    T x[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) x[k] = (T)(threadIdx.x + k);
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int k = 0; k < 8; ++k) x[k] = fma(x[k], a, b);
    }
    T s = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) s += x[k];
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

int probe_fma_tflops(int fp64, double* tflops) {
    constexpr int blocks = 256 * 8, threads = 256, iters = 4096;
    void* d = nullptr;
    if (hipMalloc(&d, (size_t)blocks * threads * 8) != hipSuccess) return -1;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    float ms = 0.f;
    for (int rep = 0; rep < 2; ++rep) {  // first launch warms up
        (void)hipEventRecord(e0, nullptr);
        if (fp64)
            k_fma_chain<double><<<blocks, threads>>>((double*)d, iters, 0.999999, 1e-7);
        else
            k_fma_chain<float><<<blocks, threads>>>((float*)d, iters, 0.999999f, 1e-7f);
        (void)hipEventRecord(e1, nullptr);
        (void)hipEventSynchronize(e1);
        (void)hipEventElapsedTime(&ms, e0, e1);
    }
    const int rc = hipGetLastError() == hipSuccess && ms > 0.f ? 0 : -1;
    *tflops = 2.0 * 8.0 * iters * (double)blocks * threads / (ms * 1e-3) / 1e12;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(d);
    return rc;
}

}  // namespace tfhe
