// capi_common.hpp -- error reporting, device allocation and FFT twiddle tables
// shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/tfhe_amd.h"
#include "device_common.hpp"
#include "fft1024.hpp"

namespace tfhe {

int fail(int code, const std::string& msg);
int hipcheck(hipError_t e, const char* what);
int check_launch(const char* what);
int dmalloc(void** p, size_t bytes);

// Twiddle tables of the wavefront FFT (see device_common.hpp for the layout).
struct DeviceTables {
    FFTTables T{};
    cplx* d_buf = nullptr;
    int init(int N) {
        if (N != 1024) return fail(TFHE_E_PARAM, "FFT tables: only N = 1024 is built");
        std::vector<cplx> h(2 * 512 + 2 * N);
        const double PI = 3.14159265358979323846264338327950288;
        for (int k = 0; k < 8; ++k)
            for (int b = 0; b < 64; ++b) {
                // w^b * exp(-2 pi i b k / 512), w = exp(i pi / N)
                const double ang = PI * b / N - 2.0 * PI * (double)((b * k) % 512) / 512.0;
                h[k * 64 + b] = {std::cos(ang), std::sin(ang)};
            }
        for (int e = 0; e < 8; ++e)
            for (int l = 0; l < 64; ++l) {
                const double ang = -2.0 * PI * (double)(((l & 7) * e) % 64) / 64.0;
                h[512 + e * 64 + l] = {std::cos(ang), std::sin(ang)};
            }
        for (int j = 0; j < 2 * N; ++j) {  // w^j, the evaluation points' powers
            const double ang = PI * j / N;
            h[1024 + j] = {std::cos(ang), std::sin(ang)};
        }
        int rc = dmalloc((void**)&d_buf, h.size() * sizeof(cplx));
        if (rc) return rc;
        rc = hipcheck(hipMemcpy(d_buf, h.data(), h.size() * sizeof(cplx), hipMemcpyHostToDevice), "tables");
        if (rc) return rc;
        T.p1t = d_buf;
        T.p2 = d_buf + 512;
        T.rot = d_buf + 1024;
        return TFHE_OK;
    }
    void release() {
        if (d_buf) (void)hipFree(d_buf);
        d_buf = nullptr;
    }
};

// Twiddle tables of the 4-wave N = 2048 FFT (fft1024.hpp): p1t, p2, p3 [4][64],
// tr [4][4][64], rot [2N]; 5888 complex values.
struct DeviceTables1024 {
    FFT1024Tables T{};
    cplx* d_buf = nullptr;
    static constexpr int kCount = 1792 + 4096;
    // host values of the tables, in device order
    static std::vector<cplx> host_values() {
        std::vector<cplx> h(kCount);
        const double PI = 3.14159265358979323846264338327950288;
        const int N = 2048;
        auto put = [&](int idx, double ang) { h[idx] = {std::cos(ang), std::sin(ang)}; };
        for (int k0 = 0; k0 < 4; ++k0)
            for (int b = 0; b < 64; ++b)  // (w^4)^b * W256^(b k0)
                put(k0 * 64 + b, PI * 4.0 * b / N - 2.0 * PI * (double)((b * k0) % 256) / 256.0);
        for (int e = 0; e < 4; ++e)
            for (int L = 0; L < 64; ++L) put(256 + e * 64 + L, -2.0 * PI * (double)(((L & 15) * e) % 64) / 64.0);
        for (int p = 0; p < 4; ++p)
            for (int L = 0; L < 64; ++L) put(512 + p * 64 + L, -2.0 * PI * (double)(((L & 3) * p) % 16) / 16.0);
        for (int r = 0; r < 4; ++r)
            for (int s = 0; s < 4; ++s)
                for (int l = 0; l < 64; ++l) {
                    const int k = (l >> 4) + 4 * ((l >> 2) & 3) + 16 * (l & 3) + 64 * s;
                    put(768 + (r * 4 + s) * 64 + l, PI * r / N - 2.0 * PI * (double)((r * k) % 1024) / 1024.0);
                }
        for (int j = 0; j < 2 * N; ++j) put(1792 + j, PI * j / N);
        return h;
    }
    int init() {
        const std::vector<cplx> h = host_values();
        int rc = dmalloc((void**)&d_buf, h.size() * sizeof(cplx));
        if (rc) return rc;
        rc = hipcheck(hipMemcpy(d_buf, h.data(), h.size() * sizeof(cplx), hipMemcpyHostToDevice), "tables1024");
        if (rc) return rc;
        T.p1t = d_buf;
        T.p2 = d_buf + 256;
        T.p3 = d_buf + 512;
        T.tr = d_buf + 768;
        T.rot = d_buf + 1792;
        return TFHE_OK;
    }
    void release() {
        if (d_buf) (void)hipFree(d_buf);
        d_buf = nullptr;
    }
};

// The same tables rounded to FP32 (the split-precision product).
struct DeviceTables1024F {
    FFT1024TablesF T{};
    cplxf* d_buf = nullptr;
    int init() {
        const std::vector<cplx> h = DeviceTables1024::host_values();
        std::vector<cplxf> f(h.size());
        for (size_t k = 0; k < h.size(); ++k) f[k] = {(float)h[k].re, (float)h[k].im};
        int rc = dmalloc((void**)&d_buf, f.size() * sizeof(cplxf));
        if (rc) return rc;
        rc = hipcheck(hipMemcpy(d_buf, f.data(), f.size() * sizeof(cplxf), hipMemcpyHostToDevice), "tables1024f");
        if (rc) return rc;
        T.p1t = d_buf;
        T.p2 = d_buf + 256;
        T.p3 = d_buf + 512;
        T.tr = d_buf + 768;
        T.rot = d_buf + 1792;
        return TFHE_OK;
    }
    void release() {
        if (d_buf) (void)hipFree(d_buf);
        d_buf = nullptr;
    }
};

// Process-lifetime twiddle tables for the stateless entry points, one set per
// device (tables in one device's memory are not read from another), each built
// once under a lock.
constexpr int kMaxDevices = 64;
template <class Tables, class Init>
int per_device_tables(Tables*& out, Init init) {
    static std::mutex mu;
    static Tables tables[kMaxDevices];
    int dev = 0;
    int rc = hipcheck(hipGetDevice(&dev), "get device");
    if (rc) return rc;
    if (dev < 0 || dev >= kMaxDevices) return fail(TFHE_E_PARAM, "device index out of range");
    std::lock_guard<std::mutex> lock(mu);
    if (!tables[dev].d_buf) {
        rc = init(tables[dev]);
        if (rc) return rc;
    }
    out = &tables[dev];
    return TFHE_OK;
}

// Batched gate bootstrap with the scratch passed in: lweN holds [batch][N+1] u32 (blind-rotation
// outputs), ks_at [N][ks_bpad(batch)] u32 (key-switch digits); cap is the largest batch they hold.
int gate_eval_with_scratch(tfhe_gate_ctx* c, int gate, const uint32_t* d_ct1, const uint32_t* d_ct2, uint32_t batch,
                           uint32_t cap, uint32_t* d_out, uint32_t* lweN, uint32_t* ks_at, void* stream);

// Looks up the no-allocation entry point's device array on the current GPU (once per device).
int noalloc_prepare_current_device();

}  // namespace tfhe
