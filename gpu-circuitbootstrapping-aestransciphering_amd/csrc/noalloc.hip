// noalloc.hip -- tfhe_gate_eval_batch_noalloc: the batched gate bootstrap of tfhe_gate_eval_batch
// with its scratch in a fixed device array of the library instead of the context's buffers, so
// that the call allocates nothing (for code under HIP graph capture, serving loops and harnesses
// that forbid hipMalloc).
//
// The scratch holds the blind rotation's N-dimension LWE outputs ([batch][N+1] u32) and the key
// switch's rounded, transposed digits ([N][ks_bpad(batch)] u32); both are written in full by the
// call before they are read, so nothing needs clearing (the key switch zero-fills its output with
// a fill kernel). Each GPU holds its own copy of the array until the process ends; every call on
// one GPU shares it, so calls that overlap on one GPU (two streams, or two host threads) race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/tfhe_amd.h"
#include "capi_common.hpp"
#include "kernels.hpp"

namespace {

constexpr size_t kGateN = 1024;  // the gate ring dimension (the only one the gate context builds)
constexpr size_t kLweWords = (size_t)TFHE_NOALLOC_MAX_GATES * (kGateN + 1);
constexpr size_t kLweWordsPadded = (kLweWords + 63) / 64 * 64;  // keeps the digit region 256-byte aligned
constexpr size_t kKsWords = kGateN * (size_t)TFHE_NOALLOC_MAX_GATES;
constexpr size_t kScratchWords = kLweWordsPadded + kKsWords;

__device__ __align__(256) uint32_t scratch[kScratchWords];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
uint32_t* g_scratch[kMaxDevices];  // device address of `scratch` on each GPU (null: not yet looked up)

// Device address of the array on the current GPU, looked up once per device. The array's name is a
// host-side stand-in, not a device address; only the looked-up pointer is passed to kernels.
int scratch_on_current_device(uint32_t** out) {
    int dev = -1;
    int rc = tfhe::hipcheck(hipGetDevice(&dev), "noalloc: hipGetDevice");
    if (rc) return rc;
    if (dev < 0 || dev >= kMaxDevices) return tfhe::fail(TFHE_E_PARAM, "noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        rc = tfhe::hipcheck(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "noalloc: hipGetSymbolAddress(scratch)");
        if (rc) return rc;
        if (!p) return tfhe::fail(TFHE_E_HIP, "noalloc: scratch has no device address");
        g_scratch[dev] = static_cast<uint32_t*>(p);
    }
    *out = g_scratch[dev];
    return TFHE_OK;
}

}  // namespace

namespace tfhe {

// Looks the array up on the current device at context set-up, so that later calls on that device
// make no runtime call beyond their kernel launches.
int noalloc_prepare_current_device() {
    uint32_t* p = nullptr;
    return scratch_on_current_device(&p);
}

}  // namespace tfhe

extern "C" int tfhe_gate_eval_batch_noalloc(tfhe_gate_ctx* ctx, int gate, const uint32_t* d_ct1,
                                            const uint32_t* d_ct2, uint32_t batch, uint32_t* d_out, void* stream) {
    if (batch > TFHE_NOALLOC_MAX_GATES)
        return tfhe::fail(TFHE_E_CAPACITY,
                          "tfhe_gate_eval_batch_noalloc: batch exceeds TFHE_NOALLOC_MAX_GATES (1024) gates");
    if (tfhe::ks_bpad(batch) > TFHE_NOALLOC_MAX_GATES)
        return tfhe::fail(TFHE_E_CAPACITY, "tfhe_gate_eval_batch_noalloc: padded batch exceeds the scratch");
    uint32_t* base = nullptr;
    int rc = scratch_on_current_device(&base);
    if (rc) return rc;
    return tfhe::gate_eval_with_scratch(ctx, gate, d_ct1, d_ct2, batch, TFHE_NOALLOC_MAX_GATES, d_out, base,
                                        base + kLweWordsPadded, stream);
}
