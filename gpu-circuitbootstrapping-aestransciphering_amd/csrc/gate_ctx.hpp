// gate_ctx.hpp -- the gate-bootstrapping context (include/tfhe_amd.h, tfhe_gate_ctx), shared by
// the gate entry points and the programmable-function flows.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../../include/tfhe_amd.h"
#include "capi_common.hpp"
#include "device_common.hpp"

// Programmable-bootstrap state, allocated on the first function call: test vectors staged
// through pinned host memory into device slots (one slot per bootstrap of a flow), and the
// flow's [max_batch][n+1] intermediates.
struct GateFuncState {
    static constexpr int kSlots = 24;          // EvalSign over Z_2^32: 8 floors x 2 + 1
    static constexpr uint32_t kSlotWords = 2048;  // ctmod <= 2N
    uint32_t* h_tv = nullptr;   // pinned [kSlots][kSlotWords]
    uint32_t* d_tv = nullptr;   // [kSlots][kSlotWords]
    hipEvent_t uploaded = nullptr;  // the last staging copy; h_tv is rewritten after it
    uint32_t* d_t = nullptr;    // [max_batch][n+1] each
    uint32_t* d_u = nullptr;
    uint32_t* d_c1 = nullptr;
    uint32_t* d_tb = nullptr;
    uint32_t* d_ks = nullptr;
    bool ready = false;
};

struct tfhe_gate_ctx {
    using cplx = tfhe::cplx;
    tfhe_gate_params p{};
    uint32_t max_batch = 0;
    bool has_keys = false;
    std::vector<int32_t> lwe_sk, rlwe_sk;
    cplx* d_bskf = nullptr;       // [n][2l][2][8][64]
    uint32_t* d_ksk = nullptr;    // [N][ks_l][B][n+1]
    uint32_t* d_lweN = nullptr;   // [max_batch][N+1]
    uint32_t* d_ks_at = nullptr;  // [N][ks_bpad(max_batch)] key-switch digits scratch
    // plain-form scratch
    uint32_t* d_acc = nullptr;    // [2][N]
    cplx* d_digf = nullptr;       // [2l][8][64]
    uint32_t* d_t = nullptr;      // [2][N]
    tfhe::DeviceTables tables;
    GateFuncState func;
};

namespace tfhe {
void gate_func_release(tfhe_gate_ctx* c);
}
