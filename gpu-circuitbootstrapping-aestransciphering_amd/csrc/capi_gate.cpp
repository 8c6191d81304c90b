// capi_gate.cpp -- C ABI of the gate-bootstrapping context (include/tfhe_amd.h).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/tfhe_amd.h"
#include "capi_common.hpp"
#include "device_common.hpp"
#include "gate_ctx.hpp"
#include "kernels.hpp"

using namespace tfhe;

static bool gate_params_supported(const tfhe_gate_params* p) {
    return p && p->N == 1024 && p->bk_logB == 7 && p->bk_l == 3 && p->ks_logB == 2 && p->ks_l == 8 &&
           p->n > 0 && p->n <= 1024 && p->lwe_nshift >= 1 && p->lwe_nshift <= 31 && p->rlwe_nshift >= 1 &&
           p->rlwe_nshift <= 31;
}

static size_t bsk_words(const tfhe_gate_params& p) { return (size_t)p.n * 2 * p.bk_l * 2 * p.N; }
static size_t ksk_words(const tfhe_gate_params& p) {
    return (size_t)p.N * p.ks_l * ((size_t)1 << p.ks_logB) * (p.n + 1);
}

// gate constants in eighths of the LWE modulus, OpenFHE BINGATE order
static int gate_const(int gate) {
    switch (gate) {
        case TFHE_OR: return 5;
        case TFHE_AND: return 7;
        case TFHE_NOR: return 1;
        case TFHE_NAND: return 3;
        case TFHE_XOR_FAST: return 5;
        case TFHE_XNOR_FAST: return 1;
        case TFHE_XOR: return 5;
        case TFHE_XNOR: return 1;
        default: return -1;
    }
}
static int gate_doubled(int gate) {
    return gate == TFHE_XOR || gate == TFHE_XNOR || gate == TFHE_XOR_FAST || gate == TFHE_XNOR_FAST;
}

extern "C" int tfhe_gate_ctx_create(const tfhe_gate_params* params, uint32_t max_batch, tfhe_gate_ctx** out) {
    if (!out) return fail(TFHE_E_PARAM, "tfhe_gate_ctx_create: out is NULL");
    *out = nullptr;
    if (!gate_params_supported(params))
        return fail(TFHE_E_PARAM, "tfhe_gate_ctx_create: only N=1024, BSK 2^7 x 3, KS 2^2 x 8 is built");
    if (max_batch == 0) max_batch = 1;
    auto* c = new tfhe_gate_ctx();
    c->p = *params;
    c->max_batch = max_batch;
    const auto& p = c->p;
    int rc = TFHE_OK;
    rc = rc ? rc : dmalloc((void**)&c->d_bskf, bsk_words(p) / 2 * sizeof(cplx));
    rc = rc ? rc : dmalloc((void**)&c->d_ksk, ksk_words(p) * sizeof(uint32_t));
    rc = rc ? rc : dmalloc((void**)&c->d_lweN, (size_t)max_batch * (p.N + 1) * sizeof(uint32_t));
    rc = rc ? rc : dmalloc((void**)&c->d_ks_at, (size_t)p.N * ks_bpad(max_batch) * sizeof(uint32_t));
    rc = rc ? rc : dmalloc((void**)&c->d_acc, 2 * (size_t)p.N * sizeof(uint32_t));
    rc = rc ? rc : dmalloc((void**)&c->d_digf, (size_t)2 * p.bk_l * 512 * sizeof(cplx));
    rc = rc ? rc : dmalloc((void**)&c->d_t, 2 * (size_t)p.N * sizeof(uint32_t));
    rc = rc ? rc : c->tables.init(p.N);
    rc = rc ? rc : noalloc_prepare_current_device();
    if (rc) {
        tfhe_gate_ctx_destroy(c);
        return rc;
    }
    *out = c;
    return TFHE_OK;
}

extern "C" int tfhe_gate_ctx_destroy(tfhe_gate_ctx* c) {
    if (!c) return TFHE_OK;
    (void)hipFree(c->d_bskf);
    (void)hipFree(c->d_ksk);
    (void)hipFree(c->d_lweN);
    (void)hipFree(c->d_ks_at);
    (void)hipFree(c->d_acc);
    (void)hipFree(c->d_digf);
    (void)hipFree(c->d_t);
    gate_func_release(c);
    c->tables.release();
    delete c;
    return TFHE_OK;
}

// Coefficient BSK (device) -> Fourier BSK in the context.
static int convert_bsk(tfhe_gate_ctx* c, const uint32_t* d_bsk, hipStream_t st) {
    const size_t npoly = bsk_words(c->p) / c->p.N;
    launch_to_fourier1024(d_bsk, npoly, c->tables.T, c->d_bskf, st);
    return check_launch("to_fourier1024");
}

extern "C" int tfhe_gen_bsk32(const tfhe_gate_params* p, uint64_t seed, const int32_t* d_s, const int32_t* d_K,
                              uint32_t* d_bsk, void* stream) {
    if (!gate_params_supported(p) || !d_s || !d_K || !d_bsk) return fail(TFHE_E_PARAM, "tfhe_gen_bsk32: bad argument");
    launch_gen_bsk32(seed, p->n, p->N, p->bk_logB, p->bk_l, p->rlwe_nshift, d_s, d_K, d_bsk, (hipStream_t)stream);
    return check_launch("gen_bsk32");
}

extern "C" int tfhe_gen_ksk32(const tfhe_gate_params* p, uint64_t seed, const int32_t* d_s, const int32_t* d_K,
                              uint32_t* d_ksk, void* stream) {
    if (!gate_params_supported(p) || !d_s || !d_K || !d_ksk) return fail(TFHE_E_PARAM, "tfhe_gen_ksk32: bad argument");
    launch_gen_ksk32(seed, p->n, p->N, p->ks_logB, p->ks_l, p->lwe_nshift, d_s, d_K, d_ksk, (hipStream_t)stream);
    return check_launch("gen_ksk32");
}

extern "C" int tfhe_gate_bt_keygen(tfhe_gate_ctx* c, const int32_t* lwe_sk, uint64_t seed) {
    if (!c || !lwe_sk) return fail(TFHE_E_PARAM, "tfhe_gate_bt_keygen: bad argument");
    const auto& p = c->p;
    c->lwe_sk.assign(lwe_sk, lwe_sk + p.n);
    c->rlwe_sk.resize(p.N);
    tfhe_rlwe_keygen(p.N, seed, DOM_RLWE1_SK, c->rlwe_sk.data());
    int32_t *d_s = nullptr, *d_K = nullptr;
    uint32_t* d_bsk = nullptr;
    int rc = dmalloc((void**)&d_s, p.n * sizeof(int32_t));
    rc = rc ? rc : dmalloc((void**)&d_K, p.N * sizeof(int32_t));
    rc = rc ? rc : dmalloc((void**)&d_bsk, bsk_words(p) * sizeof(uint32_t));
    if (!rc) rc = hipcheck(hipMemcpy(d_s, c->lwe_sk.data(), p.n * sizeof(int32_t), hipMemcpyHostToDevice), "copy sk");
    if (!rc) rc = hipcheck(hipMemcpy(d_K, c->rlwe_sk.data(), p.N * sizeof(int32_t), hipMemcpyHostToDevice), "copy K");
    if (!rc) rc = tfhe_gen_bsk32(&p, seed, d_s, d_K, d_bsk, nullptr);
    if (!rc) rc = tfhe_gen_ksk32(&p, seed, d_s, d_K, c->d_ksk, nullptr);
    if (!rc) rc = convert_bsk(c, d_bsk, nullptr);
    if (!rc) rc = hipcheck(hipDeviceSynchronize(), "keygen sync");
    (void)hipFree(d_s);
    (void)hipFree(d_K);
    (void)hipFree(d_bsk);
    if (!rc) c->has_keys = true;
    return rc;
}

extern "C" int tfhe_gate_set_keys(tfhe_gate_ctx* c, const uint32_t* bsk, const uint32_t* ksk) {
    if (!c || !bsk || !ksk) return fail(TFHE_E_PARAM, "tfhe_gate_set_keys: bad argument");
    const auto& p = c->p;
    uint32_t* d_bsk = nullptr;
    int rc = dmalloc((void**)&d_bsk, bsk_words(p) * sizeof(uint32_t));
    if (!rc) rc = hipcheck(hipMemcpy(d_bsk, bsk, bsk_words(p) * sizeof(uint32_t), hipMemcpyHostToDevice), "copy bsk");
    if (!rc) rc = hipcheck(hipMemcpy(c->d_ksk, ksk, ksk_words(p) * sizeof(uint32_t), hipMemcpyHostToDevice), "copy ksk");
    if (!rc) rc = convert_bsk(c, d_bsk, nullptr);
    if (!rc) rc = hipcheck(hipDeviceSynchronize(), "set_keys sync");
    (void)hipFree(d_bsk);
    if (!rc) c->has_keys = true;
    return rc;
}

extern "C" int tfhe_gate_get_rlwe_sk(const tfhe_gate_ctx* c, int32_t* out) {
    if (!c || !out) return fail(TFHE_E_PARAM, "tfhe_gate_get_rlwe_sk: bad argument");
    if (c->rlwe_sk.empty()) return fail(TFHE_E_STATE, "tfhe_gate_get_rlwe_sk: keys were uploaded, not generated");
    std::memcpy(out, c->rlwe_sk.data(), c->rlwe_sk.size() * sizeof(int32_t));
    return TFHE_OK;
}

static int check_gate_call(tfhe_gate_ctx* c, int gate, const uint32_t* a, const uint32_t* b, uint32_t batch,
                           uint32_t cap) {
    if (!c) return fail(TFHE_E_PARAM, "gate call: ctx is NULL");
    if (!c->has_keys) return fail(TFHE_E_STATE, "gate call: no bootstrapping keys (call tfhe_gate_bt_keygen)");
    if (gate_const(gate) < 0) return fail(TFHE_E_PARAM, "gate call: unsupported gate (AND, OR, NAND, NOR, XOR, XNOR)");
    if (!a || !b) return fail(TFHE_E_PARAM, "gate call: NULL ciphertext");
    if (a == b) return fail(TFHE_E_PARAM, "Input ciphertexts should be independent");
    if (batch > cap) return fail(TFHE_E_CAPACITY, "gate call: batch exceeds the context's max_batch");
    return TFHE_OK;
}
static int check_gate_call(tfhe_gate_ctx* c, int gate, const uint32_t* a, const uint32_t* b, uint32_t batch) {
    return check_gate_call(c, gate, a, b, batch, c ? c->max_batch : 0);
}

extern "C" int tfhe_gate_blind_rotate_batch(tfhe_gate_ctx* c, int gate, const uint32_t* d_ct1, const uint32_t* d_ct2,
                                            uint32_t batch, uint32_t* d_lweN, void* stream) {
    int rc = check_gate_call(c, gate, d_ct1, d_ct2, 0);
    if (rc) return rc;
    if (!d_lweN) return fail(TFHE_E_PARAM, "blind_rotate: NULL output");
    if (batch == 0) return TFHE_OK;
    const uint32_t q1 = (uint32_t)gate_const(gate) * (2u * c->p.N / 8u);
    launch_gate_blind_rotate(d_ct1, d_ct2, batch, c->p.n, gate_doubled(gate), q1, c->d_bskf, c->tables.T, d_lweN,
                             (hipStream_t)stream);
    return check_launch("gate_blind_rotate");
}

extern "C" int tfhe_gate_keyswitch_batch(tfhe_gate_ctx* c, const uint32_t* d_lweN, uint32_t batch, uint32_t* d_out,
                                         void* stream) {
    if (!c || !d_lweN || !d_out) return fail(TFHE_E_PARAM, "keyswitch: bad argument");
    if (!c->has_keys) return fail(TFHE_E_STATE, "keyswitch: no key-switching key");
    if (batch > c->max_batch) return fail(TFHE_E_CAPACITY, "keyswitch: batch exceeds the context capacity");
    if (batch == 0) return TFHE_OK;
    if (launch_keyswitch32(d_lweN, batch, c->p.N, c->p.n, c->p.ks_logB, c->p.ks_l, c->d_ksk, c->d_ks_at, d_out,
                           (hipStream_t)stream))
        return fail(TFHE_E_PARAM, "keyswitch: unsupported shape");
    return check_launch("keyswitch32");
}

namespace tfhe {

int gate_eval_with_scratch(tfhe_gate_ctx* c, int gate, const uint32_t* d_ct1, const uint32_t* d_ct2, uint32_t batch,
                           uint32_t cap, uint32_t* d_out, uint32_t* lweN, uint32_t* ks_at, void* stream) {
    int rc = check_gate_call(c, gate, d_ct1, d_ct2, batch, cap);
    if (rc) return rc;
    if (!d_out) return fail(TFHE_E_PARAM, "eval_batch: NULL output");
    if (batch == 0) return TFHE_OK;
    const hipStream_t st = (hipStream_t)stream;
    const uint32_t q1 = (uint32_t)gate_const(gate) * (2u * c->p.N / 8u);
    launch_gate_blind_rotate(d_ct1, d_ct2, batch, c->p.n, gate_doubled(gate), q1, c->d_bskf, c->tables.T, lweN, st);
    rc = check_launch("gate_blind_rotate");
    if (rc) return rc;
    if (launch_keyswitch32(lweN, batch, c->p.N, c->p.n, c->p.ks_logB, c->p.ks_l, c->d_ksk, ks_at, d_out, st))
        return fail(TFHE_E_PARAM, "keyswitch: unsupported shape");
    return check_launch("keyswitch32");
}

}  // namespace tfhe

extern "C" int tfhe_gate_eval_batch(tfhe_gate_ctx* c, int gate, const uint32_t* d_ct1, const uint32_t* d_ct2,
                                    uint32_t batch, uint32_t* d_out, void* stream) {
    if (!c) return fail(TFHE_E_PARAM, "gate call: ctx is NULL");
    return gate_eval_with_scratch(c, gate, d_ct1, d_ct2, batch, c->max_batch, d_out, c->d_lweN, c->d_ks_at, stream);
}

static uint32_t host_modswitch_2N(uint32_t v, int N) {
    int lg = 0;
    while ((1 << lg) < 2 * N) ++lg;
    return (uint32_t)((((uint64_t)v + (1ull << (31 - lg))) >> (32 - lg)) & (uint64_t)(2 * N - 1));
}

extern "C" int tfhe_gate_eval_plain(tfhe_gate_ctx* c, int gate, const uint32_t* d_ct1, const uint32_t* d_ct2,
                                    uint32_t* d_out, void* stream) {
    int rc = check_gate_call(c, gate, d_ct1, d_ct2, 1);
    if (rc) return rc;
    if (!d_out) return fail(TFHE_E_PARAM, "eval_plain: NULL output");
    const auto& p = c->p;
    hipStream_t st = (hipStream_t)stream;
    std::vector<uint32_t> h1(p.n + 1), h2(p.n + 1);
    rc = hipcheck(hipMemcpyAsync(h1.data(), d_ct1, (p.n + 1) * 4, hipMemcpyDeviceToHost, st), "copy ct1");
    if (!rc) rc = hipcheck(hipMemcpyAsync(h2.data(), d_ct2, (p.n + 1) * 4, hipMemcpyDeviceToHost, st), "copy ct2");
    if (!rc) rc = hipcheck(hipStreamSynchronize(st), "sync");
    if (rc) return rc;
    const int dbl = gate_doubled(gate);
    std::vector<uint32_t> prep(p.n + 1);
    for (int k = 0; k <= p.n; ++k) prep[k] = (h1[k] + h2[k]) * (dbl ? 2u : 1u);
    const uint32_t q1 = (uint32_t)gate_const(gate) * (2u * p.N / 8u);
    launch_plain_init(c->d_acc, q1, host_modswitch_2N(prep[p.n], p.N), st);
    const size_t key_stride = (size_t)2 * p.bk_l * 2 * 512;
    for (int i = 0; i < p.n; ++i) {
        const uint32_t at = host_modswitch_2N(prep[i], p.N);
        const uint32_t m = (2u * p.N - at) & (2u * p.N - 1u);
        if (m == 0) continue;
        launch_plain_step(c->d_acc, c->d_digf, c->d_t, c->d_bskf + (size_t)i * key_stride, m, c->tables.T, st);
    }
    launch_plain_extract(c->d_acc, c->d_lweN, st);
    rc = check_launch("plain steps");
    if (rc) return rc;
    return tfhe_gate_keyswitch_batch(c, c->d_lweN, 1, d_out, stream);
}

extern "C" int tfhe_polymul1024_batch(const int32_t* d_x, const uint32_t* d_y, uint32_t batch, uint32_t* d_out,
                                     void* stream) {
    if (!d_x || !d_y || !d_out) return fail(TFHE_E_PARAM, "polymul: NULL argument");
    if (batch == 0) return TFHE_OK;
    DeviceTables* tables = nullptr;
    int rc = per_device_tables(tables, [](DeviceTables& d) { return d.init(1024); });
    if (rc) return rc;
    launch_polymul1024(d_x, d_y, batch, tables->T, d_out, (hipStream_t)stream);
    return check_launch("polymul1024");
}

extern "C" int tfhe_gate_external_product(tfhe_gate_ctx* c, uint32_t key_index, uint32_t* d_acc, uint32_t* d_out,
                                          void* stream) {
    if (!c || !d_acc || !d_out) return fail(TFHE_E_PARAM, "external_product: NULL argument");
    if (!c->has_keys) return fail(TFHE_E_STATE, "external_product: no bootstrapping key");
    if (key_index >= (uint32_t)c->p.n) return fail(TFHE_E_PARAM, "external_product: key index out of range");
    const size_t key_stride = (size_t)2 * c->p.bk_l * 2 * 512;
    launch_external_product1024(d_acc, c->d_digf, d_out, c->d_bskf + key_index * key_stride, c->tables.T,
                                (hipStream_t)stream);
    return check_launch("external_product1024");
}

// ---------------------------------------------------------------------------
// Host helpers
// ---------------------------------------------------------------------------
extern "C" void tfhe_lwe_keygen(int n, uint64_t seed, int32_t* sk) {
    for (int i = 0; i < n; ++i) sk[i] = secret_bit(seed, DOM_LWE_SK, (uint32_t)i);
}

extern "C" void tfhe_rlwe_keygen(int N, uint64_t seed, uint32_t domain, int32_t* sk) {
    for (int i = 0; i < N; ++i) sk[i] = secret_bit(seed, domain, (uint32_t)i);
}

extern "C" void tfhe_lwe_encrypt(int n, const int32_t* sk, uint64_t seed, uint32_t id, uint32_t m, uint32_t p,
                                 int nshift, uint32_t* ct) {
    uint32_t b = 0;
    for (int i = 0; i < n; ++i) {
        ct[i] = rand32(seed, DOM_ENC, id, (uint32_t)i);
        b += ct[i] * (uint32_t)sk[i];
    }
    b += (uint32_t)noise(seed, DOM_ENC, id, 0, nshift);
    const uint64_t delta = (1ull << 32) / p;
    ct[n] = b + (uint32_t)(m * delta);
}

extern "C" uint32_t tfhe_lwe_decrypt(int n, const int32_t* sk, const uint32_t* ct, uint32_t p) {
    uint32_t ph = ct[n];
    for (int i = 0; i < n; ++i) ph -= ct[i] * (uint32_t)sk[i];
    // round(ph * p / 2^32) mod p
    const uint64_t v = ((uint64_t)ph * p + (1ull << 31)) >> 32;
    return (uint32_t)(v % p);
}
