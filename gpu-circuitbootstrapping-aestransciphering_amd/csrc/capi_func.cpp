// capi_func.cpp -- programmable bootstrap and the large-precision functions of the gate context
// (include/tfhe_amd.h): BootstrapFunc, EvalFunc, EvalFloor, EvalSign and EvalDecomp over batches
// (reference bitwise/src/binfhe-base-scheme.cu:255-698 and binfhe-base-scheme-batch.cu:215-698,
// after Liu, Micciancio and Polyakov, eprint 2021/1337).
//
// Every flow is a fixed sequence of batched launches on the caller's stream: the programmable
// blind rotation (one wavefront per ciphertext, FUNC mode of the gate kernel), the N -> n key
// switch, and element-wise LWE arithmetic over power-of-two moduli.  The flow's test vectors are
// computed on the host up front (each at most 2N words), staged into pinned memory and sent in
// one copy; the control flow (which bootstraps, how many floors) depends only on the moduli and
// on the table's class, never on ciphertext data, so nothing is read back.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "../../include/tfhe_amd.h"
#include "capi_common.hpp"
#include "gate_ctx.hpp"
#include "kernels.hpp"

using namespace tfhe;

namespace {

bool pow2(uint64_t v) { return v >= 2 && !(v & (v - 1)); }
int log2u(uint64_t v) {
    int l = 0;
    while ((1ull << l) < v) ++l;
    return l;
}
uint32_t mask_of(uint64_t mod) { return (uint32_t)(mod - 1u); }

int func_prepare(tfhe_gate_ctx* c) {
    auto& f = c->func;
    if (f.ready) return TFHE_OK;
    const size_t slot_bytes = (size_t)GateFuncState::kSlots * GateFuncState::kSlotWords * sizeof(uint32_t);
    const size_t lwe_bytes = (size_t)c->max_batch * (c->p.n + 1) * sizeof(uint32_t);
    int rc = hipcheck(hipHostMalloc((void**)&f.h_tv, slot_bytes, hipHostMallocDefault), "pinned test vectors");
    rc = rc ? rc : dmalloc((void**)&f.d_tv, slot_bytes);
    rc = rc ? rc : hipcheck(hipEventCreateWithFlags(&f.uploaded, hipEventDisableTiming), "event");
    rc = rc ? rc : dmalloc((void**)&f.d_t, lwe_bytes);
    rc = rc ? rc : dmalloc((void**)&f.d_u, lwe_bytes);
    rc = rc ? rc : dmalloc((void**)&f.d_c1, lwe_bytes);
    rc = rc ? rc : dmalloc((void**)&f.d_tb, lwe_bytes);
    rc = rc ? rc : dmalloc((void**)&f.d_ks, lwe_bytes);
    if (rc) {
        gate_func_release(c);
        return rc;
    }
    f.ready = true;
    return TFHE_OK;
}

// Host-side plan of one flow: its test vectors, in slot order.
struct Tables {
    tfhe_gate_ctx* c;
    int used = 0;
    // slot for tv[x] = f(x) 2^32 / fmod (x < ctmod), f given by a callable over Z_fmod
    template <class F>
    int add(uint64_t ctmod, uint64_t fmod, F fn) {
        if (used >= GateFuncState::kSlots) return -1;
        uint32_t* tv = c->func.h_tv + (size_t)used * GateFuncState::kSlotWords;
        const int sh = 32 - log2u(fmod);
        for (uint64_t x = 0; x < ctmod; ++x) tv[x] = (uint32_t)((uint64_t)(fn(x) & mask_of(fmod)) << sh);
        return used++;
    }
};

int begin_tables(tfhe_gate_ctx* c, Tables& t) {
    t.c = c;
    // the previous flow's staging copy must have read h_tv before it is rewritten
    return hipcheck(hipEventSynchronize(c->func.uploaded), "wait for the previous test-vector copy");
}

int upload_tables(tfhe_gate_ctx* c, const Tables& t, hipStream_t st) {
    const size_t bytes = (size_t)t.used * GateFuncState::kSlotWords * sizeof(uint32_t);
    int rc = hipcheck(hipMemcpyAsync(c->func.d_tv, c->func.h_tv, bytes, hipMemcpyHostToDevice, st), "test vectors");
    return rc ? rc : hipcheck(hipEventRecord(c->func.uploaded, st), "record");
}

// BootstrapFunc over the batch: in over Z_ctmod (read through ctmod - 1), out over Z_fmod,
// masked with omask (fmod - 1, or a smaller power of two minus one).
int bootstrap(tfhe_gate_ctx* c, const uint32_t* d_in, uint32_t batch, uint64_t ctmod, int slot, uint64_t fmod,
              uint32_t omask, uint32_t* d_out, hipStream_t st) {
    const auto& p = c->p;
    const int logf = log2u(2u * (uint64_t)p.N / ctmod);
    launch_func_blind_rotate(d_in, batch, p.n, logf, c->func.d_tv + (size_t)slot * GateFuncState::kSlotWords,
                             c->d_bskf, c->tables.T, c->d_lweN, st);
    int rc = check_launch("func_blind_rotate");
    if (rc) return rc;
    if (launch_keyswitch32(c->d_lweN, batch, p.N, p.n, p.ks_logB, p.ks_l, c->d_ksk, c->d_ks_at, c->func.d_ks, st))
        return fail(TFHE_E_PARAM, "keyswitch: unsupported shape");
    launch_lwe_modswitch(c->func.d_ks, 32 - log2u(fmod), omask, p.n, batch, d_out, st);
    return check_launch("lwe_modswitch");
}

int axpb(tfhe_gate_ctx* c, const uint32_t* x, uint32_t mx, const uint32_t* y, uint32_t my, uint32_t cst,
         uint32_t mo, uint32_t batch, uint32_t* out, hipStream_t st) {
    launch_lwe_axpb(x, mx, y, my, cst, mo, c->p.n, batch, out, st);
    return check_launch("lwe_axpb");
}

// The two tables of EvalFloor at inner modulus q over Z_mod (binfhe-base-scheme.cu:497-523).
int floor_tables(Tables& t, uint64_t q, uint64_t mod, int* s1, int* s2) {
    *s1 = t.add(q, mod, [&](uint64_t x) { return x < q / 2 ? mod - q / 4 : q / 4; });
    *s2 = t.add(q, mod, [&](uint64_t x) {
        return x < q / 4 ? mod - q / 2 - x : x < 3 * (q / 4) ? x : mod + q / 2 - x;
    });
    return (*s1 < 0 || *s2 < 0) ? -1 : 0;
}

// EvalFloor on d_in over Z_mod with inner modulus q: d_out = in + beta - (bootstrap residue);
// in and out may be the same buffer; uses func.d_c1 and func.d_tb.
int floor_run(tfhe_gate_ctx* c, const uint32_t* d_in, uint32_t batch, uint64_t mod, uint64_t q, int s1, int s2,
              uint32_t* d_out, hipStream_t st) {
    const uint32_t mm = mask_of(mod);
    auto& f = c->func;
    int rc = axpb(c, d_in, mm, nullptr, 0, TFHE_FUNC_BETA, mm, batch, f.d_c1, st);
    rc = rc ? rc : bootstrap(c, f.d_c1, batch, q, s1, mod, mm, f.d_tb, st);
    rc = rc ? rc : axpb(c, f.d_c1, mm, f.d_tb, mm, 0, mm, batch, f.d_c1, st);
    rc = rc ? rc : bootstrap(c, f.d_c1, batch, q, s2, mod, mm, f.d_tb, st);
    return rc ? rc : axpb(c, f.d_c1, mm, f.d_tb, mm, 0, mm, batch, d_out, st);
}

int check_func_call(tfhe_gate_ctx* c, const uint32_t* d_ct, uint32_t batch, const void* d_out, const char* what) {
    if (!c) return fail(TFHE_E_PARAM, std::string(what) + ": ctx is NULL");
    if (!c->has_keys) return fail(TFHE_E_STATE, std::string(what) + ": no bootstrapping keys");
    if (!d_ct || !d_out) return fail(TFHE_E_PARAM, std::string(what) + ": NULL ciphertext buffer");
    if (batch > c->max_batch) return fail(TFHE_E_CAPACITY, std::string(what) + ": batch exceeds the context's max_batch");
    return func_prepare(c);
}

// number of floors EvalSign / EvalDecomp take from Z_mod down to at most q_small
int floors_for(uint64_t mod) {
    int k = 0;
    for (uint64_t m = mod; m > TFHE_FUNC_Q; m = m / TFHE_FUNC_Q * 2 * TFHE_FUNC_BETA) ++k;
    return k;
}

}  // namespace

namespace tfhe {
void gate_func_release(tfhe_gate_ctx* c) {
    auto& f = c->func;
    if (f.uploaded) (void)hipEventSynchronize(f.uploaded);
    if (f.h_tv) (void)hipHostFree(f.h_tv);
    if (f.uploaded) (void)hipEventDestroy(f.uploaded);
    (void)hipFree(f.d_tv);
    (void)hipFree(f.d_t);
    (void)hipFree(f.d_u);
    (void)hipFree(f.d_c1);
    (void)hipFree(f.d_tb);
    (void)hipFree(f.d_ks);
    f = GateFuncState{};
}
}  // namespace tfhe

extern "C" int tfhe_func_property(const uint32_t* lut, uint64_t q) {
    if (!lut || !pow2(q) || q > (1ull << 32)) return -1;
    const uint64_t h = q / 2;
    bool neg = true, per = true;
    for (uint64_t x = 0; x < h; ++x) {
        if (((lut[x] + lut[x + h]) & mask_of(q)) != 0) neg = false;
        if (lut[x] != lut[x + h]) per = false;
    }
    return neg ? TFHE_FUNC_NEGACYCLIC : per ? TFHE_FUNC_PERIODIC : TFHE_FUNC_ARBITRARY;
}

extern "C" int tfhe_gate_bootstrap_func_batch(tfhe_gate_ctx* c, const uint32_t* d_ct, uint32_t batch, uint64_t ctmod,
                                              const uint32_t* lut, uint64_t fmod, uint32_t* d_out, void* stream) {
    int rc = check_func_call(c, d_ct, batch, d_out, "bootstrap_func");
    if (rc) return rc;
    if (!lut) return fail(TFHE_E_PARAM, "bootstrap_func: NULL table");
    if (!pow2(ctmod) || ctmod > 2u * (uint64_t)c->p.N || !pow2(fmod) || fmod > (1ull << 32))
        return fail(TFHE_E_PARAM, "bootstrap_func: ctmod must be a power of two <= 2N, fmod a power of two <= 2^32");
    if (batch == 0) return TFHE_OK;
    const hipStream_t st = (hipStream_t)stream;
    Tables t;
    rc = begin_tables(c, t);
    if (rc) return rc;
    const int s = t.add(ctmod, fmod, [&](uint64_t x) { return (uint64_t)lut[x]; });
    rc = upload_tables(c, t, st);
    return rc ? rc : bootstrap(c, d_ct, batch, ctmod, s, fmod, mask_of(fmod), d_out, st);
}

extern "C" int tfhe_gate_eval_func_batch(tfhe_gate_ctx* c, const uint32_t* d_ct, uint32_t batch, uint64_t q,
                                         const uint32_t* lut, uint32_t* d_out, void* stream) {
    int rc = check_func_call(c, d_ct, batch, d_out, "eval_func");
    if (rc) return rc;
    if (!lut) return fail(TFHE_E_PARAM, "eval_func: NULL table");
    if (!pow2(q) || q > 2u * (uint64_t)c->p.N) return fail(TFHE_E_PARAM, "eval_func: q must be a power of two <= 2N");
    const int prop = tfhe_func_property(lut, q);
    if (prop == TFHE_FUNC_ARBITRARY && q > (uint64_t)c->p.N)
        return fail(TFHE_E_PARAM, "eval_func: a table that is neither negacyclic nor periodic needs q <= N");
    if (batch == 0) return TFHE_OK;
    const hipStream_t st = (hipStream_t)stream;
    const uint32_t mq = mask_of(q), beta = TFHE_FUNC_BETA;
    auto& f = c->func;
    Tables t;
    rc = begin_tables(c, t);
    if (rc) return rc;
    if (prop == TFHE_FUNC_NEGACYCLIC) {  // one bootstrap (binfhe-base-scheme.cu:394-400)
        const int s = t.add(q, q, [&](uint64_t x) { return (uint64_t)lut[x]; });
        rc = upload_tables(c, t, st);
        rc = rc ? rc : axpb(c, d_ct, mq, nullptr, 0, beta, mq, batch, f.d_c1, st);
        return rc ? rc : bootstrap(c, f.d_c1, batch, q, s, q, mq, d_out, st);
    }
    if (prop == TFHE_FUNC_ARBITRARY) {  // through the modulus 2q (binfhe-base-scheme.cu:402-456)
        const uint64_t dq = 2 * q;
        const uint32_t md = mask_of(dq);
        const int s0 = t.add(dq, dq, [&](uint64_t x) { return x < q ? dq - dq / 4 : dq / 4; });
        const int s1 = t.add(dq, dq, [&](uint64_t x) { return x < q ? (uint64_t)lut[x] : dq - lut[x - q]; });
        rc = upload_tables(c, t, st);
        rc = rc ? rc : axpb(c, d_ct, mq, nullptr, 0, beta, md, batch, f.d_c1, st);
        rc = rc ? rc : bootstrap(c, f.d_c1, batch, dq, s0, dq, md, f.d_tb, st);
        rc = rc ? rc : axpb(c, d_ct, mq, f.d_tb, md, beta + (uint32_t)(dq - q / 2), md, batch, f.d_c1, st);
        return rc ? rc : bootstrap(c, f.d_c1, batch, dq, s1, dq, mq, d_out, st);
    }
    // periodic: sign bootstrap, then a negacyclic table on the shifted input (binfhe-base-scheme.cu:458-481)
    const int s0 = t.add(q, q, [&](uint64_t x) { return x < q / 2 ? q - q / 4 : q / 4; });
    const int s1 = t.add(q, q, [&](uint64_t x) { return x < q / 2 ? (uint64_t)lut[x] : q - lut[x - q / 2]; });
    rc = upload_tables(c, t, st);
    rc = rc ? rc : axpb(c, d_ct, mq, nullptr, 0, beta, mq, batch, f.d_c1, st);
    rc = rc ? rc : bootstrap(c, f.d_c1, batch, q, s0, q, mq, f.d_tb, st);
    rc = rc ? rc : axpb(c, d_ct, mq, f.d_tb, mq, beta + (uint32_t)(q - q / 4), mq, batch, f.d_c1, st);
    return rc ? rc : bootstrap(c, f.d_c1, batch, q, s1, q, mq, d_out, st);
}

extern "C" int tfhe_gate_eval_floor_batch(tfhe_gate_ctx* c, const uint32_t* d_ct, uint32_t batch, uint64_t mod,
                                          uint32_t roundbits, uint32_t* d_out, void* stream) {
    int rc = check_func_call(c, d_ct, batch, d_out, "eval_floor");
    if (rc) return rc;
    if (roundbits > 30) return fail(TFHE_E_PARAM, "eval_floor: roundbits too large");
    const uint64_t q = roundbits == 0 ? TFHE_FUNC_Q : (uint64_t)TFHE_FUNC_BETA << (roundbits + 1);
    if (!pow2(mod) || mod > (1ull << 32) || q > 2u * (uint64_t)c->p.N || q > mod)
        return fail(TFHE_E_PARAM, "eval_floor: mod must be a power of two <= 2^32 and the inner modulus <= min(2N, mod)");
    if (batch == 0) return TFHE_OK;
    const hipStream_t st = (hipStream_t)stream;
    Tables t;
    rc = begin_tables(c, t);
    if (rc) return rc;
    int s1, s2;
    if (floor_tables(t, q, mod, &s1, &s2)) return fail(TFHE_E_PARAM, "eval_floor: too many test vectors");
    rc = upload_tables(c, t, st);
    return rc ? rc : floor_run(c, d_ct, batch, mod, q, s1, s2, d_out, st);
}

extern "C" int tfhe_gate_eval_sign_batch(tfhe_gate_ctx* c, const uint32_t* d_ct, uint32_t batch, uint64_t mod,
                                         int scheme_switch, uint32_t* d_out, void* stream) {
    int rc = check_func_call(c, d_ct, batch, d_out, "eval_sign");
    if (rc) return rc;
    if (!pow2(mod) || mod > (1ull << 32) || mod <= TFHE_FUNC_Q)
        return fail(TFHE_E_PARAM, "eval_sign: mod must be a power of two in (q_small, 2^32]");
    if (batch == 0) return TFHE_OK;
    const hipStream_t st = (hipStream_t)stream;
    const uint64_t qs = TFHE_FUNC_Q;
    auto& f = c->func;
    Tables t;
    rc = begin_tables(c, t);
    if (rc) return rc;
    std::vector<int> slots;
    uint64_t m = mod;
    while (m > qs) {  // the floors (binfhe-base-scheme.cu:556-590)
        int s1, s2;
        if (floor_tables(t, qs, m, &s1, &s2)) return fail(TFHE_E_PARAM, "eval_sign: too many test vectors");
        slots.push_back(s1);
        slots.push_back(s2);
        m = m * 2 * TFHE_FUNC_BETA / qs;
    }
    const int s3 = t.add(m, qs, [&](uint64_t x) {
        const bool lo = x < m / 2;
        return (lo != (scheme_switch != 0)) ? qs / 4 : qs - qs / 4;
    });
    if (s3 < 0) return fail(TFHE_E_PARAM, "eval_sign: too many test vectors");
    rc = upload_tables(c, t, st);
    rc = rc ? rc : axpb(c, d_ct, mask_of(mod), nullptr, 0, 0, mask_of(mod), batch, f.d_t, st);
    m = mod;
    for (size_t k = 0; !rc && m > qs; k += 2) {
        rc = floor_run(c, f.d_t, batch, m, qs, slots[k], slots[k + 1], f.d_u, st);
        const uint64_t m2 = m * 2 * TFHE_FUNC_BETA / qs;
        if (!rc) {
            launch_lwe_modswitch(f.d_u, log2u(m / m2), mask_of(m2), c->p.n, batch, f.d_t, st);
            rc = check_launch("lwe_modswitch");
        }
        m = m2;
    }
    rc = rc ? rc : axpb(c, f.d_t, mask_of(m), nullptr, 0, TFHE_FUNC_BETA, mask_of(m), batch, f.d_t, st);
    if (scheme_switch) return rc ? rc : bootstrap(c, f.d_t, batch, m, s3, qs, mask_of(qs), d_out, st);
    rc = rc ? rc : bootstrap(c, f.d_t, batch, m, s3, qs, mask_of(qs), f.d_u, st);
    return rc ? rc : axpb(c, f.d_u, mask_of(qs), nullptr, 0, (uint32_t)(qs - qs / 4), mask_of(qs), batch, d_out, st);
}

extern "C" int tfhe_gate_eval_decomp_batch(tfhe_gate_ctx* c, const uint32_t* d_ct, uint32_t batch, uint64_t mod,
                                           uint32_t max_digits, uint32_t* d_out, uint32_t* n_digits,
                                           uint64_t* digit_mods, void* stream) {
    int rc = check_func_call(c, d_ct, batch, d_out, "eval_decomp");
    if (rc) return rc;
    if (!pow2(mod) || mod > (1ull << 32) || mod <= TFHE_FUNC_Q)
        return fail(TFHE_E_PARAM, "eval_decomp: mod must be a power of two in (q_small, 2^32]");
    const int floors = floors_for(mod);
    if ((uint32_t)floors + 1 > max_digits) return fail(TFHE_E_CAPACITY, "eval_decomp: max_digits is too small");
    const uint64_t qs = TFHE_FUNC_Q;
    {
        uint64_t m = mod;
        for (int k = 0; k < floors; ++k, m = m / qs * 2 * TFHE_FUNC_BETA)
            if (digit_mods) digit_mods[k] = qs;
        if (digit_mods) digit_mods[floors] = m;
    }
    if (n_digits) *n_digits = (uint32_t)floors + 1;
    if (batch == 0) return TFHE_OK;
    const hipStream_t st = (hipStream_t)stream;
    auto& f = c->func;
    Tables t;
    rc = begin_tables(c, t);
    if (rc) return rc;
    std::vector<int> slots;
    for (uint64_t m = mod; m > qs; m = m / qs * 2 * TFHE_FUNC_BETA) {
        int s1, s2;
        if (floor_tables(t, qs, m, &s1, &s2)) return fail(TFHE_E_PARAM, "eval_decomp: too many test vectors");
        slots.push_back(s1);
        slots.push_back(s2);
    }
    rc = upload_tables(c, t, st);
    const size_t digit = (size_t)batch * (c->p.n + 1);
    rc = rc ? rc : axpb(c, d_ct, mask_of(mod), nullptr, 0, 0, mask_of(mod), batch, f.d_t, st);
    uint64_t m = mod;
    for (int k = 0; !rc && k < floors; ++k) {  // binfhe-base-scheme.cu:655-690
        rc = axpb(c, f.d_t, mask_of(qs), nullptr, 0, 0, mask_of(qs), batch, d_out + k * digit, st);
        rc = rc ? rc : floor_run(c, f.d_t, batch, m, qs, slots[2 * k], slots[2 * k + 1], f.d_u, st);
        const uint64_t m2 = m / qs * 2 * TFHE_FUNC_BETA;
        if (!rc) {
            launch_lwe_modswitch(f.d_u, log2u(m / m2), mask_of(m2), c->p.n, batch, f.d_t, st);
            rc = check_launch("lwe_modswitch");
        }
        m = m2;
    }
    return rc ? rc : axpb(c, f.d_t, mask_of(m), nullptr, 0, 0, mask_of(m), batch, d_out + floors * digit, st);
}
