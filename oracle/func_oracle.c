/*
 * func_oracle.c -- CPU oracle (test infrastructure only) of the gate level's programmable
 * bootstrap and large-precision functions: the reference's GPUBootstrapFunc, GPUEvalFunc,
 * GPUEvalFloor, GPUEvalSign and GPUEvalDecomp (bitwise/src/binfhe-base-scheme.cu:255-698,
 * after eprint 2021/1337), restated on this build's gate parameters (RLWE and LWE on the
 * 2^32 torus, N = 1024, the gate BSK / KSK) with LWE ciphertexts over a power-of-two modulus
 * `mod` (2 .. 2^32) held as uint32 residues.  The reference's LWE arithmetic comes from the
 * OpenFHE fork (absent): EvalAddConstEq adds to b, EvalSubEq(ct1, ct2) sets ct1 -= ct2,
 * EvalSubEq2(ct1, ct2) sets ct2 = ct1 - ct2, SetModulus(q) reduces every element mod q,
 * ModSwitch(q', ct) rounds every element to round(x q' / q).  One key set: the reference's
 * change of gadget base between keys (EKs.size() == 3) does not arise.
 *
 * Pinned by what the mathematics fixes, not by retyping (the reference ships no vectors for
 * these): the function classes on hand-made tables; BootstrapFunc against its closed form
 * f(phase) or -f(phase + q/2) chosen by <a,s> mod q; EvalFunc on its three branches against the
 * plaintext table; EvalFloor against floor to a multiple of q; EvalSign against the top bit;
 * EvalDecomp by rebuilding the input phase from its digits.  The floor family is exact only
 * while (x + beta) mod q keeps several sigma of the bootstrap output noise (2^-8.1 of the
 * output modulus with the gate keys: BSK noise over 636 steps, then the key switch) away from 0 and q/2; nearer, a floor moves by one level.
 */
#include <stdlib.h>
#include <string.h>

#include "tfhe_oracle.h"

static uint32_t mask_of(uint64_t mod) { return (uint32_t)(mod - 1u); }

/* round(x q' / q) mod q' for powers of two q' <= q (reference RoundqQ: floor(0.5 + x q' / q)) */
static uint32_t round_down_mod(uint32_t x, uint64_t from, uint64_t to) {
    if (to >= from) return x & mask_of(to);
    uint64_t r = from / to;
    return (uint32_t)((((uint64_t)x + r / 2) / r) & mask_of(to));
}

static void lwe_modswitch(int n, const uint32_t* in, uint64_t from, uint64_t to, uint32_t* out) {
    for (int k = 0; k <= n; ++k) out[k] = round_down_mod(in[k], from, to);
}
static void lwe_reduce(int n, uint32_t* ct, uint64_t mod) {
    for (int k = 0; k <= n; ++k) ct[k] &= mask_of(mod);
}
static void lwe_add_const(int n, uint32_t* ct, uint32_t c, uint64_t mod) { ct[n] = (ct[n] + c) & mask_of(mod); }
/* out = x - y (mod), out may alias x or y */
static void lwe_sub(int n, const uint32_t* x, const uint32_t* y, uint64_t mod, uint32_t* out) {
    for (int k = 0; k <= n; ++k) out[k] = (x[k] - y[k]) & mask_of(mod);
}

/* 0: negacyclic (lut[x + q/2] = -lut[x] mod q), 1: periodic (lut[x + q/2] = lut[x]), 2: other */
int orc_func_property(const uint32_t* lut, uint64_t q) {
    const uint64_t h = q / 2;
    int neg = 1, per = 1;
    for (uint64_t x = 0; x < h; ++x) {
        if (((lut[x] + lut[x + h]) & mask_of(q)) != 0) neg = 0;
        if (lut[x] != lut[x + h]) per = 0;
    }
    return neg ? 0 : per ? 1 : 2;
}

/* BootstrapFunc (binfhe-base-scheme.cu:259-378): ct [n+1] over Z_ctmod (ctmod a power of two
 * <= 2N), f [ctmod] values over Z_fmod.  Accumulator (0, m) with m[j 2N / ctmod] = (2^32 / fmod)
 * f((b - j) mod ctmod) for j < ctmod / 2, zero elsewhere; rotations a_i 2N / ctmod (exact);
 * coefficient 0 extracted, switched to dimension n, rounded to Z_fmod. */
int orc_bootstrap_func(const orc_gate_params* p, const double* bsk_f, const uint32_t* ksk, const uint32_t* ct,
                       uint64_t ctmod, const uint32_t* f, uint64_t fmod, uint32_t* out) {
    const int n = p->n, N = p->N;
    if (ctmod < 2 || ctmod > 2u * (uint64_t)N || (ctmod & (ctmod - 1)) || fmod < 2 || fmod > (1ull << 32) ||
        (fmod & (fmod - 1)))
        return -1;
    const uint32_t factor = (uint32_t)(2u * (uint64_t)N / ctmod);
    const uint32_t sc = (uint32_t)((1ull << 32) / fmod);  /* 2^32 / fmod (1 for fmod = 2^32) */
    uint32_t* acc = (uint32_t*)calloc((size_t)2 * N, sizeof(uint32_t));
    uint32_t* at = (uint32_t*)malloc(sizeof(uint32_t) * n);
    uint32_t* lwe = (uint32_t*)malloc(sizeof(uint32_t) * (N + 1));
    uint32_t* ks = (uint32_t*)malloc(sizeof(uint32_t) * (n + 1));
    const uint32_t b = ct[n] & mask_of(ctmod);
    for (uint64_t j = 0; j < ctmod / 2; ++j)
        acc[N + j * factor] = sc * (f[(b - (uint32_t)j) & mask_of(ctmod)] & mask_of(fmod));
    for (int i = 0; i < n; ++i) at[i] = (ct[i] & mask_of(ctmod)) * factor;
    orc_blind_rotate32(p, bsk_f, at, acc);
    orc_sample_extract32(N, acc, 0, lwe);
    orc_keyswitch32(p, ksk, lwe, ks);
    lwe_modswitch(n, ks, 1ull << 32, fmod, out);
    free(acc); free(at); free(lwe); free(ks);
    return 0;
}

/* EvalFunc (binfhe-base-scheme.cu:380-481): ct over Z_q, lut [q] over Z_q; out over Z_q. */
int orc_eval_func(const orc_gate_params* p, const double* bsk_f, const uint32_t* ksk, const uint32_t* ct,
                  uint64_t q, const uint32_t* lut, uint32_t beta, uint32_t* out) {
    const int n = p->n, N = p->N;
    const size_t w = (size_t)(n + 1);
    const int prop = orc_func_property(lut, q);
    uint32_t* ct1 = (uint32_t*)malloc(sizeof(uint32_t) * w);
    uint32_t* ct2 = (uint32_t*)malloc(sizeof(uint32_t) * w);
    uint32_t* ct3 = (uint32_t*)malloc(sizeof(uint32_t) * w);
    int rc = 0;
    memcpy(ct1, ct, sizeof(uint32_t) * w);
    lwe_reduce(n, ct1, q);
    if (prop == 0) {  /* negacyclic: one bootstrap */
        lwe_add_const(n, ct1, beta, q);
        rc = orc_bootstrap_func(p, bsk_f, ksk, ct1, q, lut, q, out);
    } else if (prop == 2) {  /* arbitrary: q <= N, through the modulus 2q */
        if (q > (uint64_t)N) {
            rc = -1;
        } else {
            const uint64_t dq = 2 * q;
            uint32_t* f0 = (uint32_t*)malloc(sizeof(uint32_t) * dq);
            uint32_t* fl = (uint32_t*)malloc(sizeof(uint32_t) * dq);
            for (uint64_t x = 0; x < dq; ++x) {
                f0[x] = x < q ? (uint32_t)(dq - dq / 4) : (uint32_t)(dq / 4);
                /* LUT2 = LUT twice; x < q: LUT2[x], else dq - LUT2[x - q] */
                fl[x] = x < q ? lut[x] : (uint32_t)((dq - lut[(x - q) % q]) & mask_of(dq));
            }
            memcpy(ct2, ct1, sizeof(uint32_t) * w);  /* ct1 read mod 2q */
            lwe_add_const(n, ct2, beta, dq);
            rc = orc_bootstrap_func(p, bsk_f, ksk, ct2, dq, f0, dq, ct3);
            lwe_sub(n, ct1, ct3, dq, ct3);           /* ct3 = ct1 - ct3 */
            lwe_add_const(n, ct3, beta, dq);
            lwe_add_const(n, ct3, (uint32_t)(dq - q / 2), dq);
            if (!rc) rc = orc_bootstrap_func(p, bsk_f, ksk, ct3, dq, fl, dq, out);
            lwe_reduce(n, out, q);
            free(f0); free(fl);
        }
    } else {  /* periodic */
        uint32_t* f0 = (uint32_t*)malloc(sizeof(uint32_t) * q);
        uint32_t* fl = (uint32_t*)malloc(sizeof(uint32_t) * q);
        for (uint64_t x = 0; x < q; ++x) {
            f0[x] = x < q / 2 ? (uint32_t)(q - q / 4) : (uint32_t)(q / 4);
            fl[x] = x < q / 2 ? lut[x] : (uint32_t)((q - lut[x - q / 2]) & mask_of(q));
        }
        lwe_add_const(n, ct1, beta, q);
        rc = orc_bootstrap_func(p, bsk_f, ksk, ct1, q, f0, q, ct2);
        memcpy(ct3, ct, sizeof(uint32_t) * w);
        lwe_reduce(n, ct3, q);
        lwe_sub(n, ct3, ct2, q, ct2);              /* ct2 = ct - ct2 */
        lwe_add_const(n, ct2, beta, q);
        lwe_add_const(n, ct2, (uint32_t)(q - q / 4), q);
        if (!rc) rc = orc_bootstrap_func(p, bsk_f, ksk, ct2, q, fl, q, out);
        free(f0); free(fl);
    }
    free(ct1); free(ct2); free(ct3);
    return rc;
}

/* EvalFloor (binfhe-base-scheme.cu:484-531): ct over Z_mod; q = qsmall (roundbits = 0) or
 * beta 2^(roundbits + 1); out over Z_mod. */
int orc_eval_floor(const orc_gate_params* p, const double* bsk_f, const uint32_t* ksk, const uint32_t* ct,
                   uint64_t mod, uint32_t beta, uint32_t roundbits, uint64_t qsmall, uint32_t* out) {
    const int n = p->n;
    const size_t w = (size_t)(n + 1);
    const uint64_t q = roundbits == 0 ? qsmall : (uint64_t)beta << (roundbits + 1);
    uint32_t* ct1 = (uint32_t*)malloc(sizeof(uint32_t) * w);
    uint32_t* tq = (uint32_t*)malloc(sizeof(uint32_t) * w);
    uint32_t* tb = (uint32_t*)malloc(sizeof(uint32_t) * w);
    uint32_t* f = (uint32_t*)malloc(sizeof(uint32_t) * q);
    memcpy(ct1, ct, sizeof(uint32_t) * w);
    lwe_reduce(n, ct1, mod);
    lwe_add_const(n, ct1, beta, mod);
    memcpy(tq, ct1, sizeof(uint32_t) * w);
    lwe_reduce(n, tq, q);
    for (uint64_t x = 0; x < q; ++x) f[x] = x < q / 2 ? (uint32_t)(mod - q / 4) : (uint32_t)(q / 4);
    int rc = orc_bootstrap_func(p, bsk_f, ksk, tq, q, f, mod, tb);
    lwe_sub(n, ct1, tb, mod, ct1);
    memcpy(tq, ct1, sizeof(uint32_t) * w);
    lwe_reduce(n, tq, q);
    for (uint64_t x = 0; x < q; ++x)
        f[x] = x < q / 4 ? (uint32_t)(mod - q / 2 - x) : x < 3 * (q / 4) ? (uint32_t)x : (uint32_t)(mod + q / 2 - x);
    if (!rc) rc = orc_bootstrap_func(p, bsk_f, ksk, tq, q, f, mod, tb);
    lwe_sub(n, ct1, tb, mod, out);
    free(ct1); free(tq); free(tb); free(f);
    return rc;
}

/* EvalSign (binfhe-base-scheme.cu:533-608) with one key set: ct over Z_mod (mod > qsmall);
 * out over Z_qsmall encrypting 0 (MSB 0) or qsmall / 2 (MSB 1); scheme_switch = 1 returns
 * -q/4 / +q/4 without the final shift. */
int orc_eval_sign(const orc_gate_params* p, const double* bsk_f, const uint32_t* ksk, const uint32_t* ct,
                  uint64_t mod, uint32_t beta, uint64_t qsmall, int scheme_switch, uint32_t* out) {
    const int n = p->n;
    const size_t w = (size_t)(n + 1);
    if (mod <= qsmall) return -1;
    uint32_t* t = (uint32_t*)malloc(sizeof(uint32_t) * w);
    uint32_t* u = (uint32_t*)malloc(sizeof(uint32_t) * w);
    memcpy(t, ct, sizeof(uint32_t) * w);
    lwe_reduce(n, t, mod);
    uint64_t m = mod;
    int rc = 0;
    while (m > qsmall && !rc) {
        rc = orc_eval_floor(p, bsk_f, ksk, t, m, beta, 0, qsmall, u);
        const uint64_t m2 = (m << 1) * beta / qsmall;
        lwe_modswitch(n, u, m, m2, t);
        m = m2;
    }
    lwe_add_const(n, t, beta, m);
    uint32_t* f = (uint32_t*)malloc(sizeof(uint32_t) * m);
    for (uint64_t x = 0; x < m; ++x) {
        const int lo = x < m / 2;
        f[x] = (uint32_t)((lo ^ scheme_switch) ? qsmall / 4 : qsmall - qsmall / 4);
    }
    if (!rc) rc = orc_bootstrap_func(p, bsk_f, ksk, t, m, f, qsmall, out);
    if (!scheme_switch) lwe_add_const(n, out, (uint32_t)(qsmall - qsmall / 4), qsmall);
    free(t); free(u); free(f);
    return rc;
}

/* EvalDecomp (binfhe-base-scheme.cu:633-698) with one key set: ct over Z_mod (mod > qsmall);
 * writes the digits [count][n+1] (each over Z_qsmall except the last, over its own modulus,
 * <= qsmall) and returns the count (or -1). */
int orc_eval_decomp(const orc_gate_params* p, const double* bsk_f, const uint32_t* ksk, const uint32_t* ct,
                    uint64_t mod, uint32_t beta, uint64_t qsmall, uint32_t* out, int max_digits) {
    const int n = p->n;
    const size_t w = (size_t)(n + 1);
    if (mod <= qsmall) return -1;
    uint32_t* t = (uint32_t*)malloc(sizeof(uint32_t) * w);
    uint32_t* u = (uint32_t*)malloc(sizeof(uint32_t) * w);
    memcpy(t, ct, sizeof(uint32_t) * w);
    lwe_reduce(n, t, mod);
    uint64_t m = mod;
    int count = 0, rc = 0;
    while (m > qsmall && !rc) {
        if (count >= max_digits) { rc = -1; break; }
        memcpy(out + (size_t)count * w, t, sizeof(uint32_t) * w);
        lwe_reduce(n, out + (size_t)count * w, qsmall);
        ++count;
        rc = orc_eval_floor(p, bsk_f, ksk, t, m, beta, 0, qsmall, u);
        const uint64_t m2 = m / qsmall * 2 * beta;
        lwe_modswitch(n, u, m, m2, t);
        m = m2;
    }
    if (!rc && count < max_digits) {
        memcpy(out + (size_t)count * w, t, sizeof(uint32_t) * w);
        ++count;
    } else {
        rc = -1;
    }
    free(t); free(u);
    return rc ? -1 : count;
}
