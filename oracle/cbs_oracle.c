/*
 * cbs_oracle.c -- CPU restatement of the circuit-bootstrapping level (N = 2048):
 * multi-value programmable bootstrap (torus-64), private functional key
 * switching to torus-32 RLWE, RGSW assembly, CMux / vertical-packing LUT
 * (torus-32) and the key switch back to the level-0 LWE key.
 *
 * TEST INFRASTRUCTURE ONLY (see tfhe_oracle.h).
 *
 * Exactness: every negacyclic product of a signed digit polynomial (|d| <= 2^8)
 * with a torus-64 polynomial is computed as two FP64 FFT products, one per
 * signed 32-bit limb of the torus polynomial, rounded to integers and
 * recombined mod 2^64 (the limb scheme of the reference's split-FFT backend,
 * bitwise/src/cirbts/cirbts-init.cu:1096-1120, cirbts-fft.cu:159-226,345-355,
 * with 32-bit limbs of a 64-bit torus instead of limbs of a 54-bit prime).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "tfhe_oracle.h"

/* ------------------------------------------------------------------------- */
/* limb helpers                                                               */
/* ------------------------------------------------------------------------- */
/* v = hi * 2^32 + lo with lo, hi signed 32-bit */
static inline void split64(uint64_t v, int32_t* lo, int32_t* hi) {
    int32_t l = (int32_t)(uint32_t)v;
    *lo = l;
    *hi = (int32_t)(uint32_t)((v - (uint64_t)(int64_t)l) >> 32);
}

static inline uint64_t rnd64(double x, double* maxerr) {
    double r = nearbyint(x);
    double e = fabs(x - r);
    if (e > *maxerr) *maxerr = e;
    return (uint64_t)(int64_t)r;
}

/* Fourier form of one torus-64 polynomial: [limb][re M | im M] (4*M doubles). */
void orc_to_fourier64(const uint64_t* poly, int N, double* out) {
    int M = N / 2;
    double* x0 = (double*)malloc(sizeof(double) * N);
    double* x1 = (double*)malloc(sizeof(double) * N);
    for (int c = 0; c < N; ++c) {
        int32_t lo, hi;
        split64(poly[c], &lo, &hi);
        x0[c] = lo;
        x1[c] = hi;
    }
    orc_fft_forward(x0, out, out + M, N);
    orc_fft_forward(x1, out + 2 * M, out + 3 * M, N);
    free(x0);
    free(x1);
}

/* exact a*K mod 2^64 for binary K: two limb products through the FFT
 * (|limb| * N/2 <= 2^41, far inside the FP64 exactness range) */
static void mul_binary64(const int32_t* K, const uint64_t* a, uint64_t* out, int N) {
    int64_t* x = (int64_t*)malloc(sizeof(int64_t) * N);
    int64_t* l0 = (int64_t*)malloc(sizeof(int64_t) * N);
    int64_t* l1 = (int64_t*)malloc(sizeof(int64_t) * N);
    int64_t* p0 = (int64_t*)malloc(sizeof(int64_t) * N);
    int64_t* p1 = (int64_t*)malloc(sizeof(int64_t) * N);
    for (int c = 0; c < N; ++c) {
        int32_t lo, hi;
        split64(a[c], &lo, &hi);
        l0[c] = lo;
        l1[c] = hi;
        x[c] = K[c];
    }
    orc_nc_mul_fft(x, l0, p0, N);
    orc_nc_mul_fft(x, l1, p1, N);
    for (int c = 0; c < N; ++c) out[c] = (uint64_t)p0[c] + ((uint64_t)p1[c] << 32);
    free(x); free(l0); free(l1); free(p0); free(p1);
}

/* Fresh RGSW_K(m) with gadget 2^(64 - l logB + j logB), rows r = 2j + u ([2l][2][N]);
 * ids id0 .. id0 + 2l - 1 of the sampler domain dom. */
void orc_rgsw_encrypt64(uint64_t seed, uint32_t dom, uint32_t id0, int N, const int32_t* K, int m, int logB,
                        int l, int nshift, uint64_t* rgsw) {
    uint64_t* msg = (uint64_t*)calloc((size_t)N, sizeof(uint64_t));
    for (int r = 0; r < 2 * l; ++r) {
        int j = r >> 1, u = r & 1;
        memset(msg, 0, sizeof(uint64_t) * (size_t)N);
        msg[0] = (uint64_t)m << (64 - l * logB + j * logB);
        orc_rlwe_encrypt64(seed, dom, id0 + (uint32_t)r, N, K, u == 0 ? msg : NULL, u == 1 ? msg : NULL, nshift,
                           rgsw + (size_t)r * 2 * N);
    }
    free(msg);
}

/* ct = (a + msg_a, a*K + e + msg_b) over Z/2^64; msg pointers may be NULL. */
void orc_rlwe_encrypt64(uint64_t seed, uint32_t dom, uint32_t id, int N, const int32_t* K,
                        const uint64_t* msg_a, const uint64_t* msg_b, int nshift, uint64_t* ct) {
    uint64_t* a = ct;
    uint64_t* b = ct + N;
    for (int j = 0; j < N; ++j) a[j] = orc_rand64(seed, dom, id, (uint32_t)j);
    mul_binary64(K, a, b, N);
    for (int j = 0; j < N; ++j) {
        b[j] += (uint64_t)orc_noise(seed, dom, id, (uint32_t)j, nshift);
        if (msg_b) b[j] += msg_b[j];
        if (msg_a) a[j] += msg_a[j];
    }
}

/* BSK2 [n][2l][2][N] u64: row r = 2j + u adds s_i * 2^(64 - l logB + j logB) to component u. */
void orc_gen_bsk64(uint64_t seed, const orc_cbs_params* p, const int32_t* s, const int32_t* K2,
                   uint64_t* bsk) {
    int N = p->N, rows = 2 * p->bk_l;
    uint64_t* msg = (uint64_t*)calloc((size_t)N, sizeof(uint64_t));
    for (int i = 0; i < p->n; ++i)
        for (int r = 0; r < rows; ++r) {
            int j = r >> 1, u = r & 1;
            uint64_t g = (uint64_t)1 << (64 - p->bk_l * p->bk_logB + j * p->bk_logB);
            memset(msg, 0, sizeof(uint64_t) * (size_t)N);
            msg[0] = (uint64_t)s[i] * g;
            uint32_t id = (uint32_t)(i * rows + r);
            orc_rlwe_encrypt64(seed, ORC_DOM_BSK2, id, N, K2, u == 0 ? msg : NULL, u == 1 ? msg : NULL,
                               p->rlwe_nshift, bsk + (size_t)id * 2 * N);
        }
    free(msg);
}

/* Fourier BSK2: [n][2l][2 out][4M] (two limbs per polynomial). */
void orc_bsk64_to_fourier(const orc_cbs_params* p, const uint64_t* bsk, double* bsk_f) {
    int N = p->N, M = N / 2;
    size_t npoly = (size_t)p->n * 2 * p->bk_l * 2;
    for (size_t q = 0; q < npoly; ++q) orc_to_fourier64(bsk + q * N, N, bsk_f + q * 4 * M);
}

/* Forward transforms of the 2l signed digit polynomials of acc = (a, b): dr, di
 * [2l][M]; digit poly r = 2j + u holds digit j of component u (the row order of
 * the reference's decomposition, kernel.cu:30-49). */
static void ep64_digit_transforms(int N, int logB, int l, const uint64_t* acc, double* dr, double* di) {
    int M = N / 2, rows = 2 * l;
    double* x = (double*)malloc(sizeof(double) * N);
    int32_t* dig = (int32_t*)malloc(sizeof(int32_t) * (size_t)rows * N);
    int32_t tmp[64];
    for (int u = 0; u < 2; ++u)
        for (int c = 0; c < N; ++c) {
            orc_decompose64(acc[u * N + c], logB, l, tmp);
            for (int j = 0; j < l; ++j) dig[(size_t)(2 * j + u) * N + c] = tmp[j];
        }
    for (int r = 0; r < rows; ++r) {
        for (int c = 0; c < N; ++c) x[c] = (double)dig[(size_t)r * N + c];
        orc_fft_forward(x, dr + (size_t)r * M, di + (size_t)r * M, N);
    }
    free(x);
    free(dig);
}

/* t = sum_r D_r C_r from the digit transforms, two limbs per key polynomial,
 * each limb product rounded to integers and recombined mod 2^64 (exact).
 * rgsw_f: [2l rows][2 out][4M]. */
static double ep64_mac_inverse(int N, int l, const double* dr, const double* di, const double* rgsw_f,
                               uint64_t* out) {
    int M = N / 2, rows = 2 * l;
    double* sr = (double*)calloc((size_t)4 * M, sizeof(double)); /* [o][limb][M] */
    double* si = (double*)calloc((size_t)4 * M, sizeof(double));
    for (int r = 0; r < rows; ++r) {
        const double* xr = dr + (size_t)r * M;
        const double* xi = di + (size_t)r * M;
        for (int o = 0; o < 2; ++o)
            for (int lb = 0; lb < 2; ++lb) {
                const double* kr = rgsw_f + ((size_t)r * 2 + o) * 4 * M + (size_t)lb * 2 * M;
                const double* ki = kr + M;
                double* ar = sr + (size_t)(o * 2 + lb) * M;
                double* ai = si + (size_t)(o * 2 + lb) * M;
                for (int k = 0; k < M; ++k) {
                    ar[k] += xr[k] * kr[k] - xi[k] * ki[k];
                    ai[k] += xr[k] * ki[k] + xi[k] * kr[k];
                }
            }
    }
    double maxerr = 0.0;
    double* x = (double*)malloc(sizeof(double) * N);
    double* y = (double*)malloc(sizeof(double) * N);
    for (int o = 0; o < 2; ++o) {
        orc_fft_inverse(sr + (size_t)(o * 2) * M, si + (size_t)(o * 2) * M, x, N);
        orc_fft_inverse(sr + (size_t)(o * 2 + 1) * M, si + (size_t)(o * 2 + 1) * M, y, N);
        for (int c = 0; c < N; ++c) {
            uint64_t p0 = rnd64(x[c], &maxerr);
            uint64_t p1 = rnd64(y[c], &maxerr);
            out[o * N + c] = p0 + (p1 << 32);
        }
    }
    free(x); free(y); free(sr); free(si);
    return maxerr;
}

/* t = sum_r dec_r(acc) (*) C_r over Z/2^64, exact through two limb products.
 * rgsw_f: [2l rows][2 out][4M]; digit poly r = 2j + u holds digit j of component u. */
double orc_external_product64(int N, int logB, int l, const double* rgsw_f, const uint64_t* acc,
                              uint64_t* out) {
    int M = N / 2, rows = 2 * l;
    double* dr = (double*)malloc(sizeof(double) * (size_t)rows * M);
    double* di = (double*)malloc(sizeof(double) * (size_t)rows * M);
    ep64_digit_transforms(N, logB, l, acc, dr, di);
    double e = ep64_mac_inverse(N, l, dr, di, rgsw_f, out);
    free(dr);
    free(di);
    return e;
}

/* acc += X^m t - t on both components (reference kernels_fft.cu:90-113, over Z/2^64). */
void orc_monic_update64(int N, uint32_t m, const uint64_t* t, uint64_t* acc) {
    int k = (int)(m & (uint32_t)(N - 1));
    int hi = m >= (uint32_t)N;
    for (int o = 0; o < 2; ++o)
        for (int c = 0; c < N; ++c) {
            int idx = c - k;
            int wrapped = idx < 0;
            if (wrapped) idx += N;
            uint64_t v = t[o * N + idx];
            if (hi ^ wrapped) v = 0 - v;
            acc[o * N + c] += v - t[o * N + c];
        }
}

double orc_blind_rotate64(const orc_cbs_params* p, const double* bsk_f, const uint32_t* a_tilde,
                          uint64_t* acc) {
    int N = p->N, M = N / 2, rows = 2 * p->bk_l;
    uint64_t* t = (uint64_t*)malloc(sizeof(uint64_t) * 2 * N);
    double maxerr = 0.0;
    for (int i = 0; i < p->n; ++i) {
        uint32_t m = (2u * (uint32_t)N - a_tilde[i]) & (2u * (uint32_t)N - 1u);
        if (m == 0) continue;
        const double* key = bsk_f + (size_t)i * rows * 2 * 4 * M;
        double e = orc_external_product64(N, p->bk_logB, p->bk_l, key, acc, t);
        if (e > maxerr) maxerr = e;
        orc_monic_update64(N, m, t, acc);
    }
    free(t);
    return maxerr;
}

/* ------------------------------------------------------------------------- */
/* Two key bits per blind-rotation step (the reference's multi-bit PBS,        */
/* CIRBTS_PBS_MULTIBIT, main.cu:130; combination keys cirbts-init.cu:477-583;  */
/* step kernel.cu:4585-4771, restated for two bits and the FFT domain).        */
/* For the key-bit pair (s_2k, s_2k+1) three pattern keys RGSW(ind_p) are      */
/* generated, p = 0, 1, 2 for the patterns (1,0), (0,1), (1,1); exactly one    */
/* indicator is 1 (or none), so with m1, m2 the pair's rotations               */
/*   acc += sum_p (X^(e_p) - 1) (dec(acc) (*) RGSW(ind_p)),                    */
/*   e_p = m1, m2, m1 + m2,                                                    */
/* rotates acc by X^(m1 s_2k + m2 s_2k+1) as two single-bit steps would.  One   */
/* decomposition of acc serves all three products.                              */
/* ------------------------------------------------------------------------- */
uint32_t orc_mb2_indicator(const int32_t* s, int k, int pat) {
    uint32_t b0 = (uint32_t)s[2 * k], b1 = (uint32_t)s[2 * k + 1];
    return pat == 0 ? (b0 & (1u - b1)) : pat == 1 ? ((1u - b0) & b1) : (b0 & b1);
}

/* one row of the pattern key: RGSW q = 3 k + pat, row r = 2j + u, RLWE id q 2l + r of
 * sampler domain ORC_DOM_BSK2MB, message ind_pat * 2^(64 - l logB + j logB) on
 * component u ([2][N]) */
void orc_gen_bsk64_mb2_row(uint64_t seed, const orc_cbs_params* p, const int32_t* s, const int32_t* K2,
                           uint32_t q, int r, uint64_t* row) {
    int N = p->N, rows = 2 * p->bk_l, j = r >> 1, u = r & 1;
    uint64_t g = (uint64_t)1 << (64 - p->bk_l * p->bk_logB + j * p->bk_logB);
    uint64_t* msg = (uint64_t*)calloc((size_t)N, sizeof(uint64_t));
    msg[0] = (uint64_t)orc_mb2_indicator(s, (int)(q / 3), (int)(q % 3)) * g;
    orc_rlwe_encrypt64(seed, ORC_DOM_BSK2MB, q * (uint32_t)rows + (uint32_t)r, N, K2, u == 0 ? msg : NULL,
                       u == 1 ? msg : NULL, p->rlwe_nshift, row);
    free(msg);
}

/* the whole pattern key [n/2][3][2l][2][N] (n even) */
void orc_gen_bsk64_mb2(uint64_t seed, const orc_cbs_params* p, const int32_t* s, const int32_t* K2,
                       uint64_t* bsk) {
    int N = p->N, rows = 2 * p->bk_l;
    long total = (long)(p->n / 2) * 3 * rows;
#pragma omp parallel for schedule(dynamic, 8)
    for (long id = 0; id < total; ++id)
        orc_gen_bsk64_mb2_row(seed, p, s, K2, (uint32_t)(id / rows), (int)(id % rows), bsk + (size_t)id * 2 * N);
}

/* Fourier form of nrgsw RGSWs of 2l rows ([nrgsw][2l][2 out][4M], two limbs) */
void orc_rgsw64_to_fourier(const orc_cbs_params* p, long nrgsw, const uint64_t* bsk, double* bsk_f) {
    int N = p->N, M = N / 2;
    long npoly = nrgsw * 2 * p->bk_l * 2;
#pragma omp parallel for schedule(static)
    for (long q = 0; q < npoly; ++q) orc_to_fourier64(bsk + (size_t)q * N, N, bsk_f + (size_t)q * 4 * M);
}

/* exact two-bit blind rotation with the pattern key (Fourier, [n/2][3][2l][2][4M]) */
double orc_blind_rotate64_mb2(const orc_cbs_params* p, const double* bsk_f, const uint32_t* a_tilde,
                              uint64_t* acc) {
    int N = p->N, M = N / 2, rows = 2 * p->bk_l;
    uint32_t Q2 = 2u * (uint32_t)N;
    double* dr = (double*)malloc(sizeof(double) * (size_t)rows * M);
    double* di = (double*)malloc(sizeof(double) * (size_t)rows * M);
    uint64_t* t = (uint64_t*)malloc(sizeof(uint64_t) * 3 * 2 * N);
    double maxerr = 0.0;
    for (int k = 0; k < p->n / 2; ++k) {
        uint32_t m1 = (Q2 - a_tilde[2 * k]) & (Q2 - 1), m2 = (Q2 - a_tilde[2 * k + 1]) & (Q2 - 1);
        if ((m1 | m2) == 0) continue;
        uint32_t e[3] = {m1, m2, (m1 + m2) & (Q2 - 1)};
        ep64_digit_transforms(N, p->bk_logB, p->bk_l, acc, dr, di);
        for (int pat = 0; pat < 3; ++pat) {
            const double* key = bsk_f + ((size_t)k * 3 + pat) * rows * 2 * 4 * M;
            double err = ep64_mac_inverse(N, p->bk_l, dr, di, key, t + (size_t)pat * 2 * N);
            if (err > maxerr) maxerr = err;
        }
        /* all three products come from the accumulator before the step */
        for (int pat = 0; pat < 3; ++pat)
            if (e[pat]) orc_monic_update64(N, e[pat], t + (size_t)pat * 2 * N, acc);
    }
    free(dr); free(di); free(t);
    return maxerr;
}

/* LWE of coefficient idx over Z/2^64 (extract.cuh:22-31). */
void orc_sample_extract64(int N, const uint64_t* acc, uint32_t idx, uint64_t* lwe) {
    for (uint32_t j = 0; j < (uint32_t)N; ++j) lwe[j] = j <= idx ? acc[idx - j] : 0 - acc[(uint32_t)N + idx - j];
    lwe[N] = acc[N + idx];
}

/* round(v * 2N / (2^32 * 2^bw)) * 2^bw mod 2N: the multi-value modulus switch
 * (reference gpu_SpecilMS, bitwise/src/cirbts/cirbts-utils.cu:52-61, restated in
 * exact integer arithmetic, ties up). */
uint32_t orc_special_ms(uint32_t v, int N, int bw) {
    int lg = 0;
    while ((1 << lg) < 2 * N) ++lg;
    int sh = 32 - lg + bw;
    uint64_t r = ((uint64_t)v + ((uint64_t)1 << (sh - 1))) >> sh;
    return (uint32_t)((r << bw) & (uint64_t)(2 * N - 1));
}

/* Circuit-bootstrapping PBS with both CB levels from one blind rotation
 * (the reference's ManyLUT, cirbts-bootstrap.cu:753-839 + kernel_GenMVRLWEs):
 * input LWE_s (torus-32) of b * 2^31; b += 2^30; test vector TV[c] = -mu_{c mod 2}
 * with mu_j = 2^(64 - cb_l cb_logB + j cb_logB - 1); outputs out[j] = LWE_K2 of
 * b * 2^(64 - cb_l cb_logB + j cb_logB) (coefficient j plus mu_j), j = 0, 1. */
double orc_pbs_manylut(const orc_cbs_params* p, const double* bsk_f, const uint32_t* ct32, uint64_t* out) {
    return orc_pbs_manylut_mb(p, bsk_f, 1, ct32, out);
}

/* mb = 1: one key bit per step with the BSK2 of orc_gen_bsk64; mb = 2: two bits per
 * step with the pattern key of orc_gen_bsk64_mb2 (n even) */
double orc_pbs_manylut_mb(const orc_cbs_params* p, const double* bsk_f, int mb, const uint32_t* ct32,
                          uint64_t* out) {
    int n = p->n, N = p->N;
    uint32_t* at = (uint32_t*)malloc(sizeof(uint32_t) * n);
    uint64_t* acc = (uint64_t*)calloc((size_t)2 * N, sizeof(uint64_t));
    for (int i = 0; i < n; ++i) at[i] = orc_special_ms(ct32[i], N, 1);
    uint32_t bt = orc_special_ms(ct32[n] + (1u << 30), N, 1);
    uint64_t mu[2];
    for (int j = 0; j < 2; ++j) mu[j] = (uint64_t)1 << (64 - p->cb_l * p->cb_logB + j * p->cb_logB - 1);
    /* b-part: the rotation X^{-<a~,s>} of the blind rotation selects coefficient
     * <a~,s> of the test vector, so b~ is folded in as in the gate test vector
     * (binfhe-base-scheme.cu:120-131): TV[k] = F_{k mod 2}(b~ - (k - k mod 2))
     * with F_j(x) = -mu_j for x mod 2N in [0, N) and +mu_j otherwise. */
    for (int k = 0; k < N; ++k) {
        uint32_t arg = (bt + 2u * (uint32_t)N - (uint32_t)(k & ~1)) & (2u * (uint32_t)N - 1u);
        uint64_t v = 0 - mu[k & 1];
        if (arg >= (uint32_t)N) v = 0 - v;
        acc[N + k] = v;
    }
    double err = mb == 2 ? orc_blind_rotate64_mb2(p, bsk_f, at, acc) : orc_blind_rotate64(p, bsk_f, at, acc);
    for (int j = 0; j < 2; ++j) {
        uint64_t* lwe = out + (size_t)j * (N + 1);
        orc_sample_extract64(N, acc, (uint32_t)j, lwe);
        lwe[N] += mu[j];
    }
    free(at);
    free(acc);
    return err;
}

/* orc_blind_rotate_lut64: the accumulator part, without the offsets.
 * orc_bootstrap_many_lut: the reference's gpu_BootstrapManyLUT (cirbts-bootstrap.cu:753-839) with an
 * arbitrary test polynomial: a~_i, b~ = gpu_SpecilMS with bitwidth bw
 * (cirbts-utils.cu:52-61, no offset on b), acc = (0, LUT X^{-b~}) (the memcpy of LUT
 * and kernel_ModMulScalar by the monomial X^{-b_MS}), the blind rotation (mb key bits
 * per step) by X^{+a~_i s_i}, so that acc ends as LUT X^{-phase~} with
 * phase~ = b~ - <a~, s> mod 2N (orc_blind_rotate64 multiplies by X^{-a_tilde_i}, so
 * it is given -a~_i, the negation the reference applies in its LMKCDEY branch),
 * then Gpow_j / 2 added to body coefficient j < cb_l (kernel_ModAddGpowScaled,
 * kernel.cu:7459-7473; Gpow_j = 2^(64 - cb_l cb_logB + j cb_logB) here).
 * out: RLWE [2][N] (mask, body), 64-bit torus, coefficient domain; coefficient c of
 * its phase is LUT[c + phase~] (negated past N) plus the offsets and the noise. */
double orc_blind_rotate_lut64(const orc_cbs_params* p, const double* bsk_f, int mb, const uint32_t* ct32,
                              const uint64_t* lut, int bw, uint64_t* out) {
    int n = p->n, N = p->N;
    uint32_t Q2 = 2u * (uint32_t)N;
    uint32_t* at = (uint32_t*)malloc(sizeof(uint32_t) * n);
    for (int i = 0; i < n; ++i) at[i] = (Q2 - orc_special_ms(ct32[i], N, bw)) & (Q2 - 1);
    uint32_t bt = orc_special_ms(ct32[n], N, bw);
    for (int c = 0; c < N; ++c) {
        /* coefficient c of LUT X^{-bt}: LUT[m], m = c + bt mod 2N, negated when m >= N */
        uint32_t m = ((uint32_t)c + bt) & (Q2 - 1);
        out[c] = 0;
        out[N + c] = m < (uint32_t)N ? lut[m] : 0 - lut[m - (uint32_t)N];
    }
    double err = mb == 2 ? orc_blind_rotate64_mb2(p, bsk_f, at, out) : orc_blind_rotate64(p, bsk_f, at, out);
    free(at);
    return err;
}

double orc_bootstrap_many_lut(const orc_cbs_params* p, const double* bsk_f, int mb, const uint32_t* ct32,
                              const uint64_t* lut, int bw, uint64_t* out) {
    int N = p->N;
    double err = orc_blind_rotate_lut64(p, bsk_f, mb, ct32, lut, bw, out);
    for (int j = 0; j < p->cb_l && j < N; ++j)
        out[N + j] += (uint64_t)1 << (64 - p->cb_l * p->cb_logB + j * p->cb_logB - 1);
    return err;
}

/* ------------------------------------------------------------------------- */
/* Private functional key switching LWE_K2 (dim N, torus-64) -> RLWE_K2        */
/* (torus-32) (CGGI 2020, PrivKS):                                             */
/*   out = - sum_{i<=N} sum_t KS[u][i][t][digit_t(a~_i)]  (mod 2^32)           */
/* with a_N = b and kappa_N = -1, a~_i = round(a_i / 2^(64 - pf_l pf_logB)),   */
/* digits unsigned, least significant first, value 0 contributing nothing.     */
/* KS[u][i][t][v] = RLWE_K2(f_u(v * kappa_i * 2^(32 - pf_l pf_logB + t pf_logB)))*/
/* on the 32-bit torus (sigma 2^-pf_nshift), f_0(x) = -K2(X) * x, f_1(x) = x.  */
/* The output RLWE, the RGSW rows built from it and the CMux LUT all live on  */
/* the 32-bit torus; pf_l pf_logB = 32 keeps every bit the key can resolve.   */
/* ------------------------------------------------------------------------- */
uint32_t orc_pfks_id(const orc_cbs_params* p, int u, int i, int t, int v) {
    int B = 1 << p->pf_logB;
    return (uint32_t)((((size_t)u * (p->N + 1) + i) * p->pf_l + t) * B + v);
}

void orc_gen_pfks_entry(uint64_t seed, const orc_cbs_params* p, const int32_t* K2, int u, int i, int t,
                        int v, uint32_t* ct) {
    int N = p->N;
    uint32_t kappa = i < N ? (uint32_t)K2[i] : 0xFFFFFFFFu;
    uint32_t c = (uint32_t)v * kappa * ((uint32_t)1 << (32 - p->pf_l * p->pf_logB + t * p->pf_logB));
    uint32_t* msg = (uint32_t*)calloc((size_t)N, sizeof(uint32_t));
    if (u == 1) {
        msg[0] = c;
    } else {
        for (int k = 0; k < N; ++k) msg[k] = 0u - (uint32_t)K2[k] * c;
    }
    orc_rlwe_encrypt32(seed, ORC_DOM_PFKS, orc_pfks_id(p, u, i, t, v), N, K2, NULL, msg, p->pf_nshift, ct);
    free(msg);
}

/* Sum of the selected entries, restated by linearity: every entry is
 * (a_e, a_e * K2 + e_e + msg_e) with a_e, e_e from Philox (orc_rlwe_encrypt32),
 * so the sum is (A, A * K2 + E + sum msg_e) with A = sum a_e, E = sum e_e; the
 * messages are c_e (coefficient 0, f_1) or -K2 c_e (f_0) with scalar c_e.  All
 * of it is exact mod 2^32, so the result equals adding the ciphertexts of
 * orc_gen_pfks_entry one by one, at one polynomial product instead of one per
 * entry.  OpenMP over i with per-thread partial sums (order-free mod 2^32). */
void orc_pfks(uint64_t seed, const orc_cbs_params* p, const int32_t* K2, int u, const uint64_t* lwe,
              uint32_t* out) {
    int N = p->N, B = 1 << p->pf_logB, kb = p->pf_l * p->pf_logB;
    uint32_t* A = (uint32_t*)calloc((size_t)N, sizeof(uint32_t));
    uint32_t* E = (uint32_t*)calloc((size_t)N, sizeof(uint32_t));
    uint32_t C = 0;
#pragma omp parallel
    {
        uint32_t* al = (uint32_t*)calloc((size_t)N, sizeof(uint32_t));
        uint32_t* el = (uint32_t*)calloc((size_t)N, sizeof(uint32_t));
        uint32_t cl = 0;
#pragma omp for schedule(static)
        for (int i = 0; i <= N; ++i) {
            uint64_t at = (lwe[i] + ((uint64_t)1 << (63 - kb))) >> (64 - kb);
            uint32_t kappa = i < N ? (uint32_t)K2[i] : 0xFFFFFFFFu;
            for (int t = 0; t < p->pf_l; ++t) {
                int v = (int)((at >> (t * p->pf_logB)) & (uint64_t)(B - 1));
                if (!v) continue;
                uint32_t id = orc_pfks_id(p, u, i, t, v);
                for (int j = 0; j < N; ++j) {
                    al[j] += orc_rand32(seed, ORC_DOM_PFKS, id, 0, (uint32_t)j);
                    el[j] += (uint32_t)orc_noise(seed, ORC_DOM_PFKS, id, (uint32_t)j, p->pf_nshift);
                }
                cl += (uint32_t)v * kappa * ((uint32_t)1 << (32 - kb + t * p->pf_logB));
            }
        }
#pragma omp critical
        {
            for (int j = 0; j < N; ++j) {
                A[j] += al[j];
                E[j] += el[j];
            }
            C += cl;
        }
        free(al);
        free(el);
    }
    uint32_t* b = (uint32_t*)malloc(sizeof(uint32_t) * N);
    orc_nc_mul_exact32(K2, A, b, N);
    for (int j = 0; j < N; ++j) {
        uint32_t m = u == 1 ? (j == 0 ? C : 0u) : 0u - (uint32_t)K2[j] * C;
        out[j] = 0u - A[j];
        out[N + j] = 0u - (b[j] + E[j] + m);
    }
    free(b);
    free(A);
    free(E);
}

/* RGSW row r = 2j + u = PFKS_u(PBS_j(ct)): [2 cb_l][2][N] u32 (the circuit bootstrap). */
double orc_cbs(uint64_t seed, const orc_cbs_params* p, const int32_t* K2, const double* bsk_f,
               const uint32_t* ct32, uint32_t* rgsw) {
    int N = p->N;
    uint64_t* lwes = (uint64_t*)malloc(sizeof(uint64_t) * 2 * (N + 1));
    double err = orc_pbs_manylut(p, bsk_f, ct32, lwes);
    for (int j = 0; j < p->cb_l; ++j)
        for (int u = 0; u < 2; ++u) orc_pfks(seed, p, K2, u, lwes + (size_t)j * (N + 1), rgsw + (size_t)(2 * j + u) * 2 * N);
    free(lwes);
    return err;
}

/* ------------------------------------------------------------------------- */
/* Linear-key private functional key switching on Z/2^40 (the textbook PrivKS */
/* form, CGGI 2020 Alg. 2, with one key row per (i, t) instead of the lookup  */
/* table of all digit multiples; the reference's configs name PFKS 2^4 x 8):  */
/*   KL[u][i][t] = RLWE_K2(f_u(kappa_i * 2^(32 - kb + t logB)) * 2^8) on      */
/*   Z/2^40: a = 40 low bits of Philox words, b = a * K2 + e + msg, e with    */
/*   sigma 2^(32 - ORC_PFKL_NSHIFT) units of 2^-40 (2^-33 of the torus);      */
/*   digits signed, d_t in [-B/2, B/2), a~_i = sum_t d_t B^t mod 2^kb;         */
/*   out40 = - sum_{i<=N} sum_t d_t(a~_i) KL[u][i][t]  (mod 2^40), and the    */
/*   torus-32 output is out32 = round(out40 / 2^8).                            */
/* Noise: the sum has variance ~16392 * E[d^2] * sigma^2 = 16392 * 21.5 *     */
/* 2^-66, i.e. sigma 2^-23.8 of the torus, below the lookup form's 2^-23.0.   */
/* ------------------------------------------------------------------------- */
#define PFKL_MASK ((((uint64_t)1) << ORC_PFKL_BITS) - 1)

static uint32_t pfkl_id(const orc_cbs_params* p, int u, int i, int t) {
    return (uint32_t)(((size_t)u * (p->N + 1) + i) * p->pf_l + t);
}

/* 40-bit message scalar of entry (i, t): kappa_i * 2^(32 - kb + t logB) on the
 * 32-bit torus, times 2^8 */
static uint64_t pfkl_cval(const orc_cbs_params* p, const int32_t* K2, int i, int t) {
    int kb = p->pf_l * p->pf_logB;
    uint32_t kappa = i < p->N ? (uint32_t)K2[i] : 0xFFFFFFFFu;
    uint32_t c32 = kappa * ((uint32_t)1 << (32 - kb + t * p->pf_logB));
    return ((uint64_t)c32 << 8) & PFKL_MASK;
}

/* exact negacyclic a * K mod 2^40 for binary K (schoolbook) */
static void nc_mul_binary40(const int32_t* K, const uint64_t* a, uint64_t* out, int N) {
    for (int c = 0; c < N; ++c) {
        uint64_t acc = 0;
        for (int j = 0; j < N; ++j) {
            if (!K[j]) continue;
            acc += j <= c ? a[c - j] : (0 - a[N + c - j]);
        }
        out[c] = acc & PFKL_MASK;
    }
}

void orc_gen_pfkl_entry(uint64_t seed, const orc_cbs_params* p, const int32_t* K2, int u, int i, int t,
                        uint64_t* ct40) {
    int N = p->N;
    uint32_t id = pfkl_id(p, u, i, t);
    uint64_t cval = pfkl_cval(p, K2, i, t);
    uint64_t* a = ct40;
    uint64_t* b = ct40 + N;
    for (int j = 0; j < N; ++j) a[j] = orc_rand64(seed, ORC_DOM_PFKL, id, (uint32_t)j) & PFKL_MASK;
    nc_mul_binary40(K2, a, b, N);
    for (int j = 0; j < N; ++j) {
        uint64_t m = u == 1 ? (j == 0 ? cval : 0) : (0 - (uint64_t)K2[j] * cval);
        b[j] = (b[j] + (uint64_t)orc_noise(seed, ORC_DOM_PFKL, id, (uint32_t)j, ORC_PFKL_NSHIFT) + m) & PFKL_MASK;
    }
}

/* The sum restated by linearity, as orc_pfks does: (A, A * K2 + E + msg(C)) with
 * A = sum d a_e, E = sum d e_e, C = sum d cval_e, all mod 2^40. */
void orc_pfks_linear(uint64_t seed, const orc_cbs_params* p, const int32_t* K2, int u, const uint64_t* lwe,
                     uint32_t* out) {
    int N = p->N, B = 1 << p->pf_logB, kb = p->pf_l * p->pf_logB;
    uint64_t kmask = kb == 64 ? ~(uint64_t)0 : (((uint64_t)1 << kb) - 1);
    uint64_t off = 0;
    for (int t = 0; t < p->pf_l; ++t) off += (uint64_t)(B / 2) << (t * p->pf_logB);
    uint64_t* A = (uint64_t*)calloc((size_t)N, sizeof(uint64_t));
    uint64_t* E = (uint64_t*)calloc((size_t)N, sizeof(uint64_t));
    uint64_t C = 0;
#pragma omp parallel
    {
        uint64_t* al = (uint64_t*)calloc((size_t)N, sizeof(uint64_t));
        uint64_t* el = (uint64_t*)calloc((size_t)N, sizeof(uint64_t));
        uint64_t cl = 0;
#pragma omp for schedule(static)
        for (int i = 0; i <= N; ++i) {
            uint64_t at = (lwe[i] + ((uint64_t)1 << (63 - kb))) >> (64 - kb);
            uint64_t w = (at + off) & kmask;
            for (int t = 0; t < p->pf_l; ++t) {
                int64_t d = (int64_t)((w >> (t * p->pf_logB)) & (uint64_t)(B - 1)) - B / 2;
                if (!d) continue;
                uint32_t id = pfkl_id(p, u, i, t);
                for (int j = 0; j < N; ++j) {
                    al[j] += (uint64_t)d * (orc_rand64(seed, ORC_DOM_PFKL, id, (uint32_t)j) & PFKL_MASK);
                    el[j] += (uint64_t)d * (uint64_t)orc_noise(seed, ORC_DOM_PFKL, id, (uint32_t)j, ORC_PFKL_NSHIFT);
                }
                cl += (uint64_t)d * pfkl_cval(p, K2, i, t);
            }
        }
#pragma omp critical
        {
            for (int j = 0; j < N; ++j) {
                A[j] += al[j];
                E[j] += el[j];
            }
            C += cl;
        }
        free(al);
        free(el);
    }
    for (int j = 0; j < N; ++j) A[j] &= PFKL_MASK;
    C &= PFKL_MASK;
    uint64_t* b = (uint64_t*)malloc(sizeof(uint64_t) * N);
    nc_mul_binary40(K2, A, b, N);
    for (int j = 0; j < N; ++j) {
        uint64_t m = u == 1 ? (j == 0 ? C : 0) : (0 - (uint64_t)K2[j] * C);
        uint64_t oa = (0 - A[j]) & PFKL_MASK;
        uint64_t ob = (0 - (b[j] + E[j] + m)) & PFKL_MASK;
        out[j] = (uint32_t)((oa + 128) >> 8);
        out[N + j] = (uint32_t)((ob + 128) >> 8);
    }
    free(b);
    free(A);
    free(E);
}

/* orc_cbs with the linear-key PFKS */
double orc_cbs_linear(uint64_t seed, const orc_cbs_params* p, const int32_t* K2, const double* bsk_f,
                      const uint32_t* ct32, uint32_t* rgsw) {
    int N = p->N;
    uint64_t* lwes = (uint64_t*)malloc(sizeof(uint64_t) * 2 * (N + 1));
    double err = orc_pbs_manylut(p, bsk_f, ct32, lwes);
    for (int j = 0; j < p->cb_l; ++j)
        for (int u = 0; u < 2; ++u)
            orc_pfks_linear(seed, p, K2, u, lwes + (size_t)j * (N + 1), rgsw + (size_t)(2 * j + u) * 2 * N);
    free(lwes);
    return err;
}

/* Fourier RGSW (torus-32, one limb): [2 cb_l rows][2 out][re M | im M]. */
void orc_rgsw_to_fourier32(int N, int rows, const uint32_t* rgsw, double* rgsw_f) {
    int M = N / 2;
    double* x = (double*)malloc(sizeof(double) * N);
    for (int q = 0; q < rows * 2; ++q) {
        for (int c = 0; c < N; ++c) x[c] = (double)(int32_t)rgsw[(size_t)q * N + c];
        orc_fft_forward(x, rgsw_f + (size_t)q * 2 * M, rgsw_f + (size_t)q * 2 * M + M, N);
    }
    free(x);
}

/* CMux(C, d0, d1) = d0 + C (*) (d1 - d0) on the 32-bit torus (reference
 * rlwe_tables.cuh:79-91); the external product is the exact one-limb FFT
 * product of orc_external_product32 (|digit| <= 2^7, |key| <= 2^31, N = 2048). */
double orc_cmux32(int N, int logB, int l, const double* rgsw_f, const uint32_t* d0, const uint32_t* d1,
                  uint32_t* out) {
    uint32_t* diff = (uint32_t*)malloc(sizeof(uint32_t) * 2 * N);
    uint32_t* t = (uint32_t*)malloc(sizeof(uint32_t) * 2 * N);
    for (int k = 0; k < 2 * N; ++k) diff[k] = d1[k] - d0[k];
    double e = orc_external_product32(N, logB, l, rgsw_f, diff, t);
    for (int k = 0; k < 2 * N; ++k) out[k] = d0[k] + t[k];
    free(diff);
    free(t);
    return e;
}

/* Vertical-packing LUT on the 32-bit torus: acc = (0, table_poly); for bit
 * i = 0..7 (least significant first, as EvalLUT8x32_RLWE_CPU,
 * rlwe_tables.cuh:93-108): acc = CMux(C_i, acc, X^{-stride 2^i} acc).  After
 * the loop coefficient j holds table_poly[stride x + j].
 * rgsw_f: [8][2 cb_l][2][2M] (orc_rgsw_to_fourier32 per bit). */
double orc_vertical_packing(const orc_cbs_params* p, int stride, const double* rgsw_f, const uint32_t* table_poly,
                            uint32_t* acc) {
    int N = p->N, M = N / 2, rows = 2 * p->cb_l;
    uint32_t* rot = (uint32_t*)malloc(sizeof(uint32_t) * 2 * N);
    double maxerr = 0.0;
    memset(acc, 0, sizeof(uint32_t) * N);
    memcpy(acc + N, table_poly, sizeof(uint32_t) * N);
    for (int i = 0; i < 8; ++i) {
        uint32_t sh = (uint32_t)(stride << i); /* rotate by X^{-sh} */
        for (int o = 0; o < 2; ++o)
            for (int c = 0; c < N; ++c) {
                uint32_t src = (uint32_t)c + sh; /* (X^{-sh} a)_c = a_{c+sh} (negated past N) */
                uint32_t v = acc[o * N + (src % (uint32_t)N)];
                if ((src / (uint32_t)N) & 1) v = 0u - v;
                rot[o * N + c] = v;
            }
        double e = orc_cmux32(N, p->cb_logB, p->cb_l, rgsw_f + (size_t)i * rows * 2 * 2 * M, acc, rot, acc);
        if (e > maxerr) maxerr = e;
    }
    free(rot);
    return maxerr;
}

/* KS2: LWE_K2 (torus-64, dim N) -> LWE_s (torus-32, dim n): modulus switch each
 * word 2^64 -> 2^32 (round half up) and key switch with the torus-32 key
 * KSK2 [N][ks_l][B][n+1] (the reference's Q -> qKS then KS order,
 * extract.cuh:66-75, aes_gpu_ks.cuh:341-370). */
void orc_ks2(const orc_cbs_params* p, const uint32_t* ksk2, const uint64_t* lwe64, uint32_t* out) {
    int N = p->N;
    uint32_t* l32 = (uint32_t*)malloc(sizeof(uint32_t) * (N + 1));
    for (int i = 0; i <= N; ++i) l32[i] = (uint32_t)((lwe64[i] + ((uint64_t)1 << 31)) >> 32);
    orc_gate_params g = {p->n, p->N, 0, 0, p->ks_logB, p->ks_l, p->lwe_nshift, 0};
    orc_keyswitch32(&g, ksk2, l32, out);
    free(l32);
}

void orc_gen_ksk2(uint64_t seed, const orc_cbs_params* p, const int32_t* s, const int32_t* K2, uint32_t* ksk) {
    orc_gate_params g = {p->n, p->N, 0, 0, p->ks_logB, p->ks_l, p->lwe_nshift, 0};
    orc_gen_ksk32_dom(seed, ORC_DOM_KSK2, &g, s, K2, ksk);
}

/* ------------------------------------------------------------------------- */
/* N -> n key switch with a linear key: the same switch as orc_ks2 (reference   */
/* kernel_KeySwitchU16, aes_gpu_ks.cuh:341-370: base 2^2, 8 digits of the       */
/* rounded top 16 bits), with one key row per (i, j),                          */
/*   KL[i][j] = LWE_s(K2_i 2^(32 - ks_l ks_logB + j ks_logB)),                  */
/* and signed digits d_j in [-B/2, B/2) by the offset method:                  */
/*   a~ = round(a_i / 2^(32 - 16)), d_j = ((a~ + sum_j (B/2) B^j) >> j logB mod B) - B/2, */
/* so sum_j d_j B^j = a~ mod 2^16 and                                          */
/*   out = (0, b) - sum_i sum_j d_j KL[i][j]  (mod 2^32)                        */
/* has the lookup form's phase (the wrap of the top digit is a multiple of     */
/* 2^32).  The sum is an integer GEMM (digits x key), which the GPU runs on the */
/* matrix cores.  Key rows: domain ORC_DOM_KSKL, id i l + j, noise lwe_nshift.  */
/* ------------------------------------------------------------------------- */
void orc_gen_kskl_entry(uint64_t seed, const orc_cbs_params* p, const int32_t* s, const int32_t* K2, int i, int j,
                        uint32_t* row) {
    const int kbits = p->ks_l * p->ks_logB;
    const uint32_t mu = (uint32_t)K2[i] * ((uint32_t)1 << (32 - kbits + j * p->ks_logB));
    orc_lwe_encrypt32(seed, ORC_DOM_KSKL, (uint32_t)(i * p->ks_l + j), p->n, s, mu, p->lwe_nshift, row);
}

void orc_gen_kskl(uint64_t seed, const orc_cbs_params* p, const int32_t* s, const int32_t* K2, uint32_t* kl) {
    for (int i = 0; i < p->N; ++i)
        for (int j = 0; j < p->ks_l; ++j)
            orc_gen_kskl_entry(seed, p, s, K2, i, j, kl + ((size_t)i * p->ks_l + j) * (p->n + 1));
}

void orc_ks2_linear32(const orc_cbs_params* p, const uint32_t* kl, const uint32_t* lwe32, uint32_t* out) {
    const int n = p->n, B = 1 << p->ks_logB, kbits = p->ks_l * p->ks_logB;
    uint32_t off = 0;
    for (int j = 0; j < p->ks_l; ++j) off += (uint32_t)(B / 2) << (j * p->ks_logB);
    for (int k = 0; k < n; ++k) out[k] = 0;
    out[n] = lwe32[p->N];
    for (int i = 0; i < p->N; ++i) {
        const uint32_t at = (uint32_t)(((uint64_t)lwe32[i] + ((uint64_t)1 << (31 - kbits))) >> (32 - kbits));
        for (int j = 0; j < p->ks_l; ++j) {
            const int32_t d = (int32_t)(((at + off) >> (j * p->ks_logB)) & (uint32_t)(B - 1)) - B / 2;
            const uint32_t* row = kl + ((size_t)i * p->ks_l + j) * (n + 1);
            for (int k = 0; k <= n; ++k) out[k] -= (uint32_t)d * row[k];
        }
    }
}

void orc_ks2_linear(const orc_cbs_params* p, const uint32_t* kl, const uint64_t* lwe64, uint32_t* out) {
    const int N = p->N;
    uint32_t* l32 = (uint32_t*)malloc(sizeof(uint32_t) * (N + 1));
    for (int i = 0; i <= N; ++i) l32[i] = (uint32_t)((lwe64[i] + ((uint64_t)1 << 31)) >> 32);
    orc_ks2_linear32(p, kl, l32, out);
    free(l32);
}

/* ------------------------------------------------------------------------- */
/* HomTrace + SchemeSwitch LWE -> RGSW (the reference's conversion,            */
/* cirbts-batch.cu:1612-1790; kernels kernel.cu:7536-7569 (automorphism maps), */
/* 7776-7789 (row save), 7832-7888 (trace round), 8407-8493 (scheme switch);   */
/* keys cirbts-init.cu:734-864), restated on the 64-bit torus with exact FFT    */
/* products.  RLWE ct = (a, b), phase b - a K2.                                 */
/*   trace round k = 0 .. log2 N - 1, automorphism sigma_k: X -> X^(N/2^k + 1): */
/*     (a, b) += (ks_0, sigma_k(b) + ks_1),  ks = sum_d dig_d(sigma_k(a)) HT_k,d */
/*     HT_k,d = RLWE_K2(-sigma_k(K2) g_d): (ks_0, sigma(b) + ks_1) is sigma(ct)  */
/*     switched back to K2, so each round adds the automorphic image and the    */
/*     trace multiplies the constant coefficient by N and kills the others.     */
/*   scheme switch: (b + ks_0, ks_1), ks = sum_d dig_d(a) SS_d,                 */
/*     SS_d = RLWE_K2(K2^2 g_d): phase a K2^2 - b K2 = -K2 (b - a K2).          */
/* Gadget 2^8 x 5 (40 of 64 bits, g_d = 2^(24 + 8 d)) for both; key noise as the */
/* BSK2's (rlwe_nshift).                                                        */
/* ------------------------------------------------------------------------- */
#define ORC_HT_LOGB 8
#define ORC_HT_L 5

/* out(X) = in(X^k) mod X^N + 1, k odd: coefficient i goes to i k mod 2N (negated past N) */
void orc_automorphism64(int N, const uint64_t* in, uint32_t k, uint64_t* out) {
    uint32_t Q2 = 2u * (uint32_t)N;
    for (int i = 0; i < N; ++i) {
        uint32_t t = (uint32_t)(((uint64_t)i * k) & (Q2 - 1));
        if (t < (uint32_t)N) out[t] = in[i];
        else out[t - (uint32_t)N] = 0 - in[i];
    }
}

static int orc_log2(int N) {
    int lg = 0;
    while ((1 << lg) < N) ++lg;
    return lg;
}

/* HT key [log2 N][HT_L][2][N]: entry (k, d) = RLWE_K2(-sigma_k(K2) g_d), Philox domain
 * ORC_DOM_HTKEY, id k HT_L + d */
void orc_gen_ht_key(uint64_t seed, const orc_cbs_params* p, const int32_t* K2, uint64_t* key) {
    int N = p->N, R = orc_log2(N);
    uint64_t* k64 = (uint64_t*)malloc(sizeof(uint64_t) * N);
    uint64_t* sk = (uint64_t*)malloc(sizeof(uint64_t) * N);
    uint64_t* msg = (uint64_t*)malloc(sizeof(uint64_t) * N);
    for (int i = 0; i < N; ++i) k64[i] = (uint64_t)(int64_t)K2[i];
    for (int k = 0; k < R; ++k) {
        orc_automorphism64(N, k64, (uint32_t)((N >> k) + 1), sk);
        for (int d = 0; d < ORC_HT_L; ++d) {
            uint64_t g = (uint64_t)1 << (64 - ORC_HT_L * ORC_HT_LOGB + d * ORC_HT_LOGB);
            for (int i = 0; i < N; ++i) msg[i] = 0 - sk[i] * g;
            orc_rlwe_encrypt64(seed, ORC_DOM_HTKEY, (uint32_t)(k * ORC_HT_L + d), N, K2, NULL, msg, p->rlwe_nshift,
                               key + ((size_t)k * ORC_HT_L + d) * 2 * N);
        }
    }
    free(k64); free(sk); free(msg);
}

/* SS key [HT_L][2][N]: entry d = RLWE_K2(K2^2 g_d), domain ORC_DOM_SSKEY, id d */
void orc_gen_ss_key(uint64_t seed, const orc_cbs_params* p, const int32_t* K2, uint64_t* key) {
    int N = p->N;
    uint64_t* k64 = (uint64_t*)malloc(sizeof(uint64_t) * N);
    uint64_t* sq = (uint64_t*)malloc(sizeof(uint64_t) * N);
    uint64_t* msg = (uint64_t*)malloc(sizeof(uint64_t) * N);
    for (int i = 0; i < N; ++i) k64[i] = (uint64_t)(int64_t)K2[i];
    mul_binary64(K2, k64, sq, N);
    for (int d = 0; d < ORC_HT_L; ++d) {
        uint64_t g = (uint64_t)1 << (64 - ORC_HT_L * ORC_HT_LOGB + d * ORC_HT_LOGB);
        for (int i = 0; i < N; ++i) msg[i] = sq[i] * g;
        orc_rlwe_encrypt64(seed, ORC_DOM_SSKEY, (uint32_t)d, N, K2, NULL, msg, p->rlwe_nshift,
                           key + (size_t)d * 2 * N);
    }
    free(k64); free(sq); free(msg);
}

/* Fourier form of nkeys key-switching keys of HT_L rows each, in the external-product
 * layout [2 HT_L rows][2 out][4M] with row 2d = entry d and row 2d + 1 zero (so that
 * orc_external_product64 of (x, 0) is sum_d dig_d(x) entry_d) */
void orc_ksk64_to_fourier(const orc_cbs_params* p, int nkeys, const uint64_t* key, double* key_f) {
    int N = p->N, M = N / 2;
    size_t per = (size_t)2 * ORC_HT_L * 2 * 4 * M;
    memset(key_f, 0, sizeof(double) * per * (size_t)nkeys);
    for (int q = 0; q < nkeys; ++q)
        for (int d = 0; d < ORC_HT_L; ++d)
            for (int o = 0; o < 2; ++o)
                orc_to_fourier64(key + (((size_t)q * ORC_HT_L + d) * 2 + o) * N, N,
                                 key_f + (size_t)q * per + ((size_t)(2 * d) * 2 + o) * 4 * M);
}

/* sum_d dig_d(x) key_d, exact (x [N], out [2][N]) */
static double ks64(int N, const double* key_f, const uint64_t* x, uint64_t* out) {
    uint64_t* acc = (uint64_t*)calloc((size_t)2 * N, sizeof(uint64_t));
    memcpy(acc, x, sizeof(uint64_t) * N);
    double e = orc_external_product64(N, ORC_HT_LOGB, ORC_HT_L, key_f, acc, out);
    free(acc);
    return e;
}

/* HomTrace of ct [2][N] in place, all log2 N rounds (htkey_f: orc_ksk64_to_fourier of
 * the HT key) */
double orc_homtrace64(const orc_cbs_params* p, const double* htkey_f, uint64_t* ct) {
    int N = p->N, M = N / 2, R = orc_log2(N);
    size_t per = (size_t)2 * ORC_HT_L * 2 * 4 * M;
    uint64_t* sa = (uint64_t*)malloc(sizeof(uint64_t) * N);
    uint64_t* sb = (uint64_t*)malloc(sizeof(uint64_t) * N);
    uint64_t* ks = (uint64_t*)malloc(sizeof(uint64_t) * 2 * N);
    double maxerr = 0.0;
    for (int k = 0; k < R; ++k) {
        uint32_t e = (uint32_t)((N >> k) + 1);
        orc_automorphism64(N, ct, e, sa);
        orc_automorphism64(N, ct + N, e, sb);
        double er = ks64(N, htkey_f + (size_t)k * per, sa, ks);
        if (er > maxerr) maxerr = er;
        for (int i = 0; i < N; ++i) {
            ct[i] += ks[i];
            ct[N + i] += sb[i] + ks[N + i];
        }
    }
    free(sa); free(sb); free(ks);
    return maxerr;
}

/* SchemeSwitch: out = (b + ks_0, ks_1), ks = sum_d dig_d(a) SS_d: an RLWE of -K2 m */
double orc_schemeswitch64(const orc_cbs_params* p, const double* sskey_f, const uint64_t* ct, uint64_t* out) {
    int N = p->N;
    double e = ks64(N, sskey_f, ct, out);
    for (int i = 0; i < N; ++i) out[i] += ct[N + i];
    return e;
}

/* Circuit bootstrap through HomTrace + SchemeSwitch: the accumulator of the blind
 * rotation of LUT[k] = -2^(47 + 8 (k mod 2)) / N (b + 2^30 switched with bw = 1), then
 * for each CB level j: acc X^-j (coefficient j to 0) plus 2^(47 + 8j) / N at coefficient
 * 0 (the level's value bit 2^(48 + 8j) / N there), HomTrace (times N: an RLWE of
 * bit 2^(48 + 8j)), SchemeSwitch; rows rounded to the 32-bit torus (top 32 bits):
 * row 2j = SS (RLWE of -K2 bit 2^(16 + 8j)), row 2j + 1 = HT (bit 2^(16 + 8j)), the
 * reference's row order (kernel.cu:7776-7789) and the PFKS forms' row meaning. */
double orc_cbs_htss(const orc_cbs_params* p, const double* bsk_f, const double* htkey_f, const double* sskey_f,
                    const uint32_t* ct32, uint32_t* rgsw) {
    int N = p->N, lg = orc_log2(N);
    uint64_t* lut = (uint64_t*)malloc(sizeof(uint64_t) * N);
    uint64_t* acc = (uint64_t*)malloc(sizeof(uint64_t) * 2 * N);
    uint64_t* lv = (uint64_t*)malloc(sizeof(uint64_t) * 2 * N);
    uint64_t* ss = (uint64_t*)malloc(sizeof(uint64_t) * 2 * N);
    uint32_t* in = (uint32_t*)malloc(sizeof(uint32_t) * (p->n + 1));
    memcpy(in, ct32, sizeof(uint32_t) * p->n);
    in[p->n] = ct32[p->n] + (1u << 30);
    for (int k = 0; k < N; ++k)
        lut[k] = 0 - ((uint64_t)1 << (64 - p->cb_l * p->cb_logB + (k & 1) * p->cb_logB - 1 - lg));
    double err = orc_blind_rotate_lut64(p, bsk_f, 1, in, lut, 1, acc);
    for (int j = 0; j < p->cb_l; ++j) {
        for (int o = 0; o < 2; ++o)
            for (int i = 0; i < N; ++i)  /* acc X^-j */
                lv[o * N + i] = i + j < N ? acc[o * N + i + j] : 0 - acc[o * N + i + j - N];
        lv[N] += (uint64_t)1 << (64 - p->cb_l * p->cb_logB + j * p->cb_logB - 1 - lg);
        double e = orc_homtrace64(p, htkey_f, lv);
        if (e > err) err = e;
        e = orc_schemeswitch64(p, sskey_f, lv, ss);
        if (e > err) err = e;
        for (int i = 0; i < 2 * N; ++i) {
            rgsw[(size_t)(2 * j) * 2 * N + i] = (uint32_t)((ss[i] + ((uint64_t)1 << 31)) >> 32);
            rgsw[(size_t)(2 * j + 1) * 2 * N + i] = (uint32_t)((lv[i] + ((uint64_t)1 << 31)) >> 32);
        }
    }
    free(lut); free(acc); free(lv); free(ss); free(in);
    return err;
}
