/*
 * tfhe_oracle.c -- CPU restatement of the reference's gate-bootstrapping path.
 *
 * TEST INFRASTRUCTURE ONLY (see tfhe_oracle.h).  Reference citations are
 * relative to the reference repository root.
 */
#include "tfhe_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------- */
/* Philox4x32-10, the counter-based generator of Salmon, Moraes, Dror, Shaw,   */
/* "Parallel random numbers: as easy as 1, 2, 3" (SC'11).  The reference gets */
/* its randomness from OpenFHE; its key material is absent, so the build      */
/* derives every key and ciphertext from this published generator instead.   */
/* ------------------------------------------------------------------------- */
void orc_philox4x32_10(const uint32_t key[2], const uint32_t ctr[4], uint32_t out[4]) {
    uint32_t k0 = key[0], k1 = key[1];
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static void ph(uint64_t seed, uint32_t j, uint32_t id, uint32_t sub, uint32_t dom, uint32_t out[4]) {
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t ctr[4] = {j, id, sub, dom};
    orc_philox4x32_10(key, ctr, out);
}

uint32_t orc_rand32(uint64_t seed, uint32_t dom, uint32_t id, uint32_t sub, uint32_t j) {
    uint32_t w[4];
    ph(seed, j, id, sub, dom, w);
    return w[0];
}

uint64_t orc_rand64(uint64_t seed, uint32_t dom, uint32_t id, uint32_t j) {
    uint32_t w[4];
    ph(seed, j, id, 0, dom, w);
    return (uint64_t)w[0] | ((uint64_t)w[1] << 32);
}

int64_t orc_noise(uint64_t seed, uint32_t dom, uint32_t id, uint32_t j, int nshift) {
    uint64_t s = 0;
    for (uint32_t sub = 1; sub <= 3; ++sub) {
        uint32_t w[4];
        ph(seed, j, id, sub, dom, w);
        s += (uint64_t)w[0] + w[1] + w[2] + w[3];
    }
    int64_t z = (int64_t)s - (int64_t)6 * ((int64_t)1 << 32);
    if (nshift > 0) return (z + ((int64_t)1 << (nshift - 1))) >> nshift;
    return z * ((int64_t)1 << (-nshift));
}

int32_t orc_secret_bit(uint64_t seed, uint32_t dom, uint32_t i) {
    uint32_t w[4];
    ph(seed, i, 0, 0, dom, w);
    return (int32_t)(w[0] & 1u);
}

/* ------------------------------------------------------------------------- */
/* Negacyclic products in Z[X]/(X^N+1).                                        */
/* ------------------------------------------------------------------------- */

/* Schoolbook, exact mod 2^32 (the brute-force pin for the FFT paths). */
void orc_nc_mul_exact32(const int32_t* x, const uint32_t* y, uint32_t* out, int N) {
    uint32_t* acc = (uint32_t*)calloc((size_t)N, sizeof(uint32_t));
    for (int j = 0; j < N; ++j) {
        uint32_t xj = (uint32_t)x[j];
        if (!xj) continue;
        /* X^j * y: coefficient c gets y[c-j] for c >= j and -y[N+c-j] below */
        for (int c = j; c < N; ++c) acc[c] += xj * y[c - j];
        for (int c = 0; c < j; ++c) acc[c] -= xj * y[N + c - j];
    }
    memcpy(out, acc, sizeof(uint32_t) * (size_t)N);
    free(acc);
}

void orc_nc_mul_exact64(const int32_t* x, const uint64_t* y, uint64_t* out, int N) {
    uint64_t* acc = (uint64_t*)calloc((size_t)N, sizeof(uint64_t));
    for (int j = 0; j < N; ++j) {
        uint64_t xj = (uint64_t)(int64_t)x[j];
        if (!xj) continue;
        for (int c = j; c < N; ++c) acc[c] += xj * y[c - j];
        for (int c = 0; c < j; ++c) acc[c] -= xj * y[N + c - j];
    }
    memcpy(out, acc, sizeof(uint64_t) * (size_t)N);
    free(acc);
}

/*
 * Negacyclic FP64 FFT by folding and twisting (reference bitwise/src/fft.cu:3614-3754):
 * z_j = (x_j + i x_{j+N/2}) * w^j with w = exp(i pi / N), then an N/2-point complex
 * DFT.  The inverse undoes the DFT, multiplies by conj(w^j) and unfolds the real and
 * imaginary parts back into coefficients j and j+N/2.
 */
/* twiddle tables cached per transform size (M <= 4096) */
typedef struct { int M; double* wr; double* wi; double* tc; double* ts; } fft_tab;
static fft_tab g_tabs[8];

static const fft_tab* get_tab(int M) {
    for (int i = 0; i < 8; ++i)
        if (g_tabs[i].M == M) return &g_tabs[i];
    for (int i = 0; i < 8; ++i)
        if (g_tabs[i].M == 0) {
            fft_tab* t = &g_tabs[i];
            t->wr = (double*)malloc(sizeof(double) * M);
            t->wi = (double*)malloc(sizeof(double) * M);
            t->tc = (double*)malloc(sizeof(double) * M);
            t->ts = (double*)malloc(sizeof(double) * M);
            for (int k = 0; k < M; ++k) {
                t->wr[k] = cos(2.0 * M_PI * k / M);
                t->wi[k] = sin(2.0 * M_PI * k / M);
                t->tc[k] = cos(M_PI * k / (2.0 * M)); /* twist w^k, w = exp(i pi / N), N = 2M */
                t->ts[k] = sin(M_PI * k / (2.0 * M));
            }
            t->M = M;
            return t;
        }
    return NULL;
}

static void dft_inplace(double* re, double* im, int M, int inverse) {
    const fft_tab* T = get_tab(M);
    /* iterative radix-2, bit-reversed input permutation */
    for (int i = 1, j = 0; i < M; ++i) {
        int bit = M >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double t = re[i]; re[i] = re[j]; re[j] = t;
            t = im[i]; im[i] = im[j]; im[j] = t;
        }
    }
    const double sgn = inverse ? 1.0 : -1.0;
    for (int len = 2; len <= M; len <<= 1) {
        int half = len >> 1, step = M / len;
        for (int k = 0; k < half; ++k) {
            double wr = T->wr[k * step], wi = sgn * T->wi[k * step];
            for (int s = 0; s < M; s += len) {
                int a = s + k, b = s + k + half;
                double xr = re[b] * wr - im[b] * wi;
                double xi = re[b] * wi + im[b] * wr;
                re[b] = re[a] - xr; im[b] = im[a] - xi;
                re[a] += xr; im[a] += xi;
            }
        }
    }
}

void orc_fft_forward(const double* x, double* re, double* im, int N) {
    int M = N / 2;
    const fft_tab* T = get_tab(M);
    for (int j = 0; j < M; ++j) {
        double a = x[j], b = x[j + M];
        double c = T->tc[j], s = T->ts[j];
        re[j] = a * c - b * s;
        im[j] = a * s + b * c;
    }
    dft_inplace(re, im, M, 0);
}

void orc_fft_inverse(const double* re_in, const double* im_in, double* x, int N) {
    int M = N / 2;
    double* re = (double*)malloc(sizeof(double) * M);
    double* im = (double*)malloc(sizeof(double) * M);
    memcpy(re, re_in, sizeof(double) * M);
    memcpy(im, im_in, sizeof(double) * M);
    dft_inplace(re, im, M, 1);
    const fft_tab* T = get_tab(M);
    for (int j = 0; j < M; ++j) {
        double c = T->tc[j], s = -T->ts[j];
        double r = (re[j] * c - im[j] * s) / M;
        double i = (re[j] * s + im[j] * c) / M;
        x[j] = r;
        x[j + M] = i;
    }
    free(re);
    free(im);
}

/* Exact integer product through the FFT; returns the largest distance of an
 * inverse-FFT output from its nearest integer (exact while < 0.5). */
double orc_nc_mul_fft(const int64_t* x, const int64_t* y, int64_t* out, int N) {
    int M = N / 2;
    double* xd = (double*)malloc(sizeof(double) * N);
    double* yd = (double*)malloc(sizeof(double) * N);
    double* xr = (double*)malloc(sizeof(double) * M);
    double* xi = (double*)malloc(sizeof(double) * M);
    double* yr = (double*)malloc(sizeof(double) * M);
    double* yi = (double*)malloc(sizeof(double) * M);
    for (int j = 0; j < N; ++j) { xd[j] = (double)x[j]; yd[j] = (double)y[j]; }
    orc_fft_forward(xd, xr, xi, N);
    orc_fft_forward(yd, yr, yi, N);
    for (int k = 0; k < M; ++k) {
        double r = xr[k] * yr[k] - xi[k] * yi[k];
        double i = xr[k] * yi[k] + xi[k] * yr[k];
        xr[k] = r; xi[k] = i;
    }
    orc_fft_inverse(xr, xi, xd, N);
    double maxerr = 0.0;
    for (int j = 0; j < N; ++j) {
        double r = nearbyint(xd[j]);
        double e = fabs(xd[j] - r);
        if (e > maxerr) maxerr = e;
        out[j] = (int64_t)r;
    }
    free(xd); free(yd); free(xr); free(xi); free(yr); free(yi);
    return maxerr;
}

/* ------------------------------------------------------------------------- */
/* Signed gadget decomposition (reference bitwise/src/kernel.cu:25-52 and      */
/* bitwise/src/cirbts/kernel.cu:37-68): centre the torus value, drop the low   */
/* digits with rounding (approximate gadget, rgsw-acc.cu:595-600), then emit   */
/* l signed digits in [-B/2, B/2), least significant first.  Digit j carries   */
/* weight 2^(W - l*logB + j*logB) for a W-bit torus.                           */
/* ------------------------------------------------------------------------- */
static inline int64_t sext(int64_t v, int bits) {
    uint64_t m = (uint64_t)1 << (bits - 1);
    uint64_t u = (uint64_t)v & (((uint64_t)1 << bits) - 1);
    return (int64_t)(u ^ m) - (int64_t)m;
}

/* d' = (d - r) >> bits without overflow: (d >> bits) + (r < 0) */
static inline int64_t carry_shift(int64_t d, int64_t r, int bits) {
    return (d >> bits) + (r < 0 ? 1 : 0);
}

void orc_decompose32(uint32_t t, int logB, int l, int32_t* digits) {
    int ign = 32 - l * logB;
    int64_t d = (int64_t)(int32_t)t;
    if (ign > 0) {
        int64_t r = sext(d, ign);
        d = carry_shift(d, r, ign);
    }
    for (int j = 0; j < l; ++j) {
        int64_t r = sext(d, logB);
        digits[j] = (int32_t)r;
        d = carry_shift(d, r, logB);
    }
}

void orc_decompose64(uint64_t t, int logB, int l, int32_t* digits) {
    int ign = 64 - l * logB;
    int64_t d = (int64_t)t;
    if (ign > 0) {
        int64_t r = sext(d, ign);
        d = carry_shift(d, r, ign);
    }
    for (int j = 0; j < l; ++j) {
        int64_t r = sext(d, logB);
        digits[j] = (int32_t)r;
        d = carry_shift(d, r, logB);
    }
}

/* ------------------------------------------------------------------------- */
/* Key generation.  Reference: keys come from the OpenFHE fork                 */
/* (binfhe-base-scheme.cu:18-97 KeySwitchGen + GPUKeyGenAcc, rgsw-acc.cu:13-122)*/
/* which is absent; the layouts below are the ones the product consumes.       */
/* ------------------------------------------------------------------------- */
void orc_gen_lwe_sk(uint64_t seed, int n, int32_t* s) {
    for (int i = 0; i < n; ++i) s[i] = orc_secret_bit(seed, ORC_DOM_LWE_SK, (uint32_t)i);
}

void orc_gen_rlwe_sk(uint64_t seed, uint32_t dom, int N, int32_t* K) {
    for (int i = 0; i < N; ++i) K[i] = orc_secret_bit(seed, dom, (uint32_t)i);
}

/* ct = (a + msg_a, a*K + e + msg_b), torus-32; msg pointers may be NULL. */
void orc_rlwe_encrypt32(uint64_t seed, uint32_t dom, uint32_t id, int N, const int32_t* K,
                        const uint32_t* msg_a, const uint32_t* msg_b, int nshift, uint32_t* ct) {
    uint32_t* a = ct;
    uint32_t* b = ct + N;
    for (int j = 0; j < N; ++j) a[j] = orc_rand32(seed, dom, id, 0, (uint32_t)j);
    orc_nc_mul_exact32(K, a, b, N);
    for (int j = 0; j < N; ++j) {
        b[j] += (uint32_t)orc_noise(seed, dom, id, (uint32_t)j, nshift);
        if (msg_b) b[j] += msg_b[j];
        if (msg_a) a[j] += msg_a[j];
    }
}

void orc_lwe_encrypt32(uint64_t seed, uint32_t dom, uint32_t id, int n, const int32_t* s,
                       uint32_t mu, int nshift, uint32_t* ct) {
    uint32_t b = 0;
    for (int i = 0; i < n; ++i) {
        ct[i] = orc_rand32(seed, dom, id, 0, (uint32_t)i);
        b += ct[i] * (uint32_t)s[i];
    }
    b += (uint32_t)orc_noise(seed, dom, id, 0, nshift);
    ct[n] = b + mu;
}

uint32_t orc_lwe_phase32(int n, const int32_t* s, const uint32_t* ct) {
    uint32_t acc = ct[n];
    for (int i = 0; i < n; ++i) acc -= ct[i] * (uint32_t)s[i];
    return acc;
}

/* BSK layout [n][2*l rows][2 (a,b)][N]; row r = 2*j + u adds s_i * g_j to
 * component u, g_j = 2^(32 - l*logB + j*logB).  Row interleaving follows the
 * reference's digit-poly order (poly_idx & 1 selects a/b, poly_idx >> 1 the
 * digit: bitwise/src/kernel.cu:30-49). */
void orc_gen_bsk32(uint64_t seed, const orc_gate_params* p, const int32_t* s, const int32_t* K,
                   uint32_t* bsk) {
    int N = p->N, rows = 2 * p->bk_l;
    uint32_t* msg = (uint32_t*)calloc((size_t)N, sizeof(uint32_t));
    for (int i = 0; i < p->n; ++i) {
        for (int r = 0; r < rows; ++r) {
            int j = r >> 1, u = r & 1;
            uint32_t g = (uint32_t)1 << (32 - p->bk_l * p->bk_logB + j * p->bk_logB);
            memset(msg, 0, sizeof(uint32_t) * (size_t)N);
            msg[0] = (uint32_t)s[i] * g;
            uint32_t* ct = bsk + ((size_t)i * rows + r) * 2 * (size_t)N;
            uint32_t id = (uint32_t)(i * rows + r);
            orc_rlwe_encrypt32(seed, ORC_DOM_BSK1, id, N, K, u == 0 ? msg : NULL, u == 1 ? msg : NULL,
                               p->rlwe_nshift, ct);
        }
    }
    free(msg);
}

/* KSK layout [N][ks_l][B][n+1]: entry (i, j, v) = LWE_s(v * K_i * 2^(32 - ks_l*ks_logB + j*ks_logB)). */
void orc_gen_ksk32(uint64_t seed, const orc_gate_params* p, const int32_t* s, const int32_t* K,
                   uint32_t* ksk) {
    orc_gen_ksk32_dom(seed, ORC_DOM_KSK1, p, s, K, ksk);
}

void orc_gen_ksk32_dom(uint64_t seed, uint32_t dom, const orc_gate_params* p, const int32_t* s,
                       const int32_t* K, uint32_t* ksk) {
    int B = 1 << p->ks_logB;
    int kbits = p->ks_l * p->ks_logB;
    for (int i = 0; i < p->N; ++i)
        for (int j = 0; j < p->ks_l; ++j)
            for (int v = 0; v < B; ++v) {
                uint32_t id = (uint32_t)(((size_t)i * p->ks_l + j) * B + v);
                uint32_t mu = (uint32_t)v * (uint32_t)K[i] * ((uint32_t)1 << (32 - kbits + j * p->ks_logB));
                orc_lwe_encrypt32(seed, dom, id, p->n, s, mu, p->lwe_nshift,
                                  ksk + (size_t)id * (p->n + 1));
            }
}

/* ------------------------------------------------------------------------- */
/* Gate bootstrapping (reference bitwise/src/binfhe-base-scheme.cu:99-252,     */
/* bitwise/src/rgsw-acc.cu:580-708, bitwise/src/kernels_fft.cu:47-137).         */
/* ------------------------------------------------------------------------- */

/* Fourier form of the BSK: [n][rows][2][M] complex as (re[M], im[M]) pairs. */
void orc_bsk_to_fourier(const orc_gate_params* p, const uint32_t* bsk, double* bsk_f) {
    int N = p->N, M = N / 2, rows = 2 * p->bk_l;
    double* x = (double*)malloc(sizeof(double) * N);
    size_t npoly = (size_t)p->n * rows * 2;
    for (size_t q = 0; q < npoly; ++q) {
        const uint32_t* src = bsk + q * N;
        for (int j = 0; j < N; ++j) x[j] = (double)(int32_t)src[j];
        orc_fft_forward(x, bsk_f + q * 2 * M, bsk_f + q * 2 * M + M, N);
    }
    free(x);
}

/* round(v * 2N / 2^32) mod 2N, ties up (reference RoundqQ, binfhe-base-scheme.cu:11-17,
 * restated in exact integer arithmetic). */
uint32_t orc_modswitch_2N(uint32_t v, int N) {
    int logN2 = 0;
    while ((1 << logN2) < 2 * N) ++logN2;
    uint64_t r = ((uint64_t)v + ((uint64_t)1 << (31 - logN2))) >> (32 - logN2);
    return (uint32_t)(r & (uint64_t)(2 * N - 1));
}

/* OpenFHE BINGATE order; gate constants in eighths of the modulus. */
static int gate_const(int gate) {
    switch (gate) {
        case 0: return 5;  /* OR */
        case 1: return 7;  /* AND */
        case 2: return 1;  /* NOR */
        case 3: return 3;  /* NAND */
        case 4: return 5;  /* XOR_FAST */
        case 5: return 1;  /* XNOR_FAST */
        case 11: return 5; /* XOR */
        case 12: return 1; /* XNOR */
        default: return -1;
    }
}

static int gate_doubles(int gate) { return gate == 4 || gate == 5 || gate == 11 || gate == 12; }

/* Test vector m (b-part of the initial accumulator), reference
 * binfhe-base-scheme.cu:103-138 with q = 2N (sparse-embedding factor 1). */
void orc_gate_test_vector(int N, int gate, uint32_t b_tilde, uint32_t* tv) {
    uint32_t q = 2u * (uint32_t)N;
    uint32_t q1 = (uint32_t)gate_const(gate) * (q >> 3);
    uint32_t q2 = (q1 + (q >> 1)) % q;
    uint32_t Q2p = (1u << 29) + 1u; /* Q/(2p) + 1 with Q = 2^32, p = 4 */
    uint32_t Q2pNeg = 0u - Q2p;
    for (uint32_t j = 0; j < (uint32_t)N; ++j) {
        uint32_t temp = (b_tilde + q - j) % q;
        if (q1 < q2)
            tv[j] = (temp >= q1 && temp < q2) ? Q2pNeg : Q2p;
        else
            tv[j] = (temp >= q2 && temp < q1) ? Q2p : Q2pNeg;
    }
}

/* t = sum_r dec_r(acc) * RGSW_r, exact through the FP64 FFT and rounded mod 2^32.
 * Digit poly r = 2j+u holds digit j of component u.  Returns the max rounding error. */
double orc_external_product32(int N, int logB, int l, const double* rgsw_f, const uint32_t* acc,
                              uint32_t* out) {
    int M = N / 2, rows = 2 * l;
    double* x = (double*)malloc(sizeof(double) * N);
    double* dr = (double*)malloc(sizeof(double) * M);
    double* di = (double*)malloc(sizeof(double) * M);
    double* accr = (double*)calloc((size_t)2 * M, sizeof(double));
    double* acci = (double*)calloc((size_t)2 * M, sizeof(double));
    int32_t* dig = (int32_t*)malloc(sizeof(int32_t) * (size_t)rows * N);
    int32_t tmp[64];
    for (int u = 0; u < 2; ++u)
        for (int c = 0; c < N; ++c) {
            orc_decompose32(acc[u * N + c], logB, l, tmp);
            for (int j = 0; j < l; ++j) dig[(size_t)(2 * j + u) * N + c] = tmp[j];
        }
    for (int r = 0; r < rows; ++r) {
        for (int c = 0; c < N; ++c) x[c] = (double)dig[(size_t)r * N + c];
        orc_fft_forward(x, dr, di, N);
        for (int o = 0; o < 2; ++o) {
            const double* kr = rgsw_f + ((size_t)r * 2 + o) * 2 * M;
            const double* ki = kr + M;
            for (int k = 0; k < M; ++k) {
                accr[o * M + k] += dr[k] * kr[k] - di[k] * ki[k];
                acci[o * M + k] += dr[k] * ki[k] + di[k] * kr[k];
            }
        }
    }
    double maxerr = 0.0;
    for (int o = 0; o < 2; ++o) {
        orc_fft_inverse(accr + o * M, acci + o * M, x, N);
        for (int c = 0; c < N; ++c) {
            double rr = nearbyint(x[c]);
            double e = fabs(x[c] - rr);
            if (e > maxerr) maxerr = e;
            out[o * N + c] = (uint32_t)(uint64_t)(int64_t)rr;
        }
    }
    free(x); free(dr); free(di); free(accr); free(acci); free(dig);
    return maxerr;
}

/* acc += X^m * t - t on both components (reference kernels_fft.cu:90-113). */
void orc_monic_update32(int N, uint32_t m, const uint32_t* t, uint32_t* acc) {
    int k = (int)(m & (uint32_t)(N - 1));
    int hi = m >= (uint32_t)N;
    for (int o = 0; o < 2; ++o)
        for (int c = 0; c < N; ++c) {
            int idx = c - k;
            int wrapped = idx < 0;
            if (wrapped) idx += N;
            uint32_t v = t[o * N + idx];
            if (hi ^ wrapped) v = 0u - v;
            acc[o * N + c] += v - t[o * N + c];
        }
}

/* CGGI blind rotation with binary secrets, rotation index (2N - a~_i) mod 2N
 * (reference rgsw-acc.cu:676-708). */
double orc_blind_rotate32(const orc_gate_params* p, const double* bsk_f, const uint32_t* a_tilde,
                          uint32_t* acc) {
    int N = p->N, M = N / 2, rows = 2 * p->bk_l;
    uint32_t* t = (uint32_t*)malloc(sizeof(uint32_t) * 2 * N);
    double maxerr = 0.0;
    for (int i = 0; i < p->n; ++i) {
        uint32_t m = (2u * (uint32_t)N - a_tilde[i]) & (2u * (uint32_t)N - 1u);
        if (m == 0) continue; /* X^0 t - t = 0 */
        const double* key = bsk_f + (size_t)i * rows * 2 * 2 * M;
        double e = orc_external_product32(N, p->bk_logB, p->bk_l, key, acc, t);
        if (e > maxerr) maxerr = e;
        orc_monic_update32(N, m, t, acc);
    }
    free(t);
    return maxerr;
}

/* LWE of coefficient idx: a'_j = a_{idx-j} (j <= idx), -a_{N+idx-j} (j > idx), b' = b_idx
 * (reference bitwise/examples/aes_transcipher/extract.cuh:22-31; idx = 0 is the gate
 * path's X -> X^{2N-1} automorphism, bitwise/src/kernel.cu:241-250). */
void orc_sample_extract32(int N, const uint32_t* acc, uint32_t idx, uint32_t* lwe) {
    for (uint32_t j = 0; j < (uint32_t)N; ++j)
        lwe[j] = j <= idx ? acc[idx - j] : 0u - acc[(uint32_t)N + idx - j];
    lwe[N] = acc[N + idx];
}

/* out = (0, b) - sum_{i<N} sum_{j<l} KSK[i][j][digit_j(a_i)] with a_i rounded to
 * l*logB bits and unsigned digits, least significant first (reference
 * bitwise/src/kernel.cu:259-279 and the host B loop binfhe-base-scheme.cu:225-246). */
void orc_keyswitch32(const orc_gate_params* p, const uint32_t* ksk, const uint32_t* lweN, uint32_t* out) {
    int n = p->n, B = 1 << p->ks_logB, kbits = p->ks_l * p->ks_logB;
    for (int k = 0; k < n; ++k) out[k] = 0;
    out[n] = lweN[p->N];
    for (int i = 0; i < p->N; ++i) {
        uint32_t at = (uint32_t)(((uint64_t)lweN[i] + ((uint64_t)1 << (31 - kbits))) >> (32 - kbits));
        for (int j = 0; j < p->ks_l; ++j) {
            uint32_t v = (at >> (j * p->ks_logB)) & (uint32_t)(B - 1);
            const uint32_t* row = ksk + (((size_t)i * p->ks_l + j) * B + v) * (n + 1);
            for (int k = 0; k <= n; ++k) out[k] -= row[k];
        }
    }
}

/* ct1 + ct2, doubled for XOR/XNOR (reference binfhe-base-scheme.cu:151-176).
 * Returns -1 for gates the bootstrap does not serve. */
int orc_gate_prep(int gate, int n, const uint32_t* ct1, const uint32_t* ct2, uint32_t* prep) {
    if (gate_const(gate) < 0) return -1;
    for (int k = 0; k <= n; ++k) {
        uint32_t v = ct1[k] + ct2[k];
        prep[k] = gate_doubles(gate) ? 2u * v : v;
    }
    return 0;
}

double orc_gate_bootstrap(const orc_gate_params* p, const double* bsk_f, const uint32_t* ksk,
                          int gate, const uint32_t* ct1, const uint32_t* ct2, uint32_t* out) {
    int n = p->n, N = p->N;
    uint32_t* prep = (uint32_t*)malloc(sizeof(uint32_t) * (n + 1));
    uint32_t* at = (uint32_t*)malloc(sizeof(uint32_t) * n);
    uint32_t* acc = (uint32_t*)calloc((size_t)2 * N, sizeof(uint32_t));
    uint32_t* lwe = (uint32_t*)malloc(sizeof(uint32_t) * (N + 1));
    if (orc_gate_prep(gate, n, ct1, ct2, prep) != 0) {
        free(prep); free(at); free(acc); free(lwe);
        return -1.0;
    }
    for (int i = 0; i < n; ++i) at[i] = orc_modswitch_2N(prep[i], N);
    uint32_t bt = orc_modswitch_2N(prep[n], N);
    orc_gate_test_vector(N, gate, bt, acc + N);
    double e = orc_blind_rotate32(p, bsk_f, at, acc);
    orc_sample_extract32(N, acc, 0, lwe);
    lwe[N] += (1u << 29) + 1u; /* + Q/8 + 1 maps back to Q/4 arithmetic (binfhe-base-scheme.cu:222) */
    orc_keyswitch32(p, ksk, lwe, out);
    free(prep); free(at); free(acc); free(lwe);
    return e;
}

/* ------------------------------------------------------------------------- */
/* Plain AES-128 (FIPS-197), the decryption oracle of the transciphering path  */
/* (reference bitwise/examples/aes_transcipher/aes128.cuh:89-114, 227-333).    */
/* The S-box is computed from its definition: inverse in GF(2^8) modulo        */
/* x^8+x^4+x^3+x+1 followed by the affine map with constant 0x63.              */
/* ------------------------------------------------------------------------- */
static uint8_t gf_mul(uint8_t a, uint8_t b) {
    uint8_t r = 0;
    while (b) {
        if (b & 1) r ^= a;
        a = (uint8_t)((a << 1) ^ ((a & 0x80) ? 0x1b : 0));
        b >>= 1;
    }
    return r;
}

uint8_t orc_aes_sbox(uint8_t x) {
    uint8_t inv = 0;
    if (x) {
        /* x^254 = x^-1 */
        uint8_t r = 1, b = x;
        int e = 254;
        while (e) {
            if (e & 1) r = gf_mul(r, b);
            b = gf_mul(b, b);
            e >>= 1;
        }
        inv = r;
    }
    uint8_t s = inv;
    uint8_t res = 0x63;
    for (int i = 0; i < 5; ++i) {
        res ^= s;
        s = (uint8_t)((s << 1) | (s >> 7));
    }
    return res;
}

void orc_aes128_expand(const uint8_t key[16], uint8_t rk[176]) {
    memcpy(rk, key, 16);
    uint8_t rcon = 1;
    for (int i = 4; i < 44; ++i) {
        uint8_t t[4];
        memcpy(t, rk + 4 * (i - 1), 4);
        if (i % 4 == 0) {
            uint8_t t0 = t[0];
            t[0] = (uint8_t)(orc_aes_sbox(t[1]) ^ rcon);
            t[1] = orc_aes_sbox(t[2]);
            t[2] = orc_aes_sbox(t[3]);
            t[3] = orc_aes_sbox(t0);
            rcon = gf_mul(rcon, 2);
        }
        for (int k = 0; k < 4; ++k) rk[4 * i + k] = (uint8_t)(rk[4 * (i - 4) + k] ^ t[k]);
    }
}

void orc_aes128_encrypt(const uint8_t rk[176], const uint8_t in[16], uint8_t out[16]) {
    uint8_t s[16];
    for (int i = 0; i < 16; ++i) s[i] = (uint8_t)(in[i] ^ rk[i]);
    for (int round = 1; round <= 10; ++round) {
        uint8_t t[16];
        for (int i = 0; i < 16; ++i) s[i] = orc_aes_sbox(s[i]);
        /* ShiftRows: byte index = row + 4*col */
        for (int c = 0; c < 4; ++c)
            for (int r = 0; r < 4; ++r) t[r + 4 * c] = s[r + 4 * ((c + r) % 4)];
        if (round != 10) {
            for (int c = 0; c < 4; ++c) {
                uint8_t* a = t + 4 * c;
                uint8_t a0 = a[0], a1 = a[1], a2 = a[2], a3 = a[3];
                s[4 * c + 0] = (uint8_t)(gf_mul(a0, 2) ^ gf_mul(a1, 3) ^ a2 ^ a3);
                s[4 * c + 1] = (uint8_t)(a0 ^ gf_mul(a1, 2) ^ gf_mul(a2, 3) ^ a3);
                s[4 * c + 2] = (uint8_t)(a0 ^ a1 ^ gf_mul(a2, 2) ^ gf_mul(a3, 3));
                s[4 * c + 3] = (uint8_t)(gf_mul(a0, 3) ^ a1 ^ a2 ^ gf_mul(a3, 2));
            }
        } else {
            memcpy(s, t, 16);
        }
        for (int i = 0; i < 16; ++i) s[i] ^= rk[16 * round + i];
    }
    memcpy(out, s, 16);
}

/* ------------------------------------------------------------------------- */
/* AES-128 decryption and the T-tables of the reference's AES data flow        */
/* (aes128.cuh:227-276 BuildAesDecTables / BuildAesEncTables; the round-key     */
/* words of the Td path, key_material.cuh:12-47, rk_imc_words).                 */
/* ------------------------------------------------------------------------- */
static uint8_t rotl8(uint8_t x, int k) { return (uint8_t)((x << k) | (x >> (8 - k))); }

/* S^-1(y) = inverse in GF(2^8) of the inverse affine map of y: x = rotl(y', 1) ^
 * rotl(y', 3) ^ rotl(y', 6) with y' = y ^ 0x63 (FIPS-197 5.3.2). */
uint8_t orc_aes_inv_sbox(uint8_t y) {
    uint8_t v = (uint8_t)(y ^ 0x63);
    uint8_t x = (uint8_t)(rotl8(v, 1) ^ rotl8(v, 3) ^ rotl8(v, 6));
    if (!x) return 0;
    uint8_t r = 1, b = x;
    int e = 254;
    while (e) {
        if (e & 1) r = gf_mul(r, b);
        b = gf_mul(b, b);
        e >>= 1;
    }
    return r;
}

static uint32_t be32(uint8_t b0, uint8_t b1, uint8_t b2, uint8_t b3) {
    return ((uint32_t)b0 << 24) | ((uint32_t)b1 << 16) | ((uint32_t)b2 << 8) | b3;
}
static uint32_t rotr32(uint32_t x, int k) { return k ? (x >> k) | (x << (32 - k)) : x; }

/* T-table word: dec = 0: Te_r[x] = rotr(be(2S, S, S, 3S), 8r); dec = 1:
 * Td_r[x] = rotr(be(14 S^-1, 9 S^-1, 13 S^-1, 11 S^-1), 8r), S = S(x) / S^-1(x). */
uint32_t orc_aes_ttable(int dec, int r, uint8_t x) {
    uint32_t w;
    if (dec) {
        uint8_t s = orc_aes_inv_sbox(x);
        w = be32(gf_mul(s, 14), gf_mul(s, 9), gf_mul(s, 13), gf_mul(s, 11));
    } else {
        uint8_t s = orc_aes_sbox(x);
        w = be32(gf_mul(s, 2), s, s, gf_mul(s, 3));
    }
    return rotr32(w, 8 * r);
}

static void inv_mix_column(const uint8_t a[4], uint8_t o[4]) {
    o[0] = (uint8_t)(gf_mul(a[0], 14) ^ gf_mul(a[1], 11) ^ gf_mul(a[2], 13) ^ gf_mul(a[3], 9));
    o[1] = (uint8_t)(gf_mul(a[0], 9) ^ gf_mul(a[1], 14) ^ gf_mul(a[2], 11) ^ gf_mul(a[3], 13));
    o[2] = (uint8_t)(gf_mul(a[0], 13) ^ gf_mul(a[1], 9) ^ gf_mul(a[2], 14) ^ gf_mul(a[3], 11));
    o[3] = (uint8_t)(gf_mul(a[0], 11) ^ gf_mul(a[1], 13) ^ gf_mul(a[2], 9) ^ gf_mul(a[3], 14));
}

/* The straightforward inverse cipher (FIPS-197 5.3): AddRoundKey(10), then per round
 * InvShiftRows, InvSubBytes, AddRoundKey, InvMixColumns (rounds 9..1), final round
 * without InvMixColumns.  Byte index = row + 4 col. */
void orc_aes128_decrypt(const uint8_t rk[176], const uint8_t in[16], uint8_t out[16]) {
    uint8_t s[16], t[16];
    for (int i = 0; i < 16; ++i) s[i] = (uint8_t)(in[i] ^ rk[160 + i]);
    for (int round = 9; round >= 0; --round) {
        for (int c = 0; c < 4; ++c)
            for (int r = 0; r < 4; ++r) t[r + 4 * ((c + r) % 4)] = s[r + 4 * c];  /* InvShiftRows */
        for (int i = 0; i < 16; ++i) s[i] = (uint8_t)(orc_aes_inv_sbox(t[i]) ^ rk[16 * round + i]);
        if (round > 0)
            for (int c = 0; c < 4; ++c) {
                uint8_t o[4];
                inv_mix_column(s + 4 * c, o);
                memcpy(s + 4 * c, o, 4);
            }
    }
    memcpy(out, s, 16);
}

/* Decryption key schedule of the equivalent inverse cipher (FIPS-197 5.3.5):
 * dk[0] = rk[10], dk[r] = InvMixColumns(rk[10 - r]) for r = 1..9, dk[10] = rk[0]. */
void orc_aes128_dec_schedule(const uint8_t rk[176], uint8_t dk[176]) {
    memcpy(dk, rk + 160, 16);
    for (int r = 1; r <= 9; ++r)
        for (int c = 0; c < 4; ++c) inv_mix_column(rk + 16 * (10 - r) + 4 * c, dk + 16 * r + 4 * c);
    memcpy(dk + 160, rk, 16);
}

/* The equivalent inverse cipher on Td tables, as the reference's GPU rounds evaluate
 * it: s = in ^ dk[0]; rounds 1..9: column c of the new state = Td0[s(0, c)] ^
 * Td1[s(1, c - 1)] ^ Td2[s(2, c - 2)] ^ Td3[s(3, c - 3)] ^ dk[r] column c (big-endian
 * word: row 0 in the top byte); final round: S^-1, InvShiftRows, ^ dk[10]. */
void orc_aes128_decrypt_ttable(const uint8_t dk[176], const uint8_t in[16], uint8_t out[16]) {
    uint8_t s[16], t[16];
    for (int i = 0; i < 16; ++i) s[i] = (uint8_t)(in[i] ^ dk[i]);
    for (int round = 1; round <= 9; ++round) {
        for (int c = 0; c < 4; ++c) {
            uint32_t w = 0;
            for (int r = 0; r < 4; ++r) w ^= orc_aes_ttable(1, r, s[r + 4 * ((c - r + 4) % 4)]);
            for (int r = 0; r < 4; ++r) t[r + 4 * c] = (uint8_t)((w >> (24 - 8 * r)) ^ dk[16 * round + r + 4 * c]);
        }
        memcpy(s, t, 16);
    }
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r)
            t[r + 4 * c] = (uint8_t)(orc_aes_inv_sbox(s[r + 4 * ((c - r + 4) % 4)]) ^ dk[160 + r + 4 * c]);
    memcpy(out, t, 16);
}

/* The cipher on Te tables (reference GPU rounds of the CTR path): column c = Te0[s(0, c)]
 * ^ Te1[s(1, c + 1)] ^ Te2[s(2, c + 2)] ^ Te3[s(3, c + 3)] ^ rk[r] column c; final round
 * S, ShiftRows, ^ rk[10]. */
void orc_aes128_encrypt_ttable(const uint8_t rk[176], const uint8_t in[16], uint8_t out[16]) {
    uint8_t s[16], t[16];
    for (int i = 0; i < 16; ++i) s[i] = (uint8_t)(in[i] ^ rk[i]);
    for (int round = 1; round <= 9; ++round) {
        for (int c = 0; c < 4; ++c) {
            uint32_t w = 0;
            for (int r = 0; r < 4; ++r) w ^= orc_aes_ttable(0, r, s[r + 4 * ((c + r) % 4)]);
            for (int r = 0; r < 4; ++r) t[r + 4 * c] = (uint8_t)((w >> (24 - 8 * r)) ^ rk[16 * round + r + 4 * c]);
        }
        memcpy(s, t, 16);
    }
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) t[r + 4 * c] = (uint8_t)(orc_aes_sbox(s[r + 4 * ((c + r) % 4)]) ^ rk[160 + r + 4 * c]);
    memcpy(out, t, 16);
}
