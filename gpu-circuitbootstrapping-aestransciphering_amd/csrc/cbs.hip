// cbs.hip -- gfx950 kernels of the circuit-bootstrapping level (N = 2048):
// Fourier conversions, key generation, the multi-value programmable bootstrap
// (torus-64), private functional key switching to torus-32 RLWE, the torus-32
// CMux vertical-packing LUT and the AES linear layers.  A product of a digit
// polynomial with a torus-64 polynomial is either two FP64 products (one per
// signed 32-bit limb), rounded and recombined mod 2^64, which is exact, or (the
// one-product PBS form) one FP64 product of the polynomial scaled to the torus,
// rounded once.  With a torus-32 polynomial it is one FP64 product, exact.
#include "fft1024.hpp"
#include "kernels.hpp"

namespace tfhe {

constexpr int CB_N = 2048, CB_M = 1024;
constexpr int PBS_LOGB = 9, PBS_L = 4, PBS_ROWS = 8, PBS_IGN = 28;  // PBS gadget (see pbs_digit)

// ---------------------------------------------------------------------------
// shared helpers
// ---------------------------------------------------------------------------
// p1t, p2, p3 staged in LDS (768 cplx); the per-wave cross twiddles tr[w] live
// in registers (4 cplx per lane).
struct StagedTables {
    FFT1024Tables T;
};

__device__ __forceinline__ FFT1024Tables stage_tables(const FFT1024Tables& Tg, cplx* tl) {
    for (int k = threadIdx.x; k < 256; k += blockDim.x) {
        tl[k] = Tg.p1t[k];
        tl[256 + k] = Tg.p2[k];
        tl[512 + k] = Tg.p3[k];
    }
    __syncthreads();
    return {tl, tl + 256, tl + 512, Tg.tr, Tg.rot};
}

__device__ __forceinline__ void split_limbs(uint64_t v, double& lo, double& hi) {
    const int32_t l = (int32_t)(uint32_t)v;
    lo = (double)l;
    hi = (double)(int32_t)(uint32_t)((v - (uint64_t)(int64_t)l) >> 32);
}

// ---------------------------------------------------------------------------
// Fourier conversion: u64 polynomials -> [poly][limb][q][s][lane] (1024 cplx per limb)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_to_fourier2048(const uint64_t* __restrict__ coef, size_t npoly,
                                                       FFT1024Tables Tg, cplx* __restrict__ out) {
    __shared__ cplx tl[768];
    __shared__ cplx lds_all[4][F256_LDS_CPLX];
    __shared__ cplx xbuf[2][1024];
    const FFT1024Tables T = stage_tables(Tg, tl);
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    cplx trw[4];
    load_trw(Tg, w, lane, trw);
    int buf = 0;
    for (size_t q = blockIdx.x; q < npoly; q += gridDim.x) {
        const uint64_t* p = coef + q * CB_N;
        cplx x0[4], x1[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const int c = own_coeff(w, lane, a);
            double l0r, l1r, l0i, l1i;
            split_limbs(p[c], l0r, l1r);
            split_limbs(p[c + CB_M], l0i, l1i);
            x0[a] = {l0r, l0i};
            x1[a] = {l1r, l1i};
        }
        f1024_fwd<false>(x0, lds_all[w], xbuf[buf], T, trw, w, lane);
        buf ^= 1;
        f1024_fwd<false>(x1, lds_all[w], xbuf[buf], T, trw, w, lane);
        buf ^= 1;
        cplx* o = out + q * 2 * CB_M;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            o[(w * 4 + s) * 64 + lane] = x0[s];
            o[CB_M + (w * 4 + s) * 64 + lane] = x1[s];
        }
    }
}

// Fourier conversion of torus-32 polynomials: u32 -> [poly][q][s][lane] (1024 cplx)
__global__ __launch_bounds__(256) void k_to_fourier2048_32(const uint32_t* __restrict__ coef, size_t npoly,
                                                          FFT1024Tables Tg, cplx* __restrict__ out) {
    __shared__ cplx tl[768];
    __shared__ cplx lds_all[4][F256_LDS_CPLX];
    __shared__ cplx xbuf[2][1024];
    const FFT1024Tables T = stage_tables(Tg, tl);
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    cplx trw[4];
    load_trw(Tg, w, lane, trw);
    int buf = 0;
    for (size_t q = blockIdx.x; q < npoly; q += gridDim.x) {
        const uint32_t* p = coef + q * CB_N;
        cplx x[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const int c = own_coeff(w, lane, a);
            x[a] = {(double)(int32_t)p[c], (double)(int32_t)p[c + CB_M]};
        }
        f1024_fwd<false>(x, lds_all[w], xbuf[buf], T, trw, w, lane);
        buf ^= 1;
        cplx* o = out + q * CB_M;
#pragma unroll
        for (int s = 0; s < 4; ++s) o[(w * 4 + s) * 64 + lane] = x[s];
    }
}

// Fourier conversion of torus-64 polynomials in one limb: coefficient v (signed
// 64-bit) becomes the double v * 2^-64 / M, so that an inverse transform of a
// product with a digit polynomial lands directly in torus units (the 1/M of the
// inverse folded in).  Output [poly][q][s][lane] (1024 cplx).
__global__ __launch_bounds__(256) void k_to_fourier2048_t64(const uint64_t* __restrict__ coef, size_t npoly,
                                                           FFT1024Tables Tg, cplx* __restrict__ out) {
    __shared__ cplx tl[768];
    __shared__ cplx lds_all[4][F256_LDS_CPLX];
    __shared__ cplx xbuf[2][1024];
    const FFT1024Tables T = stage_tables(Tg, tl);
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    constexpr double SCALE = 0x1p-74;  // 2^-64 / M
    cplx trw[4];
    load_trw(Tg, w, lane, trw);
    int buf = 0;
    for (size_t q = blockIdx.x; q < npoly; q += gridDim.x) {
        const uint64_t* p = coef + q * CB_N;
        cplx x[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const int c = own_coeff(w, lane, a);
            x[a] = {(double)(int64_t)p[c] * SCALE, (double)(int64_t)p[c + CB_M] * SCALE};
        }
        f1024_fwd<false>(x, lds_all[w], xbuf[buf], T, trw, w, lane);
        buf ^= 1;
        cplx* o = out + q * CB_M;
#pragma unroll
        for (int s = 0; s < 4; ++s) o[(w * 4 + s) * 64 + lane] = x[s];
    }
}

// Negacyclic product of small-integer x by torus-64 y (N = 2048), one block per
// product.  LIMBS = 2: exact (two signed 32-bit limbs of y, each product rounded
// exactly).  LIMBS = 1: the one-product form of the PBS (y scaled to the torus,
// one FP64 product, rounded once with torus64_from_double); its difference from
// the exact product is the per-product rounding error of that form.
__device__ __forceinline__ uint64_t torus64_from_double(double y);

template <int LIMBS>
__global__ __launch_bounds__(256) void k_polymul2048(const int32_t* __restrict__ x, const uint64_t* __restrict__ y,
                                                    uint32_t batch, FFT1024Tables Tg, uint64_t* __restrict__ out) {
    __shared__ cplx tl[768];
    __shared__ cplx lds_all[4][F256_LDS_CPLX];
    __shared__ cplx xbuf[2][1024];
    const FFT1024Tables T = stage_tables(Tg, tl);
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    constexpr double SCALE = 0x1p-74;  // 2^-64 / M
    cplx trw[4];
    load_trw(Tg, w, lane, trw);
    int buf = 0;
    for (uint32_t q = blockIdx.x; q < batch; q += gridDim.x) {
        const int32_t* xp = x + (size_t)q * CB_N;
        const uint64_t* yp = y + (size_t)q * CB_N;
        cplx X[4], Y0[4], Y1[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const int c = own_coeff(w, lane, a);
            X[a] = {(double)xp[c], (double)xp[c + CB_M]};
            if constexpr (LIMBS == 2) {
                double l0r, l1r, l0i, l1i;
                split_limbs(yp[c], l0r, l1r);
                split_limbs(yp[c + CB_M], l0i, l1i);
                Y0[a] = {l0r, l0i};
                Y1[a] = {l1r, l1i};
            } else {
                Y0[a] = {(double)(int64_t)yp[c] * SCALE, (double)(int64_t)yp[c + CB_M] * SCALE};
            }
        }
        f1024_fwd<false>(X, lds_all[w], xbuf[buf], T, trw, w, lane); buf ^= 1;
        f1024_fwd<false>(Y0, lds_all[w], xbuf[buf], T, trw, w, lane); buf ^= 1;
        if constexpr (LIMBS == 2) {
            f1024_fwd<false>(Y1, lds_all[w], xbuf[buf], T, trw, w, lane);
            buf ^= 1;
        }
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            Y0[s] = cmul(Y0[s], X[s]);
            if constexpr (LIMBS == 2) Y1[s] = cmul(Y1[s], X[s]);
        }
        f1024_inv<false>(Y0, lds_all[w], xbuf[buf], T, trw, w, lane); buf ^= 1;
        if constexpr (LIMBS == 2) {
            f1024_inv<false>(Y1, lds_all[w], xbuf[buf], T, trw, w, lane);
            buf ^= 1;
        }
        uint64_t* op = out + (size_t)q * CB_N;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const int c = own_coeff(w, lane, a);
            if constexpr (LIMBS == 2) {
                op[c] = (uint64_t)round_i64(Y0[a].re, 1.0 / CB_M) + ((uint64_t)round_i64(Y1[a].re, 1.0 / CB_M) << 32);
                op[c + CB_M] =
                    (uint64_t)round_i64(Y0[a].im, 1.0 / CB_M) + ((uint64_t)round_i64(Y1[a].im, 1.0 / CB_M) << 32);
            } else {
                op[c] = torus64_from_double(Y0[a].re);
                op[c + CB_M] = torus64_from_double(Y0[a].im);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// BSK2 generation over Z/2^64: one block per RLWE row; a * K2 is exact through
// two limb products with the precomputed Fourier transform of K2.
// Row id = i*8 + r; msg = s_i * 2^(28 + 9j) on component u of coefficient 0
// (r = 2j + u).  Output row layout [2][N] (a then b).
// With mb = 2 (two key bits per blind-rotation step) RGSW q = 3 ip + pat encrypts
// the indicator of pattern pat of the key-bit pair (s_2ip, s_2ip+1): (1,0), (0,1),
// (1,1) for pat = 0, 1, 2; Philox domain DOM_BSK2MB, row id q*8 + r.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t bsk_message_bit(const int32_t* s, uint32_t q, int mb) {
    if (mb == 1) return (uint32_t)s[q];
    const uint32_t ip = q / 3, pat = q % 3;
    const uint32_t b0 = (uint32_t)s[2 * ip], b1 = (uint32_t)s[2 * ip + 1];
    return pat == 0 ? (b0 & (1u - b1)) : pat == 1 ? ((1u - b0) & b1) : (b0 & b1);
}

// What row `row` of a generated torus-64 key encrypts (RLWE_K2 rows [2][N], Philox ids = row):
//   msgpoly == NULL (BSK2): row = i * 8 + r, r = 2j + u, adds bsk_message_bit(s, i, mb) *
//     2^(64 - 4 logB + j logB) to component u at coefficient 0 (domain DOM_BSK2 or DOM_BSK2MB);
//   msgpoly != NULL (key-switching keys of HomTrace / SchemeSwitch): row = q * KS_L + d adds
//     msgpoly[q][c] * 2^(64 - KS_L KS_LOGB + d KS_LOGB) to the body (domain dom).
struct KeyRows {
    const int32_t* s = nullptr;
    int mb = 1, logB = 9;
    const int64_t* msgpoly = nullptr;
    uint32_t dom = 0;
};
constexpr int KS_LOGB = 8, KS_L = 5;  // HomTrace / SchemeSwitch gadget 2^8 x 5 on the 64-bit torus

__global__ __launch_bounds__(256) void k_gen_bsk64(uint64_t seed, uint32_t nrows, int nshift, KeyRows kr,
                                                  const cplx* __restrict__ K2f, FFT1024Tables Tg,
                                                  uint64_t* __restrict__ out) {
    __shared__ cplx tl[768];
    __shared__ cplx lds_all[4][F256_LDS_CPLX];
    __shared__ cplx xbuf[2][1024];
    const FFT1024Tables T = stage_tables(Tg, tl);
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    cplx trw[4], kf[4];
    load_trw(Tg, w, lane, trw);
#pragma unroll
    for (int sl = 0; sl < 4; ++sl) kf[sl] = K2f[(w * 4 + sl) * 64 + lane];
    int buf = 0;
    for (uint32_t row = blockIdx.x; row < nrows; row += gridDim.x) {
        const bool ks = kr.msgpoly != nullptr;
        const int r = row % 8, j = r >> 1;
        const int msg_comp = r & 1;
        const uint64_t cval =
            ks ? 0 : (uint64_t)bsk_message_bit(kr.s, row / 8, kr.mb) << (64 - PBS_L * kr.logB + j * kr.logB);
        const uint64_t* mp = ks ? reinterpret_cast<const uint64_t*>(kr.msgpoly) + (size_t)(row / KS_L) * CB_N : nullptr;
        const int gsh = 64 - KS_L * KS_LOGB + (int)(row % KS_L) * KS_LOGB;
        const uint32_t id = row;
        const uint32_t dom = ks ? kr.dom : kr.mb == 1 ? DOM_BSK2 : DOM_BSK2MB;
        uint64_t av[8];
        cplx A0[4], A1[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const int c = own_coeff(w, lane, a);
            av[a] = rand64(seed, dom, id, (uint32_t)c);
            av[a + 4] = rand64(seed, dom, id, (uint32_t)(c + CB_M));
            double l0r, l1r, l0i, l1i;
            split_limbs(av[a], l0r, l1r);
            split_limbs(av[a + 4], l0i, l1i);
            A0[a] = {l0r, l0i};
            A1[a] = {l1r, l1i};
        }
        f1024_fwd<false>(A0, lds_all[w], xbuf[buf], T, trw, w, lane); buf ^= 1;
        f1024_fwd<false>(A1, lds_all[w], xbuf[buf], T, trw, w, lane); buf ^= 1;
#pragma unroll
        for (int sl = 0; sl < 4; ++sl) {
            A0[sl] = cmul(A0[sl], kf[sl]);
            A1[sl] = cmul(A1[sl], kf[sl]);
        }
        f1024_inv<false>(A0, lds_all[w], xbuf[buf], T, trw, w, lane); buf ^= 1;
        f1024_inv<false>(A1, lds_all[w], xbuf[buf], T, trw, w, lane); buf ^= 1;
        uint64_t* oa = out + (size_t)row * 2 * CB_N;
        uint64_t* ob = oa + CB_N;
#pragma unroll
        for (int h = 0; h < 8; ++h) {
            const int a = h & 3;
            const int c = own_coeff(w, lane, a) + (h >= 4 ? CB_M : 0);
            const double re0 = h < 4 ? A0[a].re : A0[a].im, re1 = h < 4 ? A1[a].re : A1[a].im;
            uint64_t b = (uint64_t)round_i64(re0, 1.0 / CB_M) + ((uint64_t)round_i64(re1, 1.0 / CB_M) << 32);
            b += (uint64_t)noise(seed, dom, id, (uint32_t)c, nshift);
            uint64_t aval = av[h];
            if (ks) {
                b += mp[c] << gsh;
            } else if (c == 0) {
                if (msg_comp == 0) aval += cval; else b += cval;
            }
            oa[c] = aval;
            ob[c] = b;
        }
    }
}

// ---------------------------------------------------------------------------
// PFKS key generation over Z/2^32 (one block per RLWE row): row index
// ((u*(N+1) + i)*8 + t)*15 + (v-1), v = 1..15, Philox id ((u*(N+1) + i)*8 + t)*16 + v;
// a = rand32, b = a*K2 + noise + msg_b with msg_b = f_u(v * kappa_i * 2^(4t)),
// kappa_N = -1, f_1(x) = x (coefficient 0), f_0(x) = -K2(X) x.  a*K2 is one
// exact FP64 product (|a| < 2^31, K2 binary).  Output row layout [2][N] u32.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_gen_pfks32(uint64_t seed, uint32_t nrows, int nshift,
                                                   const int32_t* __restrict__ K2, const cplx* __restrict__ K2f,
                                                   FFT1024Tables Tg, uint32_t* __restrict__ out) {
    __shared__ cplx tl[768];
    __shared__ cplx lds_all[4][F256_LDS_CPLX];
    __shared__ cplx xbuf[2][1024];
    const FFT1024Tables T = stage_tables(Tg, tl);
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    cplx trw[4], kf[4];
    load_trw(Tg, w, lane, trw);
#pragma unroll
    for (int sl = 0; sl < 4; ++sl) kf[sl] = K2f[(w * 4 + sl) * 64 + lane];
    int buf = 0;
    for (uint32_t row = blockIdx.x; row < nrows; row += gridDim.x) {
        const uint32_t v1 = row % 15, rest = row / 15;  // v = v1 + 1
        const int t = rest % 8;
        const uint32_t rest2 = rest / 8;
        const int i = rest2 % (CB_N + 1);
        const int u = rest2 / (CB_N + 1);
        const uint32_t kappa = i < CB_N ? (uint32_t)K2[i] : 0xFFFFFFFFu;
        const uint32_t cval = ((v1 + 1) * kappa) << (4 * t);
        const uint32_t id = ((rest2 * 8) + t) * 16 + (v1 + 1);
        uint32_t av[8];
        cplx A[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const int c = own_coeff(w, lane, a);
            av[a] = rand32(seed, DOM_PFKS, id, (uint32_t)c);
            av[a + 4] = rand32(seed, DOM_PFKS, id, (uint32_t)(c + CB_M));
            A[a] = {(double)(int32_t)av[a], (double)(int32_t)av[a + 4]};
        }
        f1024_fwd<false>(A, lds_all[w], xbuf[buf], T, trw, w, lane); buf ^= 1;
#pragma unroll
        for (int sl = 0; sl < 4; ++sl) A[sl] = cmul(A[sl], kf[sl]);
        f1024_inv<false>(A, lds_all[w], xbuf[buf], T, trw, w, lane); buf ^= 1;
        uint32_t* oa = out + (size_t)row * 2 * CB_N;
        uint32_t* ob = oa + CB_N;
#pragma unroll
        for (int h = 0; h < 8; ++h) {
            const int a = h & 3;
            const int c = own_coeff(w, lane, a) + (h >= 4 ? CB_M : 0);
            uint32_t b = (uint32_t)round_i64(h < 4 ? A[a].re : A[a].im, 1.0 / CB_M);
            b += (uint32_t)noise(seed, DOM_PFKS, id, (uint32_t)c, nshift);
            if (u == 0) b -= (uint32_t)K2[c] * cval;
            else if (c == 0) b += cval;
            oa[c] = av[h];
            ob[c] = b;
        }
    }
}

// ===========================================================================
// Multi-value programmable bootstrap (the circuit bootstrap's blind rotation):
// one 4-wave workgroup per ciphertext, accumulator in registers in the
// FFT1024 coefficient ownership.  Input LWE_s (torus-32) of b * 2^31.
// Steps (reference ManyLUT path, cirbts-batch.cu:3293-4006, restated for the
// torus-64 CGGI accumulator): b += 2^30; a~_i = special_ms(a_i, bw = 1);
// test vector with b~ folded in; n CMux steps acc += (X^m - 1)(dec(acc) (*) BSK_i),
// m = -a~_i mod 2N (or n/2 steps over key-bit pairs, MB = 2); extract
// coefficients 0 and 1 and add mu_j = 2^(47 + 8j).
// Output [batch][2][N+1] u64: LWE_K2 of b * 2^(48 + 8j).
// ===========================================================================
// PBS gadget: PBS_L = 4 levels of 2^LOGB (LOGB = 9 for every PFKS form, 10 for the
// HomTrace + SchemeSwitch form), the low 64 - 4 LOGB bits dropped with rounding

// signed digits by the offset method on 64 bits: off = 2^(IGN-1) + sum_j (B/2) 2^(IGN + j logB)
template <int LOGB, int L>
__device__ __forceinline__ uint64_t digit_offset64() {
    constexpr int IGN = 64 - L * LOGB;
    uint64_t off = 1ull << (IGN - 1);
#pragma unroll
    for (int j = 0; j < L; ++j) off += (1ull << (LOGB - 1)) << (IGN + j * LOGB);
    return off;
}
template <int LOGB, int L>
__device__ __forceinline__ double digit64(uint64_t tp, int j) {
    constexpr int IGN = 64 - L * LOGB;
    return (double)((int32_t)((tp >> (IGN + j * LOGB)) & ((1u << LOGB) - 1)) - (1 << (LOGB - 1)));
}

template <int LOGB = PBS_LOGB>
__device__ __forceinline__ uint64_t pbs_digit_offset() {
    return digit_offset64<LOGB, PBS_L>();
}

template <int LOGB = PBS_LOGB>
__device__ __forceinline__ double pbs_digit(uint64_t tp, int j) {
    return digit64<LOGB, PBS_L>(tp, j);
}

// round(v * 2N / (2^32 2^bw)) * 2^bw mod 2N with N = 2048, bw = 1
__device__ __forceinline__ uint32_t special_ms_2048(uint32_t v) {
    return (uint32_t)((((uint64_t)v + (1ull << 20)) >> 21) << 1) & 4095u;
}
// the same for any bw in [0, 11] (the reference's gpu_SpecilMS, cirbts-utils.cu:52-61)
__device__ __forceinline__ uint32_t special_ms_2048_bw(uint32_t v, uint32_t bw) {
    return (uint32_t)((((uint64_t)v + (1ull << (19 + bw))) >> (20 + bw)) << bw) & 4095u;
}

// round(y * 2^64) mod 2^64 for a torus value held as a double.  f = y - rint(y)
// in [-1/2, 1/2] is exact; its top 32 bits are rint(f 2^32), taken with the
// 1.5 * 2^52 rounding constant, and the exact residual r = f 2^32 - rint(f 2^32)
// gives the low 32 bits the same way (borrowing one from the top half when
// rint(r 2^32) < 0, i.e. l < MAGIC).
__device__ __forceinline__ uint64_t torus64_from_double(double y) {
    constexpr double MAGIC = 6755399441055744.0, TWO32 = 4294967296.0;
    const double f = y - rint(y);
    const double h = fma(f, TWO32, MAGIC);
    const double r = fma(f, TWO32, -(h - MAGIC));
    const double l = fma(r, TWO32, MAGIC);
    const uint32_t hi = (uint32_t)__double_as_longlong(h) - (l < MAGIC ? 1u : 0u);
    return ((uint64_t)hi << 32) | (uint32_t)__double_as_longlong(l);
}

// LIMBS = 2: BSK2 Fourier rows hold the two signed 32-bit limbs of each torus-64
// coefficient and every product is rounded exactly; the monic update rotates the
// rounded product across the four waves through LDS.  LIMBS = 1: one Fourier row
// of the coefficients scaled to the torus (k_to_fourier2048_t64); the monic factor
// is applied pointwise before the inverse transform and the product is rounded
// once, to about 2^-36 of the torus per step.
// MB = 2 (one limb only): two key bits per step.  With m1, m2 the rotations of the
// pair, acc += sum_p (X^(e_p) - 1) (dec(acc) (*) RGSW(ind_p)) over the patterns
// (1,0), (0,1), (1,1) with e_p = m1, m2, m1 + m2: exactly one indicator is 1 (or
// none), so this equals the two single-bit steps.  One decomposition and two
// inverse transforms serve both bits; the factor of pattern p is applied to the
// digit transforms before they meet that pattern's key row.  The pattern keys are
// the step's largest load stream (24 complex values per lane and digit row), and the
// step waits on it unless many loads are in flight: WIDE = false keeps two waves per
// SIMD and issues each pattern's 8 values as one batch, the next pattern's before the
// current pattern's MACs; WIDE = true (launches of at most one workgroup per CU)
// issues all 24 at once with one wave per SIMD.
template <bool BF, int LIMBS, int MB = 1, bool WIDE = false, bool LUT = false, int LOGB = PBS_LOGB>
__global__ __launch_bounds__(256, (MB == 2 && !WIDE) ? 2 : 1) void k_pbs2_manylut(const uint32_t* __restrict__ lwe, uint32_t batch, int n,
                                                     const cplx* __restrict__ bskf, FFT1024Tables Tg,
                                                     uint64_t* __restrict__ out, PbsLut io) {
    __shared__ cplx tl[768];
    __shared__ cplx lds_all[4][F256_LDS_CPLX];
    __shared__ cplx xbuf[2][1024];
    __shared__ uint16_t at[1024];
    const FFT1024Tables T = stage_tables(Tg, tl);
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    cplx trw[4];
    load_trw(Tg, w, lane, trw);
    static_assert(MB == 1 || LIMBS == 1, "two key bits per step only in the one-product form");
    static_assert(!WIDE || MB == 2, "the wide key batch is a form of the two-bit step");
    constexpr bool RT = LIMBS == 1 && MB == 1;  // registers to spare for the p2 / p3 twiddles
    const Tw256Regs tw = RT ? load_tw256(Tg, lane) : Tw256Regs{};
    cplx* lds = lds_all[w];
    uint64_t* x64 = reinterpret_cast<uint64_t*>(&xbuf[0][0]);  // [2][2048] u64 scratch for the monic exchange
    const uint64_t OFF = pbs_digit_offset<LOGB>();
    static_assert(LOGB == PBS_LOGB || (LUT && LIMBS == 2 && MB == 1), "the 2^10 gadget: exact LUT-mode form only");
    constexpr uint32_t N = CB_N, Q2 = 2 * CB_N;
    const uint64_t mu0 = 1ull << 47, mu1 = 1ull << 55;
    uint32_t ek[4];  // evaluation exponents 1 - 4k (mod 2N) of this lane's frequencies
    const uint32_t kq = (uint32_t)((lane >> 4) + 4 * ((lane >> 2) & 3) + 16 * (lane & 3) + 256 * w);
#pragma unroll
    for (int s = 0; s < 4; ++s) ek[s] = (1u - 4u * (kq + 64u * (uint32_t)s)) & (Q2 - 1);

    // MB = 2, two waves per SIMD: the 64 resident workgroups of an XCD each touch one share of
    // the next step's pattern-key lines after a step's last key load (see k_pbs2_split)
    constexpr uint32_t KLINES = 3 * PBS_ROWS * 2 * CB_M * sizeof(cplx) / 128;
    const uint32_t per_xcd = (gridDim.x + 7) / 8 < 64 ? (gridDim.x + 7) / 8 : 64;
    const uint32_t share = (KLINES + per_xcd - 1) / per_xcd, part = (blockIdx.x >> 3) % per_xcd;
    uint32_t pf0 = 0, sink = 0;
    for (uint32_t cidx = blockIdx.x; cidx < batch; cidx += gridDim.x) {
        const uint32_t* ct = lwe + (size_t)cidx * (n + 1);
        __syncthreads();
        uint32_t bt;
        if constexpr (LUT) {
            // rotation by X^{+a~_i s_i} (the step multiplies by X^{-at[i]}), so that acc ends
            // as LUT X^{-phase~}
            for (int k = threadIdx.x; k < n; k += blockDim.x)
                at[k] = (uint16_t)((Q2 - special_ms_2048_bw(ct[k], io.bw)) & (Q2 - 1));
            bt = special_ms_2048_bw(ct[n] + io.b_add, io.bw);
        } else {
            for (int k = threadIdx.x; k < n; k += blockDim.x) at[k] = (uint16_t)special_ms_2048(ct[k]);
            bt = special_ms_2048(ct[n] + (1u << 30));
        }
        __syncthreads();

        uint64_t acc[2][8];
        if constexpr (LUT) {  // acc = (0, LUT X^-bt): coefficient c is LUT[m], m = c + bt mod 2N, negated for m >= N
#pragma unroll
            for (int h = 0; h < 8; ++h) {
                const uint32_t c = (uint32_t)own_coeff(w, lane, h & 3) + (h >= 4 ? CB_M : 0);
                const uint32_t m = (c + bt) & (Q2 - 1);
                const uint64_t v = io.lut[m & (N - 1)];
                acc[0][h] = 0;
                acc[1][h] = m >= N ? 0 - v : v;
            }
        } else {
#pragma unroll
            for (int h = 0; h < 8; ++h) {
                const uint32_t c = (uint32_t)own_coeff(w, lane, h & 3) + (h >= 4 ? CB_M : 0);
                const uint32_t arg = (bt + Q2 - (c & ~1u)) & (Q2 - 1);
                uint64_t v = 0 - ((c & 1) ? mu1 : mu0);
                if (arg >= N) v = 0 - v;
                acc[0][h] = 0;
                acc[1][h] = v;
            }
        }
        int buf = 0;
        for (int i = 0; i < n; i += MB) {
            const uint32_t mrot = (Q2 - (uint32_t)at[i]) & (Q2 - 1);
            uint32_t mrot2 = 0;
            if constexpr (MB == 2) mrot2 = (Q2 - (uint32_t)at[i + 1]) & (Q2 - 1);
            if ((mrot | mrot2) == 0) continue;
            const cplx* key = bskf + (size_t)(i / MB) * (MB == 2 ? 3 : 1) * PBS_ROWS * 2 * LIMBS * CB_M;
            // MB = 2: z1 = w^(m1 (1 - 4k)), z2 = w^(m2 (1 - 4k)); the pattern factors
            // z1 - 1, z1 z2 - 1 and z2 - 1 (patterns (1,0), (1,1), (0,1)), formed once per step
            cplx fp[MB == 2 ? 3 : 1][4];
            if constexpr (MB == 2) {
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const cplx z1 = T.rot[(mrot * ek[s]) & (Q2 - 1)];
                    const cplx z2 = T.rot[(mrot2 * ek[s]) & (Q2 - 1)];
                    const cplx z12 = cmul(z1, z2);
                    fp[0][s] = {z1.re - 1.0, z1.im};
                    fp[1][s] = {z12.re - 1.0, z12.im};
                    fp[2][s] = {z2.re - 1.0, z2.im};
                }
            }
            cplx macc[2][LIMBS][4];
#pragma unroll
            for (int o = 0; o < 2; ++o)
#pragma unroll
                for (int lb = 0; lb < LIMBS; ++lb)
#pragma unroll
                    for (int s = 0; s < 4; ++s) macc[o][lb][s] = {0.0, 0.0};
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                uint64_t tp[8];
#pragma unroll
                for (int h = 0; h < 8; ++h) tp[h] = acc[u][h] + OFF;
#pragma unroll
                for (int j = 0; j < PBS_L; ++j) {
                    const int r = 2 * j + u;
                    cplx x[4];
#pragma unroll
                    for (int a = 0; a < 4; ++a) x[a] = {pbs_digit<LOGB>(tp[a], j), pbs_digit<LOGB>(tp[a + 4], j)};
                    f1024_fwd<BF, RT>(x, lds, xbuf[buf], T, trw, w, lane, tw);
                    buf ^= 1;
                    if constexpr (MB == 2) {
                        cplx y[4];
                        // key values of pattern q (rows in the order (1,0), (1,1), (0,1)) for
                        // both outputs, issued as batches, which the compiler would otherwise
                        // split, waiting on each load in turn.
                        auto load_keys = [&](cplx(&kv)[2][4], int q) {
                            const int pt = q == 0 ? 0 : q == 1 ? 2 : 1;
                            const cplx* kr = key + ((size_t)pt * PBS_ROWS + r) * 2 * CB_M + (w * 4) * 64 + lane;
#pragma unroll
                            for (int o = 0; o < 2; ++o)
#pragma unroll
                                for (int s = 0; s < 4; ++s) kv[o][s] = kr[o * CB_M + s * 64];
                        };
                        // WIDE (one wave per SIMD, registers to spare): a scheduling fence keeps a
                        // batch's loads together and leaves their waits at the first use, so later
                        // batches are in flight meanwhile.  Two waves per SIMD: the empty asm
                        // statements wait for the batch at once (both batches in flight at 256
                        // registers would spill, which measured slower: 2176 inputs 50-51 ms
                        // against 38)
                        auto close_batch = [](cplx(&kv)[2][4]) {
                            if constexpr (WIDE) {
                                __builtin_amdgcn_sched_barrier(0);
                            } else {
#pragma unroll
                                for (int o = 0; o < 2; ++o)
#pragma unroll
                                    for (int k = 0; k < 4; ++k)
                                        asm volatile("" : "+v"(kv[o][k].re), "+v"(kv[o][k].im));
                            }
                        };
                        auto mac_pattern = [&](const cplx(&kv)[2][4], int q) {
#pragma unroll
                            for (int s = 0; s < 4; ++s) y[s] = cmul(x[s], fp[q][s]);
#pragma unroll
                            for (int o = 0; o < 2; ++o)
#pragma unroll
                                for (int s = 0; s < 4; ++s) cmac(macc[o][0][s], y[s], kv[o][s]);
                        };
                        if constexpr (WIDE) {  // all 24 values as one batch
                            cplx kv[3][2][4];
#pragma unroll
                            for (int q = 0; q < 3; ++q) load_keys(kv[q], q);
#pragma unroll
                            for (int q = 0; q < 3; ++q) close_batch(kv[q]);
#pragma unroll
                            for (int q = 0; q < 3; ++q) mac_pattern(kv[q], q);
                        } else {  // 8 per batch, pattern q + 1 loaded before pattern q's MACs
                            cplx ka[2][4], kb[2][4];
                            load_keys(ka, 0);
                            close_batch(ka);
                            load_keys(kb, 1);
                            close_batch(kb);
                            mac_pattern(ka, 0);
                            load_keys(ka, 2);
                            close_batch(ka);
                            mac_pattern(kb, 1);
                            mac_pattern(ka, 2);
                        }
                    } else {
                        const cplx* kr = key + (size_t)r * 2 * LIMBS * CB_M + (w * 4) * 64 + lane;
#pragma unroll
                        for (int o = 0; o < 2; ++o)
#pragma unroll
                            for (int lb = 0; lb < LIMBS; ++lb)
#pragma unroll
                                for (int s = 0; s < 4; ++s)
                                    cmac(macc[o][lb][s], x[s], kr[(o * LIMBS + lb) * CB_M + s * 64]);
                    }
                }
            }
            if constexpr (MB == 2 && !WIDE) {
                sink ^= pf0;
                const char* nk = reinterpret_cast<const char*>(i + 2 < n ? key + 3 * PBS_ROWS * 2 * CB_M : key);
                const uint32_t end = (part + 1) * share < KLINES ? (part + 1) * share : KLINES;
                const uint32_t l0 = part * share + threadIdx.x;
                pf0 = *reinterpret_cast<const uint32_t*>(nk + (size_t)(l0 < end ? l0 : end - 1) * 128);
            }
            if constexpr (LIMBS == 1) {
                // acc += (X^mrot - 1) t in the Fourier domain: frequency k is the polynomial
                // at w^(1 - 4k), so t's transform is scaled by w^(mrot (1 - 4k)) - 1 before
                // the inverse (the product at most doubles; one rounding per coefficient)
                cplx rf[4];
                if constexpr (MB == 1) {
#pragma unroll
                    for (int s = 0; s < 4; ++s) {
                        const cplx z = T.rot[(mrot * ek[s]) & (Q2 - 1)];
                        rf[s] = {z.re - 1.0, z.im};
                    }
                }
#pragma unroll
                for (int o = 0; o < 2; ++o) {
                    cplx y0[4];
#pragma unroll
                    for (int s = 0; s < 4; ++s) y0[s] = MB == 1 ? cmul(macc[o][0][s], rf[s]) : macc[o][0][s];
                    f1024_inv<BF, RT>(y0, lds, xbuf[buf], T, trw, w, lane, tw);
                    buf ^= 1;
#pragma unroll
                    for (int a = 0; a < 4; ++a) {
                        acc[o][a] += torus64_from_double(y0[a].re);
                        acc[o][a + 4] += torus64_from_double(y0[a].im);
                    }
                }
            } else {
                uint64_t t[2][8];
#pragma unroll
                for (int o = 0; o < 2; ++o) {
                    cplx y0[4], y1[4];
#pragma unroll
                    for (int s = 0; s < 4; ++s) { y0[s] = macc[o][0][s]; y1[s] = macc[o][1][s]; }
                    f1024_inv<BF>(y0, lds, xbuf[buf], T, trw, w, lane);
                    buf ^= 1;
                    f1024_inv<BF>(y1, lds, xbuf[buf], T, trw, w, lane);
                    buf ^= 1;
#pragma unroll
                    for (int a = 0; a < 4; ++a) {
                        t[o][a] = (uint64_t)round_i64(y0[a].re, 1.0 / CB_M) +
                                  ((uint64_t)round_i64(y1[a].re, 1.0 / CB_M) << 32);
                        t[o][a + 4] = (uint64_t)round_i64(y0[a].im, 1.0 / CB_M) +
                                      ((uint64_t)round_i64(y1[a].im, 1.0 / CB_M) << 32);
                    }
                }
                // acc += X^mrot t - t, rotation across the four waves through LDS
                __syncthreads();
#pragma unroll
                for (int o = 0; o < 2; ++o)
#pragma unroll
                    for (int h = 0; h < 8; ++h)
                        x64[o * CB_N + own_coeff(w, lane, h & 3) + (h >= 4 ? CB_M : 0)] = t[o][h];
                __syncthreads();
                const int kshift = (int)(mrot & (N - 1));
                const bool hi = mrot >= N;
#pragma unroll
                for (int o = 0; o < 2; ++o)
#pragma unroll
                    for (int h = 0; h < 8; ++h) {
                        int idx = own_coeff(w, lane, h & 3) + (h >= 4 ? CB_M : 0) - kshift;
                        const bool wrapped = idx < 0;
                        idx += wrapped ? (int)N : 0;
                        uint64_t v = x64[o * CB_N + idx];
                        if (hi ^ wrapped) v = 0 - v;
                        acc[o][h] += v - t[o][h];
                    }
                __syncthreads();
            }
        }
        if constexpr (LUT) {  // the whole RLWE (mask, body), Gpow_j / 2 added to body coefficient j < 2
            uint64_t* r = io.rlwe + (size_t)cidx * 2 * N;
            const uint64_t o0 = io.offsets ? mu0 : 0, o1 = io.offsets ? mu1 : 0;
#pragma unroll
            for (int h = 0; h < 8; ++h) {
                const uint32_t c = (uint32_t)own_coeff(w, lane, h & 3) + (h >= 4 ? CB_M : 0);
                r[c] = acc[0][h];
                r[N + c] = acc[1][h] + (c == 0 ? o0 : c == 1 ? o1 : 0);
            }
            continue;
        }
        // sample extraction of coefficients 0 and 1
        uint64_t* o0 = out + (size_t)cidx * 2 * (N + 1);
        uint64_t* o1 = o0 + (N + 1);
#pragma unroll
        for (int h = 0; h < 8; ++h) {
            const uint32_t c = (uint32_t)own_coeff(w, lane, h & 3) + (h >= 4 ? CB_M : 0);
            const uint64_t v = acc[0][h];
            o0[(0u - c) & (N - 1)] = (c == 0) ? v : 0 - v;
            o1[(1u - c) & (N - 1)] = (c <= 1) ? v : 0 - v;
        }
        if (lane == 0 && w == 0) o0[N] = acc[1][0] + mu0;
        if (lane == 0 && w == 1) o1[N] = acc[1][0] + mu1;
    }
    if (n < 0) out[threadIdx.x] = sink;  // never (n > 0): keeps the prefetch loads
}

// Two key bits per step with one ciphertext per 8-wave workgroup (the one-product form):
// wave group g (waves 4g .. 4g + 3) owns accumulator polynomial g.  Per step each group
// decomposes its polynomial into the digit rows r = 2j + g, transforms them and sums both
// outputs' products over its rows; the partial sum of the other group's output goes through
// LDS, and each group inverse-transforms its own output and updates its polynomial.  A wave
// runs half the forward transforms and MACs of the 4-wave form per step, and a CU holds two
// waves per SIMD with one workgroup: the form for launches of at most one workgroup per CU,
// where the 4-wave forms leave a wave per SIMD.  Each output's sum is the two groups' partial
// sums added once, where the 4-wave form runs one sum over all eight rows, so the words differ
// from that form's by FP64 rounding (the phases agree within the one-product form's bound).
template <bool LUT>
__global__ __launch_bounds__(512, 1) void k_pbs2_split(const uint32_t* __restrict__ lwe, uint32_t batch, int n,
                                                      const cplx* __restrict__ bskf, FFT1024Tables Tg,
                                                      uint64_t* __restrict__ out, PbsLut io) {
    __shared__ cplx tl[768];
    __shared__ cplx lds_all[8][F256_LDS_CPLX];
    __shared__ cplx xbuf[2][2][1024];  // [group][buffer]
    __shared__ uint16_t at[1024];
    const FFT1024Tables T = stage_tables(Tg, tl);
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int g = wv >> 2, w = wv & 3;
    cplx trw[4];
    load_trw(Tg, w, lane, trw);
    cplx* lds = lds_all[wv];
    const uint64_t OFF = pbs_digit_offset<PBS_LOGB>();
    constexpr uint32_t N = CB_N, Q2 = 2 * CB_N;
    const uint64_t mu0 = 1ull << 47, mu1 = 1ull << 55;
    uint32_t ek[4];
    const uint32_t kq = (uint32_t)((lane >> 4) + 4 * ((lane >> 2) & 3) + 16 * (lane & 3) + 256 * w);
#pragma unroll
    for (int s = 0; s < 4; ++s) ek[s] = (1u - 4u * (kq + 64u * (uint32_t)s)) & (Q2 - 1);

    // L2 prefetch of the next step's pattern keys (786 KB): workgroups are dealt to the 8 XCDs
    // in turn and run the steps nearly in lockstep, so the up to 16 workgroups of one XCD each
    // touch one share of the step's 128-byte lines (one word per line), issued after the step's
    // last key loads (loads return in order: earlier, the step's own key waits would include it)
    // and consumed a step later
    constexpr uint32_t KLINES = 3 * PBS_ROWS * 2 * CB_M * sizeof(cplx) / 128;
    const uint32_t per_xcd = (gridDim.x + 7) / 8 < 16 ? (gridDim.x + 7) / 8 : 16;
    const uint32_t share = (KLINES + per_xcd - 1) / per_xcd, part = (blockIdx.x >> 3) % per_xcd;
    uint32_t pf0 = 0, pf1 = 0, sink = 0;

    for (uint32_t cidx = blockIdx.x; cidx < batch; cidx += gridDim.x) {
        const uint32_t* ct = lwe + (size_t)cidx * (n + 1);
        __syncthreads();
        uint32_t bt;
        if constexpr (LUT) {
            for (int k = threadIdx.x; k < n; k += blockDim.x)
                at[k] = (uint16_t)((Q2 - special_ms_2048_bw(ct[k], io.bw)) & (Q2 - 1));
            bt = special_ms_2048_bw(ct[n] + io.b_add, io.bw);
        } else {
            for (int k = threadIdx.x; k < n; k += blockDim.x) at[k] = (uint16_t)special_ms_2048(ct[k]);
            bt = special_ms_2048(ct[n] + (1u << 30));
        }
        __syncthreads();

        uint64_t acc[8];  // polynomial g
#pragma unroll
        for (int h = 0; h < 8; ++h) {
            const uint32_t c = (uint32_t)own_coeff(w, lane, h & 3) + (h >= 4 ? CB_M : 0);
            uint64_t v;
            if constexpr (LUT) {
                const uint32_t m = (c + bt) & (Q2 - 1);
                v = io.lut[m & (N - 1)];
                if (m >= N) v = 0 - v;
            } else {
                const uint32_t arg = (bt + Q2 - (c & ~1u)) & (Q2 - 1);
                v = 0 - ((c & 1) ? mu1 : mu0);
                if (arg >= N) v = 0 - v;
            }
            acc[h] = g ? v : 0;
        }
        int buf = 0;
        for (int i = 0; i < n; i += 2) {
            const uint32_t mrot = (Q2 - (uint32_t)at[i]) & (Q2 - 1);
            const uint32_t mrot2 = (Q2 - (uint32_t)at[i + 1]) & (Q2 - 1);
            if ((mrot | mrot2) == 0) continue;
            const cplx* key = bskf + (size_t)(i / 2) * 3 * PBS_ROWS * 2 * CB_M;
            // z1, z2 per frequency; the pattern factors z1 - 1, z1 z2 - 1, z2 - 1 (patterns
            // (1,0), (1,1), (0,1)) are formed where they are used, which keeps 16 registers free
            cplx z1[4], z2[4];
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                z1[s] = T.rot[(mrot * ek[s]) & (Q2 - 1)];
                z2[s] = T.rot[(mrot2 * ek[s]) & (Q2 - 1)];
            }
            // sums for this group's output (mown) and the other group's (moth); the key rows
            // of each are addressed by offset, so no array is indexed by the group
            cplx mown[4], moth[4];
#pragma unroll
            for (int s = 0; s < 4; ++s) mown[s] = moth[s] = {0.0, 0.0};
            const int own_off = g * CB_M, oth_off = (1 - g) * CB_M;
            uint64_t tp[8];
#pragma unroll
            for (int h = 0; h < 8; ++h) tp[h] = acc[h] + OFF;
#pragma unroll
            for (int j = 0; j < PBS_L; ++j) {
                const int r = 2 * j + g;
                cplx x[4];
#pragma unroll
                for (int a = 0; a < 4; ++a) x[a] = {pbs_digit<PBS_LOGB>(tp[a], j), pbs_digit<PBS_LOGB>(tp[a + 4], j)};
                f1024_fwd<true>(x, lds, xbuf[g][buf], T, trw, w, lane);
                buf ^= 1;
                cplx y[4];
                auto load_keys = [&](cplx(&kv)[2][4], int q) {
                    const int pt = q == 0 ? 0 : q == 1 ? 2 : 1;
                    const cplx* kr = key + ((size_t)pt * PBS_ROWS + r) * 2 * CB_M + (w * 4) * 64 + lane;
#pragma unroll
                    for (int s = 0; s < 4; ++s) {
                        kv[0][s] = kr[own_off + s * 64];
                        kv[1][s] = kr[oth_off + s * 64];
                    }
                };
                // a scheduling fence keeps a batch's loads together and leaves their waits at
                // the first use, so the next batch is in flight meanwhile (1-block AES 18.5 ->
                // 21.1 blocks/s against waiting for each batch at once)
                auto close_batch = [](cplx(&)[2][4]) { __builtin_amdgcn_sched_barrier(0); };
                auto mac_pattern = [&](const cplx(&kv)[2][4], int q) {
#pragma unroll
                    for (int s = 0; s < 4; ++s) {
                        const cplx z = q == 0 ? z1[s] : q == 2 ? z2[s] : cmul(z1[s], z2[s]);
                        y[s] = cmul(x[s], cplx{z.re - 1.0, z.im});
                    }
#pragma unroll
                    for (int s = 0; s < 4; ++s) {
                        cmac(mown[s], y[s], kv[0][s]);
                        cmac(moth[s], y[s], kv[1][s]);
                    }
                };
                cplx ka[2][4], kb[2][4];
                load_keys(ka, 0);
                close_batch(ka);
                load_keys(kb, 1);
                close_batch(kb);
                mac_pattern(ka, 0);
                load_keys(ka, 2);
                close_batch(ka);
                mac_pattern(kb, 1);
                mac_pattern(ka, 2);
            }
            {  // branch-free: the last step re-touches its own key, lines past the share are clamped
                sink ^= pf0 ^ pf1;  // the previous step's prefetch, long landed
                const char* nk = reinterpret_cast<const char*>(i + 2 < n ? key + 3 * PBS_ROWS * 2 * CB_M : key);
                const uint32_t end = (part + 1) * share < KLINES ? (part + 1) * share : KLINES;
                const uint32_t l0 = part * share + threadIdx.x, l1 = l0 + blockDim.x;
                pf0 = *reinterpret_cast<const uint32_t*>(nk + (size_t)(l0 < end ? l0 : end - 1) * 128);
                pf1 = *reinterpret_cast<const uint32_t*>(nk + (size_t)(l1 < end ? l1 : end - 1) * 128);
            }
            // the other group's output: this group's partial through LDS (the buffer the
            // third transform used; the fourth transform's barrier ended its reads)
            cplx* ex = xbuf[g][buf];
#pragma unroll
            for (int s = 0; s < 4; ++s) ex[(w * 4 + s) * 64 + lane] = moth[s];
            __syncthreads();
            const cplx* oth = xbuf[1 - g][buf];
            cplx y0[4];
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                y0[s] = cadd(mown[s], oth[(w * 4 + s) * 64 + lane]);
            }
            buf ^= 1;
            // the exchange buffer is read by the other group until that group's inverse
            // transform reaches its barrier: this transform uses the other buffer
            f1024_inv<true>(y0, lds, xbuf[g][buf], T, trw, w, lane);
            buf ^= 1;
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                acc[a] += torus64_from_double(y0[a].re);
                acc[a + 4] += torus64_from_double(y0[a].im);
            }
        }
        if constexpr (LUT) {
            uint64_t* r = io.rlwe + (size_t)cidx * 2 * N;
            const uint64_t o0 = io.offsets ? mu0 : 0, o1 = io.offsets ? mu1 : 0;
#pragma unroll
            for (int h = 0; h < 8; ++h) {
                const uint32_t c = (uint32_t)own_coeff(w, lane, h & 3) + (h >= 4 ? CB_M : 0);
                r[g * N + c] = acc[h] + (g && c == 0 ? o0 : g && c == 1 ? o1 : 0);
            }
            continue;
        }
        uint64_t* o0 = out + (size_t)cidx * 2 * (N + 1);
        uint64_t* o1 = o0 + (N + 1);
        if (g == 0) {  // sample extraction of coefficients 0 and 1 from the mask polynomial
#pragma unroll
            for (int h = 0; h < 8; ++h) {
                const uint32_t c = (uint32_t)own_coeff(w, lane, h & 3) + (h >= 4 ? CB_M : 0);
                const uint64_t v = acc[h];
                o0[(0u - c) & (N - 1)] = (c == 0) ? v : 0 - v;
                o1[(1u - c) & (N - 1)] = (c <= 1) ? v : 0 - v;
            }
        } else {  // bodies: coefficient 0 sits in wave 0, coefficient 1 in wave 1 (lane 0, h = 0)
            if (lane == 0 && w == 0) o0[N] = acc[0] + mu0;
            if (lane == 0 && w == 1) o1[N] = acc[0] + mu1;
        }
    }
    if (n < 0) out[threadIdx.x] = sink;  // never (n > 0): keeps the prefetch loads
}

// ===========================================================================
// HomTrace + SchemeSwitch LWE -> RGSW (the reference's conversion, cirbts-batch.cu:
// 1612-1790; trace round kernel.cu:7832-7888, scheme switch 8407-8493, row save
// 7776-7789), on the 64-bit torus with exact two-limb key-switching products.
// One workgroup (4 waves) per (input, CB level j): the PBS accumulator times X^-j plus
// (Gpow_j / 2) / N at coefficient 0, then log2 N trace rounds
//   (a, b) += (ks_0, sigma_k(b) + ks_1),  ks = sum_d dig_d(sigma_k(a)) HT_k,d,
//   sigma_k: X -> X^(N/2^k + 1),
// then the scheme switch (b + ks_0, ks_1), ks = sum_d dig_d(a) SS_d; rows rounded to
// the 32-bit torus: row 2j = SS, row 2j + 1 = HT.
// Keys (two-limb Fourier, the BSK2's layout): HT [log2 N][KS_L][2 out][2 limbs][1024],
// SS [KS_L][2 out][2 limbs][1024].
// ===========================================================================
// k^-1 mod 2^12 for odd k (Newton's iteration x <- x (2 - k x) doubles the correct bits)
__device__ __forceinline__ uint32_t inv_mod_4096(uint32_t k) {
    uint32_t x = k;
#pragma unroll
    for (int i = 0; i < 4; ++i) x *= 2u - k * x;
    return x & 4095u;
}

// exact sum_d dig_d(x) key_d for the wave-distributed coefficients x (8 per lane):
// ks[o][h], o = output component
__device__ __forceinline__ void htss_keyswitch(const uint64_t (&x)[8], const cplx* __restrict__ key, cplx* lds,
                                               cplx (*xbuf)[1024], int& buf, const FFT1024Tables& T,
                                               const cplx (&trw)[4], int w, int lane, uint64_t (&ks)[2][8]) {
    const uint64_t OFF = digit_offset64<KS_LOGB, KS_L>();
    uint64_t tp[8];
#pragma unroll
    for (int h = 0; h < 8; ++h) tp[h] = x[h] + OFF;
    cplx macc[2][2][4];
#pragma unroll
    for (int o = 0; o < 2; ++o)
#pragma unroll
        for (int lb = 0; lb < 2; ++lb)
#pragma unroll
            for (int s = 0; s < 4; ++s) macc[o][lb][s] = {0.0, 0.0};
#pragma unroll 1
    for (int d = 0; d < KS_L; ++d) {
        cplx xd[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) xd[a] = {digit64<KS_LOGB, KS_L>(tp[a], d), digit64<KS_LOGB, KS_L>(tp[a + 4], d)};
        f1024_fwd<false>(xd, lds, xbuf[buf], T, trw, w, lane);
        buf ^= 1;
        const cplx* kr = key + (size_t)d * 2 * 2 * CB_M + (w * 4) * 64 + lane;
#pragma unroll
        for (int o = 0; o < 2; ++o)
#pragma unroll
            for (int lb = 0; lb < 2; ++lb)
#pragma unroll
                for (int s = 0; s < 4; ++s) cmac(macc[o][lb][s], xd[s], kr[(o * 2 + lb) * CB_M + s * 64]);
    }
#pragma unroll
    for (int o = 0; o < 2; ++o) {
        cplx y0[4], y1[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) { y0[s] = macc[o][0][s]; y1[s] = macc[o][1][s]; }
        f1024_inv<false>(y0, lds, xbuf[buf], T, trw, w, lane);
        buf ^= 1;
        f1024_inv<false>(y1, lds, xbuf[buf], T, trw, w, lane);
        buf ^= 1;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            ks[o][a] = (uint64_t)round_i64(y0[a].re, 1.0 / CB_M) + ((uint64_t)round_i64(y1[a].re, 1.0 / CB_M) << 32);
            ks[o][a + 4] = (uint64_t)round_i64(y0[a].im, 1.0 / CB_M) + ((uint64_t)round_i64(y1[a].im, 1.0 / CB_M) << 32);
        }
    }
}

__global__ __launch_bounds__(256) void k_htss(const uint64_t* __restrict__ acc_in, uint32_t batch,
                                             const cplx* __restrict__ htkey, const cplx* __restrict__ sskey,
                                             FFT1024Tables Tg, uint32_t* __restrict__ rgsw) {
    __shared__ cplx tl[768];
    __shared__ cplx lds_all[4][F256_LDS_CPLX];
    __shared__ cplx xbuf[2][1024];
    const FFT1024Tables T = stage_tables(Tg, tl);
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    cplx trw[4];
    load_trw(Tg, w, lane, trw);
    cplx* lds = lds_all[w];
    uint64_t* x64 = reinterpret_cast<uint64_t*>(&xbuf[0][0]);  // [2][N] u64: the automorphism exchange
    constexpr uint32_t N = CB_N, Q2 = 2 * CB_N;
    constexpr int LOGN = 11;
    int buf = 0;
    for (uint32_t item = blockIdx.x; item < 2 * batch; item += gridDim.x) {
        const uint32_t cidx = item >> 1, j = item & 1;
        const uint64_t* src = acc_in + (size_t)cidx * 2 * N;
        uint64_t lv[2][8];
#pragma unroll
        for (int h = 0; h < 8; ++h) {  // acc X^-j
            const uint32_t c = (uint32_t)own_coeff(w, lane, h & 3) + (h >= 4 ? CB_M : 0);
            const uint32_t sidx = c + j;
#pragma unroll
            for (int o = 0; o < 2; ++o) lv[o][h] = sidx < N ? src[o * N + sidx] : 0 - src[o * N + sidx - N];
            if (c == 0) lv[1][h] += 1ull << (47 + 8 * j - LOGN);  // (Gpow_j / 2) / N
        }
#pragma unroll 1
        for (int k = 0; k < LOGN; ++k) {
            const uint32_t einv = inv_mod_4096((N >> k) + 1);
            __syncthreads();  // the previous round's transforms are done with xbuf
#pragma unroll
            for (int o = 0; o < 2; ++o)
#pragma unroll
                for (int h = 0; h < 8; ++h) x64[o * N + own_coeff(w, lane, h & 3) + (h >= 4 ? CB_M : 0)] = lv[o][h];
            __syncthreads();
            uint64_t sa[8], sb[8];  // sigma(a), sigma(b): coefficient t is +-c[t e^-1 mod 2N]
#pragma unroll
            for (int h = 0; h < 8; ++h) {
                const uint32_t t = (uint32_t)own_coeff(w, lane, h & 3) + (h >= 4 ? CB_M : 0);
                const uint32_t u = (t * einv) & (Q2 - 1);
                const uint32_t si = u < N ? u : u - N;
                const uint64_t va = x64[si], vb = x64[N + si];
                sa[h] = u < N ? va : 0 - va;
                sb[h] = u < N ? vb : 0 - vb;
            }
            __syncthreads();  // every wave has read x64 before the transforms reuse xbuf
            uint64_t ks[2][8];
            htss_keyswitch(sa, htkey + (size_t)k * KS_L * 2 * 2 * CB_M, lds, xbuf, buf, T, trw, w, lane, ks);
#pragma unroll
            for (int h = 0; h < 8; ++h) {
                lv[0][h] += ks[0][h];
                lv[1][h] += sb[h] + ks[1][h];
            }
        }
        uint64_t ks[2][8];
        htss_keyswitch(lv[0], sskey, lds, xbuf, buf, T, trw, w, lane, ks);
        uint32_t* out = rgsw + (size_t)cidx * 4 * 2 * N;
        uint32_t* ss_row = out + (size_t)(2 * j) * 2 * N;
        uint32_t* ht_row = out + (size_t)(2 * j + 1) * 2 * N;
#pragma unroll
        for (int h = 0; h < 8; ++h) {
            const uint32_t c = (uint32_t)own_coeff(w, lane, h & 3) + (h >= 4 ? CB_M : 0);
            constexpr uint64_t R = 1ull << 31;  // round to the top 32 bits
            ss_row[c] = (uint32_t)((lv[1][h] + ks[0][h] + R) >> 32);
            ss_row[N + c] = (uint32_t)((ks[1][h] + R) >> 32);
            ht_row[c] = (uint32_t)((lv[0][h] + R) >> 32);
            ht_row[N + c] = (uint32_t)((lv[1][h] + R) >> 32);
        }
    }
}

// ===========================================================================
// Private functional key switching LWE_K2 (torus-64) -> RLWE_K2 (torus-32),
// CGGI PrivKS, for function u:
//   out = - sum_{i<=N} sum_{t<8} KS_u[i][t][digit_t(a~_i)]  (mod 2^32),
// a_N = b, a~_i = round(a_i / 2^32), unsigned base-16 digits, least
// significant first; digit 0 contributes nothing.  Key layout
// [u][i][t][v-1][2N] u32 (v = 1..15).
//
// k_pfks_digits rounds the inputs and transposes them to [i][Bpad] so that
// one i-row of a wave's PF_CT inputs is a uniform (scalar-load) read.
// k_pfks32: a workgroup of PF_WAVES waves shares one 128-column tile of the
// key; per (i, t) the 15 candidate rows are loaded once by the workgroup into
// an LDS table (row 0 = 0, double-buffered, one barrier per step) and every
// wave selects, for each of its PF_CT inputs, the row named by that input's
// digit: one LDS read and one subtraction per (input, column).  blockIdx.z
// picks an i-slice and the function; slices add into the zero-filled output
// with 32-bit atomics (addition mod 2^32 is order-free, so the result is
// exact).  The key is read once per (column tile, input group) from L2, and
// the workgroups of one key tile share an XCD (their linear ids differ by
// multiples of 64), so HBM sees it about once.
// ===========================================================================
constexpr int PF_WAVES = 8, PF_CT = 32, PF_GROUP = PF_WAVES * PF_CT;

__global__ __launch_bounds__(256) void k_pfks_digits(const uint64_t* __restrict__ lwe64, uint32_t B, uint32_t Bpad,
                                                    uint32_t* __restrict__ at) {
    __shared__ uint32_t tile[64][65];
    constexpr int N = CB_N;
    const int i0 = blockIdx.x * 64, k0 = blockIdx.y * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int r = ty; r < 64; r += 4) {  // r = input offset, tx = i offset
        const uint32_t k = k0 + r;
        const int i = i0 + tx;
        uint32_t v = 0;
        if (k < B && i <= N) v = (uint32_t)((lwe64[(size_t)k * (N + 1) + i] + (1ull << 31)) >> 32);
        tile[r][tx] = v;
    }
    __syncthreads();
    for (int r = ty; r < 64; r += 4) {  // r = i offset, tx = input offset
        const int i = i0 + r;
        const uint32_t k = k0 + tx;
        if (i <= N && k < Bpad) at[(size_t)i * Bpad + k] = tile[tx][r];
    }
}

__global__ __launch_bounds__(64 * PF_WAVES) void k_pfks32(const uint32_t* __restrict__ at, uint32_t B, uint32_t Bpad,
                                                         const uint32_t* __restrict__ key, size_t key_u_stride,
                                                         int u0, int slices, uint32_t* __restrict__ out,
                                                         size_t out_stride, size_t out_u_offset) {
    // two adjacent columns per lane: one LDS address and one 8-byte read feed two subtractions
    __shared__ uint2 tab[2][16][64];
    constexpr int N = CB_N, COLS = 2 * CB_N;
    const int ISLICE = (N + 1 + slices - 1) / slices;
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int slice = blockIdx.z % slices;
    const int u = u0 + (int)(blockIdx.z / slices);
    const int ibeg = slice * ISLICE;
    const int iend = (ibeg + ISLICE) < (N + 1) ? (ibeg + ISLICE) : (N + 1);
    if (ibeg >= iend) return;  // empty slice (uniform over the workgroup)
    const uint32_t cbase = blockIdx.y * PF_GROUP + w * PF_CT;
    const bool active = cbase < B;  // idle waves still take part in the barriers
    // this thread's share of a tile: candidate v = 1 + w (all waves) and v = 9 + w (waves 0..6)
    const uint2* kcol = reinterpret_cast<const uint2*>(key + (size_t)u * key_u_stride + blockIdx.x * 128) + lane;
    const bool second = w < 7;
    if (tid < 64) {
        tab[0][0][tid] = make_uint2(0, 0);
        tab[1][0][tid] = make_uint2(0, 0);
    }
    uint32_t acc[PF_CT][2];
#pragma unroll
    for (int k = 0; k < PF_CT; ++k) acc[k][0] = acc[k][1] = 0;
    // Tile m (m = 8 (i - ibeg) + t) is loaded into register slot m % PF_RING at step
    // m - PF_RING, stored into tab[m & 1] at the start of step m - 1 and read at step m:
    // PF_RING - 1 steps of latency cover the key stream.  With the t loop unrolled by 8
    // every slot index is a compile-time constant.
    constexpr int PF_RING = 4;
    uint2 kv[PF_RING][2];
    auto load_tile = [&](int slot, int i, int t) {
        uint64_t off = ((uint64_t)(i * 8 + t) * 15 + w) * (COLS / 2);
        asm volatile("" : "+s"(off));  // keep the prefetch where it is written
        const uint2* kr = kcol + off;
        kv[slot][0] = kr[0];
        kv[slot][1] = second ? kr[(size_t)8 * (COLS / 2)] : make_uint2(0, 0);
    };
    auto store_tile = [&](int slot, int b) {
        tab[b][1 + w][lane] = kv[slot][0];
        if (second) tab[b][9 + w][lane] = kv[slot][1];
    };
#pragma unroll
    for (int m = 0; m < PF_RING; ++m)
        if (ibeg + m / 8 < iend) load_tile(m, ibeg + m / 8, m % 8);
    store_tile(0, 0);
    __syncthreads();
    for (int i = ibeg; i < iend; ++i) {
        uint32_t a[PF_CT];
        if (active) {
            const uint32_t* ar = at + (size_t)i * Bpad + cbase;
#pragma unroll
            for (int k = 0; k < PF_CT; ++k) a[k] = ar[k];
        }
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            // store tile m + 1 (loaded PF_RING - 1 steps ago), refill its slot with tile m + PF_RING
            const int t1 = t + 1, tr = t + PF_RING;
            if (i + t1 / 8 < iend) store_tile(t1 % PF_RING, t1 & 1);
            if (i + tr / 8 < iend) load_tile(tr % PF_RING, i + tr / 8, tr % 8);
            if (active) {
#pragma unroll
                for (int k = 0; k < PF_CT; ++k) {
                    const uint2 v = tab[t & 1][(a[k] >> (4 * t)) & 15u][lane];
                    acc[k][0] -= v.x;
                    acc[k][1] -= v.y;
                }
            }
#pragma unroll
            for (int k = 0; k < PF_CT; ++k) asm volatile("" : "+v"(acc[k][0]), "+v"(acc[k][1]));
            __syncthreads();
        }
    }
    if (active) {
        uint32_t* o = out + (size_t)(u - u0) * out_u_offset + blockIdx.x * 128 + 2 * lane;
#pragma unroll
        for (int k = 0; k < PF_CT; ++k)
            if (cbase + k < B) {
                atomicAdd(&o[(size_t)(cbase + k) * out_stride], acc[k][0]);
                atomicAdd(&o[(size_t)(cbase + k) * out_stride + 1], acc[k][1]);
            }
    }
}

// Pseudo-random words (a 32-bit integer hash of the index) for timing inputs: a blind
// rotation over zero data would skip every step.
__global__ void k_fill_hash32(uint32_t* __restrict__ p, size_t count, uint32_t seed) {
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < count; e += (size_t)gridDim.x * blockDim.x) {
        uint32_t x = (uint32_t)e * 0x9E3779B9u ^ seed ^ (uint32_t)(e >> 32);
        x ^= x >> 16;
        x *= 0x7FEB352Du;
        x ^= x >> 15;
        x *= 0x846CA68Bu;
        x ^= x >> 16;
        p[e] = x;
    }
}

// gathers LWEs given as a mask array (rows a_stride words apart) and a body array,
// both NativeInt-wide (u64) with values on the 32-bit torus, into [batch][n+1] u32
__global__ void k_pack_lwe64(const uint64_t* __restrict__ a, const uint64_t* __restrict__ b, uint32_t batch, int n,
                             uint32_t a_stride, uint32_t* __restrict__ out) {
    const size_t total = (size_t)batch * (n + 1);
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const size_t k = e / (size_t)(n + 1);
        const int i = (int)(e - k * (size_t)(n + 1));
        out[e] = (uint32_t)(i < n ? a[k * a_stride + i] : b[k]);
    }
}

// zero `count` rows of `width` u32 spaced `stride` apart
__global__ void k_zero_rows32(uint32_t* __restrict__ out, uint32_t count, size_t stride, uint32_t width) {
    const size_t total = (size_t)count * width;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x)
        out[(e / width) * stride + (e % width)] = 0;
}

// ===========================================================================
// Linear-key private functional key switch as an int8 GEMM on the matrix cores.
// Key KL[u][i][t] = RLWE_K2(f_u(kappa_i 2^(4t)) 2^8) on Z/2^40 (sigma 2^7 units,
// 2^-33 of the torus), one row per (i, t); input digits signed, d_t in [-8, 8):
//   out40[L][col] = - sum_{k = 8i + t} d[L][k] KL[u][k][col]   (mod 2^40)
//   out32 = round(out40 / 2^8)
// K = 8 (N + 1) = 16392, padded with zero digits to PFKL_KP = 513 x 32.  Each
// 40-bit key word is held as 5 signed bytes s_0..s_4 (sum s_b 2^(8b) = word mod
// 2^40), so the sum is 5 int8 products of the same digit matrix: |sum| <= 16392
// x 8 x 128 < 2^31 in the int32 accumulators, recombined as sum_b acc_b 2^(8b).
// Both operands are stored in the MFMA fragment order, so that one operand load of
// a wave is one contiguous 1 KB run: element (r, k) of a 32-row block (digit rows =
// inputs, key rows = columns) sits in fragment (block, k / 32) at byte
// 16 (r + 32 ((k % 32) / 16)) + k % 16 (pfkl_frag_off):
//   digits D [Mpad / 32][PFKL_KP / 32][1 KB], key KT [2][5][2N / 32][PFKL_KP / 32][1 KB].
// ===========================================================================
constexpr int PFKL_SLICES = 5, PFKL_KP = 16416, PFKL_NSHIFT = 25;
constexpr uint64_t PFKL_MASK = (1ull << 40) - 1;
static_assert(PFKL_KP % 32 == 0 && PFKL_KP >= 8 * (CB_N + 1), "padded digit count");
constexpr int PFKL_KS = PFKL_KP / 32;  // k steps of 32

// byte offset of element (r, k) of a run of 32-row blocks in fragment order
__host__ __device__ inline size_t pfkl_frag_off(size_t r, uint32_t k) {
    return (((r / 32) * PFKL_KS + k / 32) * 64 + (r % 32) + 32 * ((k % 32) / 16)) * 16 + k % 16;
}

// one workgroup per key row (u, i, t), Philox id (u (N + 1) + i) 8 + t: a = 40 low
// bits of Philox words, b = a * K2 (schoolbook over the support of the binary K2) + e
// + msg, both polynomials written as signed byte slices at k = 8 i + t (columns c for
// a, N + c for b).  The padded k of the key buffer are zeroed by the caller.
__global__ __launch_bounds__(256) void k_gen_pfkl40(uint64_t seed, const int32_t* __restrict__ K2,
                                                   int8_t* __restrict__ KT) {
    __shared__ uint64_t as[CB_N];
    __shared__ uint16_t sup[CB_N];
    __shared__ uint32_t nsup;
    constexpr int N = CB_N;
    const uint32_t row = blockIdx.x;
    const int t = (int)(row % 8);
    const uint32_t ui = row / 8;
    const int i = (int)(ui % (N + 1)), u = (int)(ui / (N + 1));
    if (threadIdx.x == 0) nsup = 0;
    __syncthreads();
    for (int c = threadIdx.x; c < N; c += blockDim.x) {
        as[c] = rand64(seed, DOM_PFKL, row, (uint32_t)c) & PFKL_MASK;
        if (K2[c]) sup[atomicAdd(&nsup, 1u)] = (uint16_t)c;  // LDS atomic: order-free sum below
    }
    __syncthreads();
    const uint32_t ns = nsup;
    const uint32_t kappa = i < N ? (uint32_t)K2[i] : 0xFFFFFFFFu;
    const uint64_t cval = ((uint64_t)(kappa << (4 * t)) << 8) & PFKL_MASK;
    const uint32_t kpos = 8u * i + t;
    int8_t* ka = KT + (size_t)u * PFKL_SLICES * 2 * N * PFKL_KP;
    const size_t sstride = (size_t)2 * N * PFKL_KP;
    for (int c = threadIdx.x; c < N; c += blockDim.x) {
        uint64_t b = 0;
        for (uint32_t q = 0; q < ns; ++q) {
            const int j = sup[q];
            b += j <= c ? as[c - j] : (0 - as[N + c - j]);
        }
        b += (uint64_t)noise(seed, DOM_PFKL, row, (uint32_t)c, PFKL_NSHIFT);
        if (u == 1) {
            if (c == 0) b += cval;
        } else {
            b -= (uint64_t)K2[c] * cval;
        }
        uint64_t v[2] = {as[c], b & PFKL_MASK};
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            int carry = 0;
            int8_t* o = ka + pfkl_frag_off((size_t)h * N + c, kpos);
#pragma unroll
            for (int sl = 0; sl < PFKL_SLICES; ++sl) {
                int x = (int)((v[h] >> (8 * sl)) & 255u) + carry;
                carry = x >= 128;
                o[sl * sstride] = (int8_t)(x - (carry << 8));
            }
        }
    }
}

// signed base-16 digits of the rounded inputs: D[m][8 q + t] = d_t(round(lwe64[m][q] / 2^32))
// for q <= N and m < B, zero elsewhere in [Mpad][PFKL_KP]
__global__ __launch_bounds__(256) void k_pfkl_digits(const uint64_t* __restrict__ lwe64, uint32_t B, uint32_t Mpad,
                                                    int8_t* __restrict__ D) {
    constexpr int N = CB_N, Q = PFKL_KP / 8;
    const size_t total = (size_t)Mpad * Q;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const size_t m = e / Q;
        const int q = (int)(e - m * Q);
        uint64_t word = 0;
        if (m < B && q <= N) {
            const uint32_t at = (uint32_t)((lwe64[m * (N + 1) + q] + (1ull << 31)) >> 32);
            const uint32_t w = at + 0x88888888u;
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const int d = (int)((w >> (4 * t)) & 15u) - 8;
                word |= (uint64_t)(uint8_t)(int8_t)d << (8 * t);
            }
        }
        *reinterpret_cast<uint64_t*>(D + pfkl_frag_off(m, 8u * q)) = word;  // 8 bytes of one 16-byte run
    }
}

typedef int pfkl_v4i __attribute__((ext_vector_type(4)));
typedef int pfkl_v16i __attribute__((ext_vector_type(16)));

// v_mfma_i32_32x32x32_i8: lane l holds 16 bytes of A row (l % 32) and of B column
// (l % 32) at k = 16 (l / 32) + 0..15; accumulator register j of lane l is
// C[8 (j / 4) + 4 (l / 32) + j % 4][l % 32].
// Grid x = Mpad / 128 input tiles, y = nu * 2N / 64 column tiles.  Wave w computes
// 64 inputs (m0 + 0..63, two 32-row blocks) x 32 columns x 5 slices: ten 16-register
// accumulators (160 registers, held in AGPRs), one 16-byte digit load per row block and
// one key load per slice per 32 k.  One wave per SIMD leaves the registers for four
// operand sets in flight: the loads are latency-bound, and two steps of prefetch beat
// two waves per SIMD with one (3.34 against 3.78 ms per 17-block round).  The pairs of
// waves that read the same digit or key fragment are kept within four steps of each
// other by a workgroup barrier, so that the second read of a fragment finds it in L1
// (8192 inputs: 6.13 -> 5.96 ms per function pair, same words).
__global__ __launch_bounds__(256, 1) void k_pfkl_gemm(const int8_t* __restrict__ D, uint32_t B,
                                                  const int8_t* __restrict__ KT, int u0,
                                                  uint32_t* __restrict__ out, size_t out_stride,
                                                  size_t out_u_offset, int cols_first) {
    constexpr int COLS = 2 * CB_N;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // cols_first: blockIdx.x = column tile, blockIdx.y = row block, so that the row blocks of one
    // key tile have linear workgroup ids 128 apart, on one XCD (a launch of at most two row
    // blocks runs whole, and the second read of each key tile hits that XCD's L2)
    const uint32_t rb = cols_first ? blockIdx.y : blockIdx.x, ct = cols_first ? blockIdx.x : blockIdx.y;
    const uint32_t m0 = rb * 128 + (w & 1) * 64;
    const uint32_t cg = ct * 64 + (w >> 1) * 32;
    const int uu = (int)(cg / COLS), col0 = (int)(cg % COLS);
    // fragment (block, step) of lane l at ((block * PFKL_KS + step) * 64 + l) * 16: step k / 32
    // advances a pointer by 1 KB
    const int8_t* a0 = D + ((size_t)(m0 / 32) * PFKL_KS * 64 + lane) * 16;
    const int8_t* a1 = a0 + (size_t)PFKL_KS * 1024;
    const size_t sstride = (size_t)COLS * PFKL_KP;
    const int8_t* bp = KT + (size_t)(u0 + uu) * PFKL_SLICES * sstride + ((size_t)(col0 / 32) * PFKL_KS * 64 + lane) * 16;
    pfkl_v16i acc[2][PFKL_SLICES];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int sl = 0; sl < PFKL_SLICES; ++sl) acc[r][sl] = pfkl_v16i{};
    // four operand sets rotate over the steps (513 = 4 x 128 + 1), each loaded three steps
    // ahead of its products (the scheduling barriers keep the loads in place), so a load
    // has thirty products (~960 cycles) to arrive.  Loads past the last step reread step
    // 0 (in bounds, unused).
    struct Ops { pfkl_v4i a[2], b[PFKL_SLICES]; };
    auto load = [&](Ops& x, int ks) {  // fragment step ks at byte 1024 ks
        const size_t fo = (size_t)(ks < PFKL_KS ? ks : 0) * 1024;
        x.a[0] = *reinterpret_cast<const pfkl_v4i*>(a0 + fo);
        x.a[1] = *reinterpret_cast<const pfkl_v4i*>(a1 + fo);
#pragma unroll
        for (int sl = 0; sl < PFKL_SLICES; ++sl) x.b[sl] = *reinterpret_cast<const pfkl_v4i*>(bp + sl * sstride + fo);
    };
    auto mma = [&](const Ops& x) {
#pragma unroll
        for (int sl = 0; sl < PFKL_SLICES; ++sl) {
            acc[0][sl] = __builtin_amdgcn_mfma_i32_32x32x32_i8(x.a[0], x.b[sl], acc[0][sl], 0, 0, 0);
            acc[1][sl] = __builtin_amdgcn_mfma_i32_32x32x32_i8(x.a[1], x.b[sl], acc[1][sl], 0, 0, 0);
        }
    };
    static_assert(PFKL_KS % 4 == 1, "the step count is four-step rounds plus one");
    Ops X, Y, Z, W;
    load(X, 0);
    load(Y, 1);
    load(Z, 2);
#pragma unroll 1
    for (int ks = 0; ks < PFKL_KS - 1; ks += 4) {
        __syncthreads();
        load(W, ks + 3);
        __builtin_amdgcn_sched_barrier(0);
        mma(X);
        __builtin_amdgcn_sched_barrier(0);
        load(X, ks + 4);
        __builtin_amdgcn_sched_barrier(0);
        mma(Y);
        __builtin_amdgcn_sched_barrier(0);
        load(Y, ks + 5);
        __builtin_amdgcn_sched_barrier(0);
        mma(Z);
        __builtin_amdgcn_sched_barrier(0);
        load(Z, ks + 6);
        __builtin_amdgcn_sched_barrier(0);
        mma(W);
        __builtin_amdgcn_sched_barrier(0);
    }
    mma(X);  // the last step, PFKL_KS - 1
    uint32_t* o = out + (size_t)uu * out_u_offset + col0 + (lane & 31);
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const uint32_t m = m0 + 32 * r + 8 * (j / 4) + 4 * (lane >> 5) + (j & 3);
            if (m < B) {
                int64_t tsum = 0;
#pragma unroll
                for (int sl = 0; sl < PFKL_SLICES; ++sl) tsum += (int64_t)acc[r][sl][j] * ((int64_t)1 << (8 * sl));
                const uint64_t o40 = (0 - (uint64_t)tsum) & PFKL_MASK;
                o[(size_t)m * out_stride] = (uint32_t)((o40 + 128) >> 8);
            }
        }
}

// The same product with the operands staged through LDS, for launches of many input tiles:
// 8 waves per workgroup, 256 inputs (eight 32-row blocks) x 64 columns (two 32-column
// blocks) x 5 slices; wave w computes row blocks 2 (w % 4) + 0, 1 of column block w / 4, as
// a wave of k_pfkl_gemm does.  Each step's 18 fragments (8 digit blocks, 2 x 5 key blocks,
// 1 KB each) are read from global memory once per workgroup: wave w stages fragments w,
// w + 8 and (waves 0, 1) w + 16 through registers three steps ahead, into a ring of three
// LDS buffers (54 KB).  Step t stores step t + 1's fragments into buffer (t + 1) % 3, whose
// last reads (step t - 2's operands, read at step t - 3) are complete at step t - 2's
// barrier, then passes its own barrier and reads step t + 1's operands, which are in flight
// while step t's products issue: one barrier per step.  The operands live in LDS, so a wave
// holds its 160 accumulators and about 90 VGPRs, and a CU holds two waves per SIMD.  Same
// sums, same words.
__global__ __launch_bounds__(512, 1) void k_pfkl_gemm_lds(const int8_t* __restrict__ D, uint32_t B,
                                                      const int8_t* __restrict__ KT, int u0,
                                                      uint32_t* __restrict__ out, size_t out_stride,
                                                      size_t out_u_offset) {
    constexpr int COLS = 2 * CB_N, NF = 8 + 2 * PFKL_SLICES, NBUF = 3;
    static_assert(PFKL_KS % NBUF == 0, "the step count splits into ring rounds");
    __shared__ pfkl_v4i ring[NBUF][NF][64];
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t m0 = blockIdx.x * 256;
    const uint32_t cg = blockIdx.y * 64;
    const int uu = (int)(cg / COLS), col0 = (int)(cg % COLS);
    const size_t sstride = (size_t)COLS * PFKL_KP;
    // fragment f at step 0 (a step advances the pointer by 1 KB): f < 8 digit row block f,
    // else key column block (f - 8) / 5, slice (f - 8) % 5
    auto src = [&](int f) -> const pfkl_v4i* {
        const int8_t* p;
        if (f < 8) {
            p = D + ((size_t)(m0 / 32 + f) * PFKL_KS * 64 + lane) * 16;
        } else {
            const int j = (f - 8) / PFKL_SLICES, sl = (f - 8) % PFKL_SLICES;
            p = KT + (size_t)(u0 + uu) * PFKL_SLICES * sstride + (size_t)sl * sstride +
                ((size_t)(col0 / 32 + j) * PFKL_KS * 64 + lane) * 16;
        }
        return reinterpret_cast<const pfkl_v4i*>(p);
    };
    const bool third = w + 16 < NF;  // waves 0 and 1 stage a third fragment
    const pfkl_v4i* g0 = src(w);
    const pfkl_v4i* g1 = src(w + 8);
    const pfkl_v4i* g2 = src(third ? w + 16 : w);
    pfkl_v4i st[NBUF][3];
    auto fetch = [&](pfkl_v4i(&r)[3], int ks) {  // step ks (past the last: step 0, unused)
        const size_t fo = (size_t)(ks < PFKL_KS ? ks : 0) * 64;
        r[0] = g0[fo];
        r[1] = g1[fo];
        if (third) r[2] = g2[fo];
    };
    auto stage = [&](const pfkl_v4i(&r)[3], int b) {
        ring[b][w][lane] = r[0];
        ring[b][w + 8][lane] = r[1];
        if (third) ring[b][w + 16][lane] = r[2];
    };
    pfkl_v16i acc[2][PFKL_SLICES];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int sl = 0; sl < PFKL_SLICES; ++sl) acc[r][sl] = pfkl_v16i{};
    const int ra = 2 * (w & 3), cb = 8 + (w >> 2) * PFKL_SLICES;
    struct Opnd { pfkl_v4i a[2], b[PFKL_SLICES]; };
    Opnd op[2];
    auto read = [&](Opnd& x, int b) {
        x.a[0] = ring[b][ra][lane];
        x.a[1] = ring[b][ra + 1][lane];
#pragma unroll
        for (int sl = 0; sl < PFKL_SLICES; ++sl) x.b[sl] = ring[b][cb + sl][lane];
    };
    auto mma = [&](const Opnd& x) {
#pragma unroll
        for (int sl = 0; sl < PFKL_SLICES; ++sl) {
            acc[0][sl] = __builtin_amdgcn_mfma_i32_32x32x32_i8(x.a[0], x.b[sl], acc[0][sl], 0, 0, 0);
            acc[1][sl] = __builtin_amdgcn_mfma_i32_32x32x32_i8(x.a[1], x.b[sl], acc[1][sl], 0, 0, 0);
        }
    };
    // step s (s % 6 = u, so that the register set, buffer and operand indices are constants):
    // stage step s + 1, fetch step s + 4, barrier, read step s + 1's operands (staged before
    // this barrier, so they are in flight during step s's products), products of step s
    auto step = [&](int u, int s) {
        const int nb = (u + 1) % NBUF;
        stage(st[nb], nb);
        fetch(st[nb], s + 1 + NBUF);
        __syncthreads();
        read(op[(u + 1) & 1], nb);
        mma(op[u & 1]);
    };
    fetch(st[0], 0);
    fetch(st[1], 1);
    fetch(st[2], 2);
    stage(st[0], 0);
    fetch(st[0], 3);
    __syncthreads();
    read(op[0], 0);
    static_assert(PFKL_KS % 6 == 3, "six-step rounds and a tail of three");
#pragma unroll 1
    for (int t = 0; t < PFKL_KS - 3; t += 6) {
#pragma unroll
        for (int u = 0; u < 6; ++u) step(u, t + u);
    }
#pragma unroll
    for (int u = 0; u < 3; ++u) step(u, PFKL_KS - 3 + u);
    const uint32_t mw = m0 + 64 * (w & 3);
    uint32_t* o = out + (size_t)uu * out_u_offset + col0 + 32 * (w >> 2) + (lane & 31);
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const uint32_t m = mw + 32 * r + 8 * (j / 4) + 4 * (lane >> 5) + (j & 3);
            if (m < B) {
                int64_t tsum = 0;
#pragma unroll
                for (int sl = 0; sl < PFKL_SLICES; ++sl) tsum += (int64_t)acc[r][sl][j] * ((int64_t)1 << (8 * sl));
                const uint64_t o40 = (0 - (uint64_t)tsum) & PFKL_MASK;
                o[(size_t)m * out_stride] = (uint32_t)((o40 + 128) >> 8);
            }
        }
}

// ===========================================================================
// N -> n key switch with a linear key on the matrix cores (the switch of
// aes_gpu_ks.cuh:341-370, base 2^2 x 8 on the rounded top 16 bits, with one key row
// per (i, j) and signed digits instead of the lookup key's row per digit value):
//   out[m] = (0, b_m) - sum_{k = 8 i + j} d[m][k] KL[k]   (mod 2^32),
//   d_j = ((a~ + 0xAAAA) >> 2 j mod 4) - 2,  KL[i][j] = LWE_s(K2_i 2^(16 + 2 j)).
// The digit sum is a GEMM [inputs x 8N] x [8N x (n + 1)] with k padded to PFKL_KP
// = 513 x 32, in the linear PFKS's fragment geometry (pfkl_frag_off): digits
// D [Mpad / 32][PFKL_KS][1 KB], key KT [4 slices][640 / 32 column blocks][PFKL_KS][1 KB],
// each 32-bit key word held as 4 signed bytes (sum s_b 2^(8 b) = the word as an int32),
// |sum| <= 8N x 2 x 128 < 2^23 per int32 accumulator.  Columns n + 1 .. 639 and
// k >= 8N are zero.
// ===========================================================================
constexpr int KSL_SLICES = 4, KSL_COLS = 640;
static_assert(PFKL_KP >= 8 * CB_N, "padded digit count of the key switch");

// key words kl [N][8][n+1] (= [k][n+1]) -> signed byte slices of KT (zeroed by the caller)
__global__ __launch_bounds__(256) void k_kskl_pack(const uint32_t* __restrict__ kl, int n, int8_t* __restrict__ KT) {
    const size_t total = (size_t)8 * CB_N * (n + 1);
    const size_t sstride = (size_t)KSL_COLS * PFKL_KP;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const uint32_t k = (uint32_t)(e / (size_t)(n + 1));
        const int c = (int)(e - (size_t)k * (n + 1));
        int64_t x = (int32_t)kl[e];
        int8_t* o = KT + pfkl_frag_off((size_t)c, k);
#pragma unroll
        for (int sl = 0; sl < KSL_SLICES; ++sl) {
            const int8_t b = (int8_t)(uint8_t)(x & 255);
            o[sl * sstride] = b;
            x = (x - b) >> 8;
        }
    }
}

// signed base-4 digits of the rounded top 16 bits: D[m][8 q + j] for q < N and m < B,
// zero elsewhere in [Mpad][PFKL_KP]
__global__ __launch_bounds__(256) void k_ksl_digits(const uint32_t* __restrict__ lwe32, uint32_t B, uint32_t Mpad,
                                                   int8_t* __restrict__ D) {
    constexpr int N = CB_N, Q = PFKL_KP / 8;
    const size_t total = (size_t)Mpad * Q;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const size_t m = e / Q;
        const int q = (int)(e - m * Q);
        uint64_t word = 0;
        if (m < B && q < N) {
            const uint32_t at = (uint32_t)(((uint64_t)lwe32[m * (N + 1) + q] + (1u << 15)) >> 16);
            const uint32_t w = at + 0xAAAAu;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int d = (int)((w >> (2 * j)) & 3u) - 2;
                word |= (uint64_t)(uint8_t)(int8_t)d << (8 * j);
            }
        }
        *reinterpret_cast<uint64_t*>(D + pfkl_frag_off(m, 8u * q)) = word;  // 8 bytes of one 16-byte run
    }
}

// Grid x = Mpad / 128 input tiles, y = 640 / 64 column tiles.  Wave w computes 64 inputs
// (two 32-row blocks) x 32 columns x 4 slices: eight 16-register accumulators, three
// operand sets in flight as in k_pfkl_gemm.
// blockIdx.z = K-slice of kstep steps (a multiple of 3): with more than one slice
// (gridDim.z > 1) each adds its partial sum into the zero-filled output with 32-bit atomics
// (the sums are mod 2^32, so the order does not change a word), slice 0 adding b
__global__ __launch_bounds__(256, 1) void k_ksl_gemm(const int8_t* __restrict__ D, uint32_t B,
                                                 const int8_t* __restrict__ KT, const uint32_t* __restrict__ lwe32,
                                                 int n, uint32_t* __restrict__ out, int kstep) {
    constexpr int N = CB_N;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t m0 = blockIdx.x * 128 + (w & 1) * 64;
    const int col0 = (int)blockIdx.y * 64 + (w >> 1) * 32;
    const int8_t* a0 = D + ((size_t)(m0 / 32) * PFKL_KS * 64 + lane) * 16;
    const int8_t* a1 = a0 + (size_t)PFKL_KS * 1024;
    const size_t sstride = (size_t)KSL_COLS * PFKL_KP;
    const int8_t* bp = KT + ((size_t)(col0 / 32) * PFKL_KS * 64 + lane) * 16;
    pfkl_v16i acc[2][KSL_SLICES];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int sl = 0; sl < KSL_SLICES; ++sl) acc[r][sl] = pfkl_v16i{};
    // three operand sets rotate over the steps (513 = 3 x 171), each loaded two steps ahead
    // of its products; loads past the slice's last step reread its first (in bounds, unused)
    const int kbeg = (int)blockIdx.z * kstep, kend = kbeg + kstep;
    struct Ops { pfkl_v4i a[2], b[KSL_SLICES]; };
    auto load = [&](Ops& x, int ks) {
        const size_t fo = (size_t)(ks < kend ? ks : kbeg) * 1024;
        x.a[0] = *reinterpret_cast<const pfkl_v4i*>(a0 + fo);
        x.a[1] = *reinterpret_cast<const pfkl_v4i*>(a1 + fo);
#pragma unroll
        for (int sl = 0; sl < KSL_SLICES; ++sl) x.b[sl] = *reinterpret_cast<const pfkl_v4i*>(bp + sl * sstride + fo);
    };
    auto mma = [&](const Ops& x) {
#pragma unroll
        for (int sl = 0; sl < KSL_SLICES; ++sl) {
            acc[0][sl] = __builtin_amdgcn_mfma_i32_32x32x32_i8(x.a[0], x.b[sl], acc[0][sl], 0, 0, 0);
            acc[1][sl] = __builtin_amdgcn_mfma_i32_32x32x32_i8(x.a[1], x.b[sl], acc[1][sl], 0, 0, 0);
        }
    };
    Ops X, Y, Z;
    load(X, kbeg);
    load(Y, kbeg + 1);
#pragma unroll 1
    for (int ks = kbeg; ks < kend; ks += 3) {
        load(Z, ks + 2);
        __builtin_amdgcn_sched_barrier(0);
        mma(X);
        __builtin_amdgcn_sched_barrier(0);
        load(X, ks + 3);
        __builtin_amdgcn_sched_barrier(0);
        mma(Y);
        __builtin_amdgcn_sched_barrier(0);
        load(Y, ks + 4);
        __builtin_amdgcn_sched_barrier(0);
        mma(Z);
        __builtin_amdgcn_sched_barrier(0);
    }
    const int col = col0 + (lane & 31);
    if (col > n) return;
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const uint32_t m = m0 + 32 * r + 8 * (j / 4) + 4 * (lane >> 5) + (j & 3);
            if (m < B) {
                uint32_t t = 0;
#pragma unroll
                for (int sl = 0; sl < KSL_SLICES; ++sl) t += (uint32_t)acc[r][sl][j] << (8 * sl);
                uint32_t v = 0u - t;
                if (col == n && blockIdx.z == 0) v += lwe32[(size_t)m * (N + 1) + N];
                if (gridDim.z > 1)
                    atomicAdd(&out[(size_t)m * (n + 1) + col], v);
                else
                    out[(size_t)m * (n + 1) + col] = v;
            }
        }
}

// ===========================================================================
// Vertical-packing LUT on the 32-bit torus (reference CMux semantics,
// rlwe_tables.cuh:79-108): acc = (0, table); for bit i = 0..7:
// acc = CMux(C_i, acc, X^{-stride 2^i} acc) = acc + C_i (*) (X^{-stride 2^i} acc - acc),
// gadget 2^8 x 2 (16 bits dropped).  Each digit polynomial times a torus-32
// RGSW row is one exact FP64 product.  Then coefficients 0..out_bits-1 are
// sample-extracted: out [lut][out_bits][N+1] u32, ready for the N -> n key
// switch.  rgsw_f [lut][8][4 rows][2 out][1024] cplx.
// ===========================================================================
constexpr int CB_LOGB = 8, CB_L = 2, CB_ROWS = 4, CB_IGN = 16;

__device__ __forceinline__ uint32_t cb_digit_offset() {
    uint32_t off = 1u << (CB_IGN - 1);
#pragma unroll
    for (int j = 0; j < CB_L; ++j) off += (1u << (CB_LOGB - 1)) << (CB_IGN + j * CB_LOGB);
    return off;
}

__global__ __launch_bounds__(256) void k_lut_vertical(const cplx* __restrict__ rgsw_f, uint32_t nlut,
                                                     const uint32_t* __restrict__ table, size_t table_stride,
                                                     int stride, int out_bits, int parts, int tperiod,
                                                     FFT1024Tables Tg, uint32_t* __restrict__ out) {
    __shared__ cplx tl[768];
    __shared__ cplx lds_all[4][F256_LDS_CPLX];
    __shared__ cplx xbuf[2][1024];
    const FFT1024Tables T = stage_tables(Tg, tl);
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    cplx trw[4];
    load_trw(Tg, w, lane, trw);
    cplx* lds = lds_all[w];
    uint32_t* x32 = reinterpret_cast<uint32_t*>(&xbuf[0][0]);
    const uint32_t OFF = cb_digit_offset();
    constexpr uint32_t N = CB_N;

    for (uint32_t lut = blockIdx.x; lut < nlut; lut += gridDim.x) {
        // parts > 1: `parts` table polynomials share one group of 8 RGSWs (an 8 -> 8 parts
        // bit LUT); LUT lut reads group lut / parts and table polynomial lut % parts, of
        // table set (group mod tperiod) when tperiod > 0 (the T-table round: one table per
        // state row, byte = row + 4 col)
        const uint32_t group = parts > 1 ? lut / (uint32_t)parts : lut;
        const uint32_t tpoly = parts > 1 ? lut % (uint32_t)parts : lut;
        const uint32_t tset = tperiod > 0 ? (group % (uint32_t)tperiod) * (uint32_t)parts : 0u;
        const uint32_t* tab = table + (size_t)(tset + tpoly) * table_stride;
        uint32_t acc[2][8];
#pragma unroll
        for (int h = 0; h < 8; ++h) {
            acc[0][h] = 0;
            acc[1][h] = tab[own_coeff(w, lane, h & 3) + (h >= 4 ? CB_M : 0)];
        }
        int buf = 0;
        for (int bit = 0; bit < 8; ++bit) {
            const uint32_t sh = (uint32_t)(stride << bit);
            // diff = X^{-sh} acc - acc, rotation through LDS
            __syncthreads();
#pragma unroll
            for (int o = 0; o < 2; ++o)
#pragma unroll
                for (int h = 0; h < 8; ++h) x32[o * CB_N + own_coeff(w, lane, h & 3) + (h >= 4 ? CB_M : 0)] = acc[o][h];
            __syncthreads();
            uint32_t tp[2][8];
#pragma unroll
            for (int o = 0; o < 2; ++o)
#pragma unroll
                for (int h = 0; h < 8; ++h) {
                    const uint32_t src = (uint32_t)(own_coeff(w, lane, h & 3) + (h >= 4 ? CB_M : 0)) + sh;
                    uint32_t v = x32[o * CB_N + (src & (N - 1))];
                    if ((src / N) & 1) v = 0u - v;
                    tp[o][h] = v - acc[o][h] + OFF;
                }
            __syncthreads();
            const cplx* key = rgsw_f + ((size_t)group * 8 + bit) * CB_ROWS * 2 * CB_M;
            cplx macc[2][4];
#pragma unroll
            for (int o = 0; o < 2; ++o)
#pragma unroll
                for (int s = 0; s < 4; ++s) macc[o][s] = {0.0, 0.0};
#pragma unroll
            for (int u = 0; u < 2; ++u)
#pragma unroll
                for (int j = 0; j < CB_L; ++j) {
                    const int r = 2 * j + u;
                    const int sh2 = CB_IGN + j * CB_LOGB;
                    cplx x[4];
#pragma unroll
                    for (int a = 0; a < 4; ++a)
                        x[a] = {(double)((int32_t)((tp[u][a] >> sh2) & 255u) - 128),
                                (double)((int32_t)((tp[u][a + 4] >> sh2) & 255u) - 128)};
                    f1024_fwd<false>(x, lds, xbuf[buf], T, trw, w, lane);
                    buf ^= 1;
                    const cplx* kr = key + (size_t)r * 2 * CB_M + (w * 4) * 64 + lane;
#pragma unroll
                    for (int o = 0; o < 2; ++o)
#pragma unroll
                        for (int s = 0; s < 4; ++s) cmac(macc[o][s], x[s], kr[o * CB_M + s * 64]);
                }
#pragma unroll
            for (int o = 0; o < 2; ++o) {
                f1024_inv<false>(macc[o], lds, xbuf[buf], T, trw, w, lane);
                buf ^= 1;
#pragma unroll
                for (int a = 0; a < 4; ++a) {
                    acc[o][a] += (uint32_t)round_i64(macc[o][a].re, 1.0 / CB_M);
                    acc[o][a + 4] += (uint32_t)round_i64(macc[o][a].im, 1.0 / CB_M);
                }
            }
        }
        // extract coefficients 0..out_bits-1
        uint32_t* ob = out + (size_t)lut * out_bits * (N + 1);
        for (int idx = 0; idx < out_bits; ++idx) {
            uint32_t* o = ob + (size_t)idx * (N + 1);
#pragma unroll
            for (int h = 0; h < 8; ++h) {
                const uint32_t c = (uint32_t)own_coeff(w, lane, h & 3) + (h >= 4 ? CB_M : 0);
                o[((uint32_t)idx - c) & (N - 1)] = (c <= (uint32_t)idx) ? acc[0][h] : 0u - acc[0][h];
            }
            if (lane == idx / 4 && w == idx % 4) o[N] = acc[1][0];
        }
    }
}

// ===========================================================================
// AES transciphering linear pieces on level-0 LWE bits (b * 2^31):
// state init  = Enc(rk0 bit) + (0, counter bit * 2^31)  (reference
//               kernel_InitLWEStateFromRoundKey, gpu_lut.cuh:425-452)
// linear layer = ShiftRows + MixColumns + AddRoundKey as LWE additions:
//               out[bit] = rk[bit] + sum_{s in src[bit]} in[s]  (XOR = addition
//               of p = 2 encodings, rlwe_tables.cuh:26-35)
// ===========================================================================
__global__ void k_aes_init(const uint32_t* __restrict__ rk0, const uint8_t* __restrict__ ctr, uint32_t blocks, int n,
                           uint32_t* __restrict__ state) {
    const uint32_t blk = blockIdx.y, bit = blockIdx.x;
    if (blk >= blocks) return;
    const uint32_t pb = (ctr[blk * 16 + bit / 8] >> (bit % 8)) & 1u;
    const uint32_t* src = rk0 + (size_t)bit * (n + 1);
    uint32_t* dst = state + ((size_t)blk * 128 + bit) * (n + 1);
    for (int k = threadIdx.x; k <= n; k += blockDim.x) dst[k] = src[k] + (k == n ? (pb << 31) : 0u);
}

__global__ void k_aes_linear(const uint32_t* __restrict__ in, const uint32_t* __restrict__ rk,
                             const int16_t* __restrict__ src, int nsrc, uint32_t blocks, int n,
                             uint32_t* __restrict__ out) {
    const uint32_t blk = blockIdx.y, bit = blockIdx.x;
    if (blk >= blocks) return;
    const uint32_t* base = in + (size_t)blk * 128 * (n + 1);
    uint32_t* dst = out + ((size_t)blk * 128 + bit) * (n + 1);
    for (int k = threadIdx.x; k <= n; k += blockDim.x) {
        uint32_t v = rk[(size_t)bit * (n + 1) + k];
        for (int q = 0; q < nsrc; ++q) {
            const int s = src[bit * nsrc + q];
            if (s >= 0) v += base[(size_t)s * (n + 1) + k];
        }
        dst[k] = v;
    }
}

// T-table round at dimension N (the reference's word XOR, kernel_RLWE_AddInplace_Batch with
// its word map, main.cu:1480-1523): output bit 8 (p + 4 c) + j of a block (row p, column c)
// is the sum of the four LUT outputs lutw [blocks 16 bytes][32 bits][N+1], bit 8 p + j, of
// the input bytes (r, c + r) (encryption, ShiftRows) or (r, c - r) (decryption, InvShiftRows),
// r = 0..3: the XOR of the four T-table words' bits.
__global__ void k_aes_ttable_xor(const uint32_t* __restrict__ lutw, uint32_t blocks, int dec,
                                 uint32_t* __restrict__ out) {
    constexpr uint32_t W = CB_N + 1;
    const uint32_t blk = blockIdx.y, ob = blockIdx.x;
    if (blk >= blocks) return;
    const uint32_t byte = ob >> 3, j = ob & 7, p = byte & 3, c = byte >> 2;
    const uint32_t* src[4];
#pragma unroll
    for (uint32_t r = 0; r < 4; ++r) {
        const uint32_t cin = dec ? (c + 4 - r) & 3 : (c + r) & 3;
        src[r] = lutw + (((size_t)blk * 16 + r + 4 * cin) * 32 + 8 * p + j) * W;
    }
    uint32_t* dst = out + ((size_t)blk * 128 + ob) * W;
    for (uint32_t k = threadIdx.x; k < W; k += blockDim.x) dst[k] = src[0][k] + src[1][k] + src[2][k] + src[3][k];
}

// public bits XOR: ct[bit] += (0, pub_bit * 2^31)
__global__ void k_xor_public(uint32_t* __restrict__ ct, const uint8_t* __restrict__ pub, uint32_t nbits, int n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nbits) return;
    ct[(size_t)i * (n + 1) + n] += ((uint32_t)((pub[i / 8] >> (i % 8)) & 1u)) << 31;
}

// ===========================================================================
// Host launchers
// ===========================================================================
static unsigned grid_for(size_t work, size_t cap) { return (unsigned)(work < cap ? work : cap); }

void launch_lut_vertical(const cplx* rgsw_f, uint32_t nlut, const uint32_t* table, size_t table_stride, int stride,
                         int out_bits, const FFT1024Tables& T, uint32_t* out, hipStream_t st, int parts, int tperiod) {
    k_lut_vertical<<<grid_for(nlut, 2048), 256, 0, st>>>(rgsw_f, nlut, table, table_stride, stride, out_bits, parts,
                                                           tperiod, T, out);
}

void launch_htss(const uint64_t* acc, uint32_t batch, const cplx* htkey, const cplx* sskey, const FFT1024Tables& T,
                 uint32_t* rgsw, hipStream_t st) {
    k_htss<<<grid_for(2 * batch, 2048), 256, 0, st>>>(acc, batch, htkey, sskey, T, rgsw);
}

void launch_aes_ttable_xor(const uint32_t* lutw, uint32_t blocks, int dec, uint32_t* out, hipStream_t st) {
    k_aes_ttable_xor<<<dim3(128, blocks), 256, 0, st>>>(lutw, blocks, dec, out);
}

void launch_aes_init(const uint32_t* rk0, const uint8_t* ctr, uint32_t blocks, int n, uint32_t* state, hipStream_t st) {
    k_aes_init<<<dim3(128, blocks), 256, 0, st>>>(rk0, ctr, blocks, n, state);
}

void launch_aes_linear(const uint32_t* in, const uint32_t* rk, const int16_t* src, int nsrc, uint32_t blocks, int n,
                       uint32_t* out, hipStream_t st) {
    k_aes_linear<<<dim3(128, blocks), 256, 0, st>>>(in, rk, src, nsrc, blocks, n, out);
}

void launch_xor_public(uint32_t* ct, const uint8_t* pub, uint32_t nbits, int n, hipStream_t st) {
    k_xor_public<<<(nbits + 255) / 256, 256, 0, st>>>(ct, pub, nbits, n);
}

int pbs_model_form(uint32_t batch, int limbs, int mb) {
    // Two limbs: the branch-free FFT form needs 267 registers (one wave per SIMD),
    // which pays only up to 256 ciphertexts; beyond, the branchy form keeps two waves
    // per SIMD.  One limb: the branch-free form fits two waves per SIMD (220
    // registers) and was the faster one at every size measured.  Two key bits per
    // step (branch-free FFT): 2 = split over 8 waves, for launches of at most one
    // workgroup per CU (128 inputs 4.67 ms against 6.20 wide and 6.47 at two waves per
    // SIMD; 256 inputs 5.08 against 6.22), else 0 = two waves per SIMD (2176 inputs 37.9
    // against 43.8 split and 54.7 wide); 1 = the wide key batch at one wave per SIMD.
    if (mb == 2) return batch <= 256 ? 2 : 0;
    if (limbs == 2) return batch <= 256 ? 1 : 0;
    return 1;
}

template <bool LUT>
static void launch_pbs2_t(const uint32_t* lwe, uint32_t batch, int n, const cplx* bskf, int limbs, int mb,
                          const FFT1024Tables& T, uint64_t* out, hipStream_t st, int form, const PbsLut& io) {
    // one workgroup per ciphertext; form: 1 = branch-free cross-wave FFT step,
    // 0 = branchy (two key bits per step: 1 = wide key batch, 0 = two waves per
    // SIMD, 2 = split over 8 waves), < 0 = pbs_model_form's choice
    const unsigned big = grid_for(batch, 2048);
    if (form < 0) form = pbs_model_form(batch, limbs, mb);
    if constexpr (LUT) {
        if (limbs == 2 && io.logB == 10) {  // the HomTrace + SchemeSwitch form's 2^10 x 4 gadget
            if (form == 1)
                k_pbs2_manylut<true, 2, 1, false, true, 10><<<batch <= 256 ? batch : big, 256, 0, st>>>(
                    lwe, batch, n, bskf, T, out, io);
            else
                k_pbs2_manylut<false, 2, 1, false, true, 10><<<big, 256, 0, st>>>(lwe, batch, n, bskf, T, out, io);
            return;
        }
    }
    if (limbs == 2) {
        if (form == 1)
            k_pbs2_manylut<true, 2, 1, false, LUT><<<batch <= 256 ? batch : big, 256, 0, st>>>(lwe, batch, n, bskf, T,
                                                                                               out, io);
        else
            k_pbs2_manylut<false, 2, 1, false, LUT><<<big, 256, 0, st>>>(lwe, batch, n, bskf, T, out, io);
    } else if (mb == 1) {
        if (form == 0)
            k_pbs2_manylut<false, 1, 1, false, LUT><<<big, 256, 0, st>>>(lwe, batch, n, bskf, T, out, io);
        else
            k_pbs2_manylut<true, 1, 1, false, LUT><<<big, 256, 0, st>>>(lwe, batch, n, bskf, T, out, io);
    } else if (form == 1) {
        k_pbs2_manylut<true, 1, 2, true, LUT><<<batch <= 256 ? batch : big, 256, 0, st>>>(lwe, batch, n, bskf, T, out,
                                                                                          io);
    } else if (form == 2) {
        k_pbs2_split<LUT><<<batch <= 256 ? batch : big, 512, 0, st>>>(lwe, batch, n, bskf, T, out, io);
    } else {
        k_pbs2_manylut<true, 1, 2, false, LUT><<<big, 256, 0, st>>>(lwe, batch, n, bskf, T, out, io);
    }
}

void launch_pbs2(const uint32_t* lwe, uint32_t batch, int n, const cplx* bskf, int limbs, int mb,
                 const FFT1024Tables& T, uint64_t* out, hipStream_t st, int form, const PbsLut& io) {
    if (io.lut)
        launch_pbs2_t<true>(lwe, batch, n, bskf, limbs, mb, T, out, st, form, io);
    else
        launch_pbs2_t<false>(lwe, batch, n, bskf, limbs, mb, T, out, st, form, io);
}

uint32_t pfks_bpad(uint32_t B) { return (B + PF_GROUP - 1) / PF_GROUP * PF_GROUP; }

void launch_pfks_digits(const uint64_t* lwe64, uint32_t B, uint32_t* at, hipStream_t st) {
    const uint32_t Bpad = pfks_bpad(B);
    k_pfks_digits<<<dim3((CB_N + 1 + 63) / 64, Bpad / 64), 256, 0, st>>>(lwe64, B, Bpad, at);
}

void launch_fill_hash32(uint32_t* p, size_t count, uint32_t seed, hipStream_t st) {
    k_fill_hash32<<<grid_for((count + 255) / 256, 8192), 256, 0, st>>>(p, count, seed);
}

void launch_pack_lwe64(const uint64_t* a, const uint64_t* b, uint32_t batch, int n, uint32_t a_stride, uint32_t* out,
                       hipStream_t st) {
    const size_t total = (size_t)batch * (n + 1);
    k_pack_lwe64<<<grid_for((total + 255) / 256, 8192), 256, 0, st>>>(a, b, batch, n, a_stride, out);
}

void launch_zero_rows32(uint32_t* out, uint32_t count, size_t stride, uint32_t width, hipStream_t st) {
    const size_t total = (size_t)count * width;
    k_zero_rows32<<<grid_for((total + 255) / 256, 8192), 256, 0, st>>>(out, count, stride, width);
}

int pfks_model_slices(uint32_t B, int nu) {
    // i-slices (4..32): the launch runs in rounds of as many workgroups as fit at once
    // (two per CU: 92 VGPRs and 8 waves each), and a workgroup's time is its slice
    // length, so pick the slice count with the fewest rounds x coefficients per
    // slice; ties go to fewer slices (each slice is one more atomic pass over the outputs)
    const unsigned tiles = 2 * CB_N / 128, groups = pfks_bpad(B) / PF_GROUP;
    const unsigned slots = 2 * device_cus(), per_slice = tiles * groups * (unsigned)nu;
    int slices = 4;
    unsigned long long best = ~0ull;
    for (unsigned sl = 4; sl <= 32; ++sl) {
        const unsigned long long rounds = ((unsigned long long)per_slice * sl + slots - 1) / slots;
        const unsigned long long cost = rounds * ((CB_N + 1 + sl - 1) / sl);
        if (cost < best) {
            best = cost;
            slices = (int)sl;
        }
    }
    return slices;
}

void launch_pfks32(const uint32_t* at, uint32_t B, const uint32_t* key, size_t key_u_stride, int u0, int nu,
                   uint32_t* out, size_t out_stride, size_t out_u_offset, hipStream_t st, int slices) {
    if (slices <= 0) slices = pfks_model_slices(B, nu);
    if (slices > CB_N + 1) slices = CB_N + 1;
    const unsigned tiles = 2 * CB_N / 128, groups = pfks_bpad(B) / PF_GROUP;
    dim3 grid(tiles, groups, (unsigned)(slices * nu));
    k_pfks32<<<grid, 64 * PF_WAVES, 0, st>>>(at, B, pfks_bpad(B), key, key_u_stride, u0, slices, out, out_stride,
                                              out_u_offset);
}

size_t pfkl_key_bytes() { return (size_t)2 * PFKL_SLICES * 2 * CB_N * PFKL_KP; }
size_t pfkl_key_offset(int u, int slice, int col, uint32_t k) {
    return ((size_t)u * PFKL_SLICES + slice) * 2 * CB_N * PFKL_KP + pfkl_frag_off((size_t)col, k);
}
size_t pfkl_digit_bytes(uint32_t rows) { return (size_t)((rows + 255) / 256 * 256) * PFKL_KP; }
uint32_t pfkl_kp() { return PFKL_KP; }

void launch_gen_pfkl40(uint64_t seed, const int32_t* K2, int8_t* KT, hipStream_t st) {
    k_gen_pfkl40<<<2 * (CB_N + 1) * 8, 256, 0, st>>>(seed, K2, KT);
}

void launch_pfkl(const uint64_t* lwe64, uint32_t B, int8_t* D, const int8_t* KT, int u0, int nu, uint32_t* out,
                 size_t out_stride, size_t out_u_offset, hipStream_t st) {
    // many input tiles: the LDS-staged form (256-input tiles, one 8-wave workgroup per CU);
    // few: 128-input tiles, for twice the workgroups
    const bool staged = B > 1792;
    const uint32_t Mpad = staged ? (B + 255) / 256 * 256 : (B + 127) / 128 * 128;
    k_pfkl_digits<<<grid_for((size_t)Mpad * (PFKL_KP / 8) / 256 + 1, 8192), 256, 0, st>>>(lwe64, B, Mpad, D);
    const unsigned rows = Mpad / 128, tiles = (unsigned)nu * 2 * CB_N / 64;
    if (staged)
        k_pfkl_gemm_lds<<<dim3(Mpad / 256, tiles), 512, 0, st>>>(D, B, KT, u0, out, out_stride, out_u_offset);
    else if (rows <= 2)
        k_pfkl_gemm<<<dim3(tiles, rows), 256, 0, st>>>(D, B, KT, u0, out, out_stride, out_u_offset, 1);
    else
        k_pfkl_gemm<<<dim3(rows, tiles), 256, 0, st>>>(D, B, KT, u0, out, out_stride, out_u_offset, 0);
}

size_t kskl_key_bytes() { return (size_t)KSL_SLICES * KSL_COLS * PFKL_KP; }

int launch_kskl_pack(const uint32_t* kl, int n, int8_t* KT, hipStream_t st) {
    if (n + 1 > KSL_COLS) return -1;
    k_kskl_pack<<<grid_for(((size_t)8 * CB_N * (n + 1) + 255) / 256, 16384), 256, 0, st>>>(kl, n, KT);
    return 0;
}

int launch_ks_linear(const uint32_t* lwe32, uint32_t count, int n, const int8_t* KT, int8_t* D, uint32_t chunk,
                     uint32_t* out, hipStream_t st) {
    if (n + 1 > KSL_COLS || chunk == 0) return -1;
    for (uint32_t off = 0; off < count; off += chunk) {
        const uint32_t B = count - off < chunk ? count - off : chunk;
        const uint32_t Mpad = (B + 127) / 128 * 128;
        const uint32_t* in = lwe32 + (size_t)off * (CB_N + 1);
        k_ksl_digits<<<grid_for((size_t)Mpad * (PFKL_KP / 8) / 256 + 1, 8192), 256, 0, st>>>(in, B, Mpad, D);
        // small batches: the K steps split over 3, 9, 19 or 57 workgroups (513 = 27 x 19), so
        // that at least 160 workgroups stream the key (128 inputs: 10 column tiles x 19)
        const unsigned tiles = (Mpad / 128) * (KSL_COLS / 64);
        int kslices = 1;
        for (int ks : {1, 3, 9, 19, 57}) {
            kslices = ks;
            if (tiles * ks >= 160) break;
        }
        uint32_t* o = out + (size_t)off * (n + 1);
        if (kslices > 1) launch_zero_rows32(o, B, n + 1, n + 1, st);
        k_ksl_gemm<<<dim3(Mpad / 128, KSL_COLS / 64, kslices), 256, 0, st>>>(D, B, KT, in, n, o, PFKL_KS / kslices);
    }
    return 0;
}

void launch_to_fourier2048(const uint64_t* coef, size_t npoly, const FFT1024Tables& T, cplx* out, hipStream_t st) {
    k_to_fourier2048<<<grid_for(npoly, 4096), 256, 0, st>>>(coef, npoly, T, out);
}

void launch_to_fourier2048_t64(const uint64_t* coef, size_t npoly, const FFT1024Tables& T, cplx* out, hipStream_t st) {
    k_to_fourier2048_t64<<<grid_for(npoly, 8192), 256, 0, st>>>(coef, npoly, T, out);
}

void launch_to_fourier2048_32(const uint32_t* coef, size_t npoly, const FFT1024Tables& T, cplx* out, hipStream_t st) {
    k_to_fourier2048_32<<<grid_for(npoly, 8192), 256, 0, st>>>(coef, npoly, T, out);
}

void launch_polymul2048(const int32_t* x, const uint64_t* y, uint32_t batch, const FFT1024Tables& T, uint64_t* out,
                        hipStream_t st, int limbs) {
    if (limbs == 2)
        k_polymul2048<2><<<grid_for(batch, 4096), 256, 0, st>>>(x, y, batch, T, out);
    else
        k_polymul2048<1><<<grid_for(batch, 4096), 256, 0, st>>>(x, y, batch, T, out);
}

void launch_gen_bsk64(uint64_t seed, int n, int nshift, const int32_t* s, const int32_t* K2, const cplx* K2f,
                      const FFT1024Tables& T, uint64_t* bsk, hipStream_t st, int mb, int logB) {
    (void)K2;
    const uint32_t rows = (uint32_t)(mb == 1 ? n : n / 2 * 3) * 8;
    KeyRows kr;
    kr.s = s;
    kr.mb = mb;
    kr.logB = logB;
    k_gen_bsk64<<<grid_for(rows, 8192), 256, 0, st>>>(seed, rows, nshift, kr, K2f, T, bsk);
}

void launch_gen_ks64(uint64_t seed, uint32_t dom, uint32_t npoly, int nshift, const int64_t* msgpoly, const cplx* K2f,
                     const FFT1024Tables& T, uint64_t* key, hipStream_t st) {
    const uint32_t rows = npoly * KS_L;
    KeyRows kr;
    kr.msgpoly = msgpoly;
    kr.dom = dom;
    k_gen_bsk64<<<grid_for(rows, 8192), 256, 0, st>>>(seed, rows, nshift, kr, K2f, T, key);
}

void launch_gen_pfks(uint64_t seed, int nshift, const int32_t* K2, const cplx* K2f, const FFT1024Tables& T,
                     uint32_t* key, hipStream_t st) {
    const uint32_t rows = 2u * (CB_N + 1) * 8 * 15;
    k_gen_pfks32<<<grid_for(rows, 8192), 256, 0, st>>>(seed, rows, nshift, K2, K2f, T, key);
}

}  // namespace tfhe
