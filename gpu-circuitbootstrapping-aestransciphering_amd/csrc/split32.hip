// split32.hip -- split-precision FP32 negacyclic product (N = 2048) of a small-digit
// polynomial by a torus-32 polynomial, the external-product arithmetic of the
// reference's split-FFT backend (limb rule cirbts-init.cu:1096-1120, limb extraction
// cirbts-fft.cu:345-355, recombination cirbts-fft.cu:159-226) carried to FP32.
//
// The 32-bit polynomial y is cut into balanced signed limbs (widths w_k, least
// significant first, sum 32): y = sum_k l_k 2^(o_k) mod 2^32 with
// l_k in [-2^(w_k - 1), 2^(w_k - 1)).  Each limb polynomial is multiplied by x through
// the 4-wave FFT1024 in FP32, the inverse rounded to an integer, and the limbs
// recombined in integer arithmetic mod 2^32.  A limb's product is exact while the
// FP32 rounding error of its transform stays below 1/2, which depends on the
// magnitudes: the rms of a product coefficient is about rms(x) rms(l) sqrt(N), against
// the 24-bit significand.
//   FORM 0: three limbs of 11, 11, 10 bits (the configs' "3 limbs of <= 11 bits").
//   FORM 1: the same 32 bits as 11, 6, 5, 5, 5: every limb above the lowest small
//           enough that its product is exact with overwhelming probability, so that
//           rounding errors stay in the lowest limb (weight 1) instead of reaching the
//           top limbs (weight 2^22 or 2^11).
//   FORM 2: one FP64 product of the whole 32-bit polynomial (exact), the baseline.
#include "fft1024.hpp"
#include "kernels.hpp"

namespace tfhe {

namespace {

constexpr int SP_N = 2048, SP_M = 1024;

template <int FORM>
struct LimbSpec;
template <>
struct LimbSpec<0> {
    static constexpr int L = 3;
    static constexpr int W[3] = {11, 11, 10};
};
template <>
struct LimbSpec<1> {
    static constexpr int L = 5;
    static constexpr int W[5] = {11, 6, 5, 5, 5};
};

// balanced limb k of the 32-bit value v (limbs below k already removed from d)
__device__ __forceinline__ int32_t take_limb(int64_t& d, int w, bool last) {
    const int32_t low = (int32_t)(d & ((1ll << w) - 1));
    const int32_t r = (low << (32 - w)) >> (32 - w);  // sign-extend w bits
    d = last ? 0 : (d - r) >> w;
    return r;
}

template <class C>
__device__ __forceinline__ FFT1024TablesT<C> stage_tables_t(const FFT1024TablesT<C>& Tg, C* tl) {
    for (int k = threadIdx.x; k < 256; k += blockDim.x) {
        tl[k] = Tg.p1t[k];
        tl[256 + k] = Tg.p2[k];
        tl[512 + k] = Tg.p3[k];
    }
    __syncthreads();
    return {tl, tl + 256, tl + 512, Tg.tr, Tg.rot};
}

// FP32 split forms (FORM 0, 1): one block of 4 waves per product
template <int FORM>
__global__ __launch_bounds__(256) void k_polymul2048_split32(const int32_t* __restrict__ x,
                                                            const uint32_t* __restrict__ y, uint32_t batch,
                                                            FFT1024TablesF Tg, uint32_t* __restrict__ out) {
    using LS = LimbSpec<FORM>;
    __shared__ cplxf tl[768];
    __shared__ cplxf lds_all[4][F256_LDS_CPLX];
    __shared__ cplxf xbuf[2][1024];
    const FFT1024TablesF T = stage_tables_t(Tg, tl);
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    cplxf trw[4];
    load_trw(Tg, w, lane, trw);
    int buf = 0;
    for (uint32_t q = blockIdx.x; q < batch; q += gridDim.x) {
        const int32_t* xp = x + (size_t)q * SP_N;
        const uint32_t* yp = y + (size_t)q * SP_N;
        cplxf X[4];
        int64_t dre[4], dim[4];
        uint32_t acc[8];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const int c = own_coeff(w, lane, a);
            X[a] = {(float)xp[c], (float)xp[c + SP_M]};
            dre[a] = (int64_t)yp[c];
            dim[a] = (int64_t)yp[c + SP_M];
            acc[2 * a] = acc[2 * a + 1] = 0u;
        }
        f1024_fwd<false>(X, lds_all[w], xbuf[buf], T, trw, w, lane);
        buf ^= 1;
        int off = 0;
#pragma unroll
        for (int k = 0; k < LS::L; ++k) {
            const int wk = LS::W[k];
            const bool last = k == LS::L - 1;
            cplxf Y[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const int32_t lr = take_limb(dre[a], wk, last);
                const int32_t li = take_limb(dim[a], wk, last);
                Y[a] = {(float)lr, (float)li};
            }
            f1024_fwd<false>(Y, lds_all[w], xbuf[buf], T, trw, w, lane);
            buf ^= 1;
#pragma unroll
            for (int s = 0; s < 4; ++s) Y[s] = cmul(Y[s], X[s]);
            f1024_inv<false>(Y, lds_all[w], xbuf[buf], T, trw, w, lane);
            buf ^= 1;
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                // the inverse returns M x the product; 1/M is exact in FP32
                const int32_t pr = (int32_t)rintf(Y[a].re * (1.0f / SP_M));
                const int32_t pi = (int32_t)rintf(Y[a].im * (1.0f / SP_M));
                acc[2 * a] += (uint32_t)pr << off;
                acc[2 * a + 1] += (uint32_t)pi << off;
            }
            off += wk;
        }
        uint32_t* op = out + (size_t)q * SP_N;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const int c = own_coeff(w, lane, a);
            op[c] = acc[2 * a];
            op[c + SP_M] = acc[2 * a + 1];
        }
    }
}

// FP64 baseline (FORM 2): the whole 32-bit polynomial as signed values, one product
__global__ __launch_bounds__(256) void k_polymul2048_fp64_32(const int32_t* __restrict__ x,
                                                            const uint32_t* __restrict__ y, uint32_t batch,
                                                            FFT1024Tables Tg, uint32_t* __restrict__ out) {
    __shared__ cplx tl[768];
    __shared__ cplx lds_all[4][F256_LDS_CPLX];
    __shared__ cplx xbuf[2][1024];
    const FFT1024Tables T = stage_tables_t(Tg, tl);
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    cplx trw[4];
    load_trw(Tg, w, lane, trw);
    int buf = 0;
    for (uint32_t q = blockIdx.x; q < batch; q += gridDim.x) {
        const int32_t* xp = x + (size_t)q * SP_N;
        const uint32_t* yp = y + (size_t)q * SP_N;
        cplx X[4], Y[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const int c = own_coeff(w, lane, a);
            X[a] = {(double)xp[c], (double)xp[c + SP_M]};
            Y[a] = {(double)(int32_t)yp[c], (double)(int32_t)yp[c + SP_M]};
        }
        f1024_fwd<false>(X, lds_all[w], xbuf[buf], T, trw, w, lane);
        buf ^= 1;
        f1024_fwd<false>(Y, lds_all[w], xbuf[buf], T, trw, w, lane);
        buf ^= 1;
#pragma unroll
        for (int s = 0; s < 4; ++s) Y[s] = cmul(Y[s], X[s]);
        f1024_inv<false>(Y, lds_all[w], xbuf[buf], T, trw, w, lane);
        buf ^= 1;
        uint32_t* op = out + (size_t)q * SP_N;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const int c = own_coeff(w, lane, a);
            op[c] = round_mod32(Y[a].re, 1.0 / SP_M);
            op[c + SP_M] = round_mod32(Y[a].im, 1.0 / SP_M);
        }
    }
}

unsigned split32_grid(uint32_t batch) {
    const unsigned cap = 16 * device_cus();  // several rounds of resident blocks, then a grid-stride loop
    return batch < cap ? batch : cap;
}

}  // namespace

int launch_polymul2048_split32(const int32_t* x, const uint32_t* y, uint32_t batch, const FFT1024TablesF& TF,
                               const FFT1024Tables& T, uint32_t* out, hipStream_t st, int form) {
    if (!batch) return 0;
    const unsigned g = split32_grid(batch);
    switch (form) {
        case 0: k_polymul2048_split32<0><<<g, 256, 0, st>>>(x, y, batch, TF, out); return 0;
        case 1: k_polymul2048_split32<1><<<g, 256, 0, st>>>(x, y, batch, TF, out); return 0;
        case 2: k_polymul2048_fp64_32<<<g, 256, 0, st>>>(x, y, batch, T, out); return 0;
        default: return -1;
    }
}

}  // namespace tfhe
