// fft1024.hpp -- negacyclic FP64 FFT for N = 2048 (M = 1024 complex points)
// distributed over the four wavefronts of a workgroup.
//
// Decimation in time by 4: wave r owns the folded points z_j, j = 4m + r
// (m = 64a + b: lane b, slot a < 4), i.e. coefficients 4m + r (real part) and
// 4m + r + 1024 (imaginary part).  Each wave runs an LDS-transposed radix-4
// 256-point FFT on its subsequence (twisted by (w^4)^m, w = exp(i pi / N)),
// multiplies by T_r[k] = w^r W1024^(r k), and one cross-wave radix-4 step gives
// wave q the frequencies Z[k + 256 q].  Frequency layout of a transformed
// polynomial: [q][slot s < 4][lane] with k = k0 + 4e + 16p + 64s for lane
// 16 k0 + 4e + p.  The inverse runs the steps backwards and returns
// 1024 x the folded pairs in the same coefficient ownership, so a polynomial
// never leaves the wave that decomposes it.
#pragma once
#include "device_common.hpp"

namespace tfhe {

constexpr int F256_LDS_CPLX = 320;  // per-wave transpose buffer (bank-staggered rows)

struct FFT1024Tables {
    const cplx* p1t;  // [4 k0][64 b]  (w^4)^b * W256^(b k0)
    const cplx* p2;   // [4 e][64 L]   W64^((L mod 16) e)
    const cplx* p3;   // [4 p][64 L]   W16^((L mod 4) p)
    const cplx* tr;   // [4 r][4 s][64 lane]  w^r W1024^(r k(lane, s))
    const cplx* rot;  // [2N] w^j: frequency k (wave q, slot s, lane) is the polynomial at w^(1 - 4k),
                      // k = (lane >> 4) + 4 ((lane >> 2) & 3) + 16 (lane & 3) + 64 s + 256 q
};

// (w^4)^(64 a) = exp(i pi a / 8), a < 4
__device__ __forceinline__ cplx twist4_const(int a) {
    constexpr double C[4][2] = {{1.0, 0.0},
                                {0.92387953251128674, 0.38268343236508978},
                                {0.70710678118654757, 0.70710678118654746},
                                {0.38268343236508984, 0.92387953251128674}};
    return {C[a][0], C[a][1]};
}

template <bool INV>
__device__ __forceinline__ void dft4(cplx (&x)[4]) {
    cplx a0 = cadd(x[0], x[2]), a1 = csub(x[0], x[2]);
    cplx b0 = cadd(x[1], x[3]), b1 = mul_mi<INV>(csub(x[1], x[3]));
    x[0] = cadd(a0, b0);
    x[2] = csub(a0, b0);
    x[1] = cadd(a1, b1);
    x[3] = csub(a1, b1);
}

__device__ __forceinline__ int f256_a1(int k0, int b) { return k0 * 80 + (b >> 4) * 20 + (b & 15); }
__device__ __forceinline__ int f256_a2(int k0, int e, int d) { return k0 * 80 + e * 20 + d; }
__device__ __forceinline__ int f256_a3(int k0, int e, int p, int h) { return k0 * 80 + e * 20 + p * 5 + h; }

// forward 256-point: x = untwisted folded points (lane b, slot a -> point 64a + b)
// p2 and p3 twiddles of one lane held in registers (RT = true below): p2[e - 1] =
// T.p2[64 e + lane], p3[p - 1] = T.p3[64 p + lane]; they are the same for every
// transform a lane runs, so a kernel with registers to spare loads them once.
struct Tw256Regs {
    cplx p2[3], p3[3];
};

__device__ __forceinline__ Tw256Regs load_tw256(const FFT1024Tables& T, int lane) {
    Tw256Regs r;
#pragma unroll
    for (int e = 1; e < 4; ++e) {
        r.p2[e - 1] = T.p2[e * 64 + lane];
        r.p3[e - 1] = T.p3[e * 64 + lane];
    }
    return r;
}

template <bool RT>
__device__ __forceinline__ void fft256_fwd_t(cplx (&x)[4], cplx* lds, const FFT1024Tables& T, const Tw256Regs& tw,
                                             int lane) {
#pragma unroll
    for (int a = 1; a < 4; ++a) x[a] = cmul(x[a], twist4_const(a));
    dft4<false>(x);
#pragma unroll
    for (int k = 0; k < 4; ++k) x[k] = cmul(x[k], T.p1t[k * 64 + lane]);
#pragma unroll
    for (int k = 0; k < 4; ++k) lds[f256_a1(k, lane)] = x[k];
    wave_sync();
    const int k0 = lane >> 4, d = lane & 15;
#pragma unroll
    for (int c = 0; c < 4; ++c) x[c] = lds[f256_a1(k0, 16 * c + d)];
    wave_sync();
    dft4<false>(x);
#pragma unroll
    for (int e = 1; e < 4; ++e) x[e] = cmul(x[e], RT ? tw.p2[e - 1] : T.p2[e * 64 + lane]);
#pragma unroll
    for (int e = 0; e < 4; ++e) lds[f256_a2(k0, e, d)] = x[e];
    wave_sync();
    const int e = (lane >> 2) & 3, h = lane & 3;
#pragma unroll
    for (int g = 0; g < 4; ++g) x[g] = lds[f256_a2(k0, e, 4 * g + h)];
    wave_sync();
    dft4<false>(x);
#pragma unroll
    for (int p = 1; p < 4; ++p) x[p] = cmul(x[p], RT ? tw.p3[p - 1] : T.p3[p * 64 + lane]);
#pragma unroll
    for (int p = 0; p < 4; ++p) lds[f256_a3(k0, e, p, h)] = x[p];
    wave_sync();
    const int pp = lane & 3;
#pragma unroll
    for (int hh = 0; hh < 4; ++hh) x[hh] = lds[f256_a3(k0, e, pp, hh)];
    wave_sync();
    dft4<false>(x);
}

__device__ __forceinline__ void fft256_fwd(cplx (&x)[4], cplx* lds, const FFT1024Tables& T, int lane) {
    fft256_fwd_t<false>(x, lds, T, Tw256Regs{}, lane);
}

// inverse 256-point: frequency layout in, 256 x folded points out
template <bool RT>
__device__ __forceinline__ void fft256_inv_t(cplx (&x)[4], cplx* lds, const FFT1024Tables& T, const Tw256Regs& tw,
                                             int lane) {
    dft4<true>(x);
    const int k0 = lane >> 4, e = (lane >> 2) & 3, low = lane & 3;
#pragma unroll
    for (int hh = 0; hh < 4; ++hh) lds[f256_a3(k0, e, low, hh)] = x[hh];
    wave_sync();
#pragma unroll
    for (int p = 0; p < 4; ++p) x[p] = lds[f256_a3(k0, e, p, low)];
    wave_sync();
#pragma unroll
    for (int p = 1; p < 4; ++p) x[p] = cmulc(x[p], RT ? tw.p3[p - 1] : T.p3[p * 64 + lane]);
    dft4<true>(x);
#pragma unroll
    for (int g = 0; g < 4; ++g) lds[f256_a2(k0, e, 4 * g + low)] = x[g];
    wave_sync();
    const int d = lane & 15;
#pragma unroll
    for (int ee = 0; ee < 4; ++ee) x[ee] = lds[f256_a2(k0, ee, d)];
    wave_sync();
#pragma unroll
    for (int ee = 1; ee < 4; ++ee) x[ee] = cmulc(x[ee], RT ? tw.p2[ee - 1] : T.p2[ee * 64 + lane]);
    dft4<true>(x);
#pragma unroll
    for (int c = 0; c < 4; ++c) lds[f256_a1(k0, 16 * c + d)] = x[c];
    wave_sync();
#pragma unroll
    for (int k = 0; k < 4; ++k) x[k] = lds[f256_a1(k, lane)];
    wave_sync();
#pragma unroll
    for (int k = 0; k < 4; ++k) x[k] = cmulc(x[k], T.p1t[k * 64 + lane]);
    dft4<true>(x);
#pragma unroll
    for (int a = 1; a < 4; ++a) x[a] = cmulc(x[a], twist4_const(a));
}

__device__ __forceinline__ void fft256_inv(cplx (&x)[4], cplx* lds, const FFT1024Tables& T, int lane) {
    fft256_inv_t<false>(x, lds, T, Tw256Regs{}, lane);
}

// Output w of the 4-point DFT across the four waves of an FFT1024,
// X_w = sum_r y_r W^(r w) with W = -i (forward) or +i (inverse), written as
// X_w = (y0 + e y2) + W^w (y1 + e y3), e = (-1)^w, W^w = c + i t.  The wave index
// is uniform, so a select on it compiles to scalar branches that cut the FFT into
// basic blocks; the constants e, c, t in {0, +-1} make it straight-line FMAs
// with the same rounding points as the add/subtract form (products by 0 and +-1
// are exact).
template <bool INV>
__device__ __forceinline__ cplx xwave_dft4(cplx y0, cplx y1, cplx y2, cplx y3, int w) {
    const double e = (w & 1) ? -1.0 : 1.0;
    const double c = (w == 0) ? 1.0 : (w == 2) ? -1.0 : 0.0;
    const double tf = (w == 1) ? -1.0 : (w == 3) ? 1.0 : 0.0;  // sin(-pi w / 2)
    const double t = INV ? -tf : tf;
    const cplx A = {fma(e, y2.re, y0.re), fma(e, y2.im, y0.im)};
    const cplx B = {fma(e, y3.re, y1.re), fma(e, y3.im, y1.im)};
    return {fma(-t, B.im, fma(c, B.re, A.re)), fma(t, B.re, fma(c, B.im, A.im))};
}

// round(x * scale) mod 2^64 through two FP64 limbs is done by the callers;
// exact int64 of a rounded double with |x| < 2^52.
__device__ __forceinline__ int64_t round_i64(double x, double scale) {
    const double y = fma(x, scale, 6755399441055744.0);  // 1.5 * 2^52
    return __double_as_longlong(y) - 0x4338000000000000ll;
}

}  // namespace tfhe
