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

// The transform is written once for a complex type C: cplx (FP64, every form of the
// bootstraps) or cplxf (FP32, the split-precision product); tables hold C values.
template <class C>
struct FFT1024TablesT {
    const C* p1t;  // [4 k0][64 b]  (w^4)^b * W256^(b k0)
    const C* p2;   // [4 e][64 L]   W64^((L mod 16) e)
    const C* p3;   // [4 p][64 L]   W16^((L mod 4) p)
    const C* tr;   // [4 r][4 s][64 lane]  w^r W1024^(r k(lane, s))
    const C* rot;  // [2N] w^j: frequency k (wave q, slot s, lane) is the polynomial at w^(1 - 4k),
                   // k = (lane >> 4) + 4 ((lane >> 2) & 3) + 16 (lane & 3) + 64 s + 256 q
};
using FFT1024Tables = FFT1024TablesT<cplx>;
using FFT1024TablesF = FFT1024TablesT<cplxf>;

// (w^4)^(64 a) = exp(i pi a / 8), a < 4
template <class C = cplx>
__device__ __forceinline__ C twist4_const(int a) {
    using S = decltype(C{}.re);
    constexpr double K[4][2] = {{1.0, 0.0},
                                {0.92387953251128674, 0.38268343236508978},
                                {0.70710678118654757, 0.70710678118654746},
                                {0.38268343236508984, 0.92387953251128674}};
    return {(S)K[a][0], (S)K[a][1]};
}

template <bool INV, class C>
__device__ __forceinline__ void dft4(C (&x)[4]) {
    C a0 = cadd(x[0], x[2]), a1 = csub(x[0], x[2]);
    C b0 = cadd(x[1], x[3]), b1 = mul_mi<INV>(csub(x[1], x[3]));
    x[0] = cadd(a0, b0);
    x[2] = csub(a0, b0);
    x[1] = cadd(a1, b1);
    x[3] = csub(a1, b1);
}

__device__ __forceinline__ int f256_a1(int k0, int b) { return k0 * 80 + (b >> 4) * 20 + (b & 15); }

// The first transpose of the forward 256-point step (and the last of the inverse, its own
// inverse): slot c of lane 16 q + d takes slot q of lane 16 c + d, a 4 x 4 transpose of
// (slot, 16-lane row) done in registers by gfx950's row swaps: v_permlane32_swap exchanges
// rows 2, 3 of one register with rows 0, 1 of another, v_permlane16_swap rows 1, 3 with
// rows 0, 2.  Slots (0, 2), (1, 3) through the first, then (0, 1), (2, 3) through the second.
template <class C>
__device__ __forceinline__ void row_transpose4(C (&x)[4]) {
    constexpr int W = (int)(sizeof(C) / 4);  // dwords per value
    uint32_t d[4][W];
#pragma unroll
    for (int s = 0; s < 4; ++s) __builtin_memcpy(d[s], &x[s], sizeof(C));
#pragma unroll
    for (int w = 0; w < W; ++w) {
        const auto a = __builtin_amdgcn_permlane32_swap(d[0][w], d[2][w], false, false);
        const auto b = __builtin_amdgcn_permlane32_swap(d[1][w], d[3][w], false, false);
        const auto c = __builtin_amdgcn_permlane16_swap(a[0], b[0], false, false);
        const auto e = __builtin_amdgcn_permlane16_swap(a[1], b[1], false, false);
        d[0][w] = c[0];
        d[1][w] = c[1];
        d[2][w] = e[0];
        d[3][w] = e[1];
    }
#pragma unroll
    for (int s = 0; s < 4; ++s) __builtin_memcpy(&x[s], d[s], sizeof(C));
}
__device__ __forceinline__ int f256_a2(int k0, int e, int d) { return k0 * 80 + e * 20 + d; }
__device__ __forceinline__ int f256_a3(int k0, int e, int p, int h) { return k0 * 80 + e * 20 + p * 5 + h; }

// forward 256-point: x = untwisted folded points (lane b, slot a -> point 64a + b)
// p2 and p3 twiddles of one lane held in registers (RT = true below): p2[e - 1] =
// T.p2[64 e + lane], p3[p - 1] = T.p3[64 p + lane]; they are the same for every
// transform a lane runs, so a kernel with registers to spare loads them once.
template <class C>
struct Tw256RegsT {
    C p2[3], p3[3];
};
using Tw256Regs = Tw256RegsT<cplx>;

template <class C>
__device__ __forceinline__ Tw256RegsT<C> load_tw256(const FFT1024TablesT<C>& T, int lane) {
    Tw256RegsT<C> r;
#pragma unroll
    for (int e = 1; e < 4; ++e) {
        r.p2[e - 1] = T.p2[e * 64 + lane];
        r.p3[e - 1] = T.p3[e * 64 + lane];
    }
    return r;
}

template <bool RT, class C>
__device__ __forceinline__ void fft256_fwd_t(C (&x)[4], C* lds, const FFT1024TablesT<C>& T, const Tw256RegsT<C>& tw,
                                             int lane) {
#pragma unroll
    for (int a = 1; a < 4; ++a) x[a] = cmul(x[a], twist4_const<C>(a));
    dft4<false>(x);
#pragma unroll
    for (int k = 0; k < 4; ++k) x[k] = cmul(x[k], T.p1t[k * 64 + lane]);
    row_transpose4(x);
    const int k0 = lane >> 4, d = lane & 15;
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
template <bool RT, class C>
__device__ __forceinline__ void fft256_inv_t(C (&x)[4], C* lds, const FFT1024TablesT<C>& T, const Tw256RegsT<C>& tw,
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
    row_transpose4(x);
#pragma unroll
    for (int k = 0; k < 4; ++k) x[k] = cmulc(x[k], T.p1t[k * 64 + lane]);
    dft4<true>(x);
#pragma unroll
    for (int a = 1; a < 4; ++a) x[a] = cmulc(x[a], twist4_const<C>(a));
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
template <bool INV, class C>
__device__ __forceinline__ C xwave_dft4(C y0, C y1, C y2, C y3, int w) {
    using S = decltype(C{}.re);
    const S e = (w & 1) ? S(-1) : S(1);
    const S c = (w == 0) ? S(1) : (w == 2) ? S(-1) : S(0);
    const S tf = (w == 1) ? S(-1) : (w == 3) ? S(1) : S(0);  // sin(-pi w / 2)
    const S t = INV ? -tf : tf;
    const C A = {sfma(e, y2.re, y0.re), sfma(e, y2.im, y0.im)};
    const C B = {sfma(e, y3.re, y1.re), sfma(e, y3.im, y1.im)};
    return {sfma(-t, B.im, sfma(c, B.re, A.re)), sfma(t, B.re, sfma(c, B.im, A.im))};
}

// Cross-wave step of the FFT1024 (wave w holds frequencies 256 w ..): forward and
// inverse, with the cross twiddles trw taken from registers.  The cross-wave
// radix-4 output of wave w is either selected on w (BF = false: uniform branches,
// which also bound the scheduler's reordering and keep the one-limb / two-limb
// kernels within 256 registers, i.e. two waves per SIMD) or computed branch-free
// (BF = true: fewer instructions and a longer straight-line schedule, but more
// registers; faster where there is one wave per SIMD anyway).  Both round at the
// same points, so the outputs are identical.  xbuf: 1024 C of LDS shared by the
// four waves.
template <bool BF, bool RT = false, class C>
__device__ __forceinline__ void f1024_fwd(C (&x)[4], C* lds, C* xbuf, const FFT1024TablesT<C>& T,
                                          const C (&trw)[4], int w, int lane,
                                          const Tw256RegsT<C>& tw = Tw256RegsT<C>{}) {
    fft256_fwd_t<RT>(x, lds, T, tw, lane);
#pragma unroll
    for (int s = 0; s < 4; ++s) xbuf[(w * 4 + s) * 64 + lane] = cmul(x[s], trw[s]);
    __syncthreads();
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        C y0 = xbuf[(0 * 4 + s) * 64 + lane], y1 = xbuf[(1 * 4 + s) * 64 + lane];
        C y2 = xbuf[(2 * 4 + s) * 64 + lane], y3 = xbuf[(3 * 4 + s) * 64 + lane];
        if constexpr (BF) {
            x[s] = xwave_dft4<false>(y0, y1, y2, y3, w);
        } else {
            C s02 = cadd(y0, y2), d02 = csub(y0, y2), s13 = cadd(y1, y3), d13 = mul_mi<false>(csub(y1, y3));
            x[s] = (w == 0) ? cadd(s02, s13) : (w == 1) ? cadd(d02, d13) : (w == 2) ? csub(s02, s13) : csub(d02, d13);
        }
    }
}

template <bool BF, bool RT = false, class C>
__device__ __forceinline__ void f1024_inv(C (&x)[4], C* lds, C* xbuf, const FFT1024TablesT<C>& T,
                                          const C (&trw)[4], int w, int lane,
                                          const Tw256RegsT<C>& tw = Tw256RegsT<C>{}) {
#pragma unroll
    for (int s = 0; s < 4; ++s) xbuf[(w * 4 + s) * 64 + lane] = x[s];
    __syncthreads();
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        C z0 = xbuf[(0 * 4 + s) * 64 + lane], z1 = xbuf[(1 * 4 + s) * 64 + lane];
        C z2 = xbuf[(2 * 4 + s) * 64 + lane], z3 = xbuf[(3 * 4 + s) * 64 + lane];
        C y;
        if constexpr (BF) {
            y = xwave_dft4<true>(z0, z1, z2, z3, w);
        } else {
            C s02 = cadd(z0, z2), d02 = csub(z0, z2), s13 = cadd(z1, z3), d13 = mul_mi<true>(csub(z1, z3));
            y = (w == 0) ? cadd(s02, s13) : (w == 1) ? cadd(d02, d13) : (w == 2) ? csub(s02, s13) : csub(d02, d13);
        }
        x[s] = cmulc(y, trw[s]);
    }
    fft256_inv_t<RT>(x, lds, T, tw, lane);
}

template <class C>
__device__ __forceinline__ void load_trw(const FFT1024TablesT<C>& Tg, int w, int lane, C (&trw)[4]) {
#pragma unroll
    for (int s = 0; s < 4; ++s) trw[s] = Tg.tr[(w * 4 + s) * 64 + lane];
}

// coefficient owned by (wave w, lane, slot a): 4 (64 a + lane) + w; slot a + 4 is +1024
__device__ __forceinline__ int own_coeff(int w, int lane, int a) { return 4 * (64 * a + lane) + w; }

// round(x * scale) mod 2^64 through two FP64 limbs is done by the callers;
// exact int64 of a rounded double with |x| < 2^52.
__device__ __forceinline__ int64_t round_i64(double x, double scale) {
    const double y = fma(x, scale, 6755399441055744.0);  // 1.5 * 2^52
    return __double_as_longlong(y) - 0x4338000000000000ll;
}

}  // namespace tfhe
