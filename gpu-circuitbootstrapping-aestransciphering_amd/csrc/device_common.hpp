// device_common.hpp -- device-side building blocks shared by every gfx950 kernel:
// the Philox4x32-10 key-material sampler, signed gadget decomposition, complex
// helpers and the wavefront-resident negacyclic FFT.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tfhe {

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11).  Counter = (j, id, sub, domain), key = seed.
// ---------------------------------------------------------------------------
struct u32x4 { uint32_t x, y, z, w; };

__host__ __device__ inline u32x4 philox4x32_10(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2,
                                               uint32_t c3) {
    uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    return {c0, c1, c2, c3};
}

enum : uint32_t {
    DOM_LWE_SK = 1, DOM_RLWE1_SK = 2, DOM_RLWE2_SK = 3, DOM_BSK1 = 4, DOM_KSK1 = 5,
    DOM_BSK2 = 6, DOM_PFKS = 7, DOM_KSK2 = 8, DOM_BSK2MB = 9, DOM_PFKL = 10, DOM_HTKEY = 11, DOM_SSKEY = 12,
    DOM_KSKL = 13,
    DOM_ENC = 16
};

__host__ __device__ inline uint32_t rand32(uint64_t seed, uint32_t dom, uint32_t id, uint32_t j) {
    return philox4x32_10(seed, j, id, 0, dom).x;
}

__host__ __device__ inline uint64_t rand64(uint64_t seed, uint32_t dom, uint32_t id, uint32_t j) {
    u32x4 w = philox4x32_10(seed, j, id, 0, dom);
    return (uint64_t)w.x | ((uint64_t)w.y << 32);
}

// Irwin-Hall(12) noise with unit variance scaled by 2^32, rounded to 2^-nshift of that.
__host__ __device__ inline int64_t noise(uint64_t seed, uint32_t dom, uint32_t id, uint32_t j, int nshift) {
    uint64_t s = 0;
    for (uint32_t sub = 1; sub <= 3; ++sub) {
        u32x4 w = philox4x32_10(seed, j, id, sub, dom);
        s += (uint64_t)w.x + w.y + w.z + w.w;
    }
    int64_t z = (int64_t)s - (int64_t)6 * ((int64_t)1 << 32);
    if (nshift > 0) return (z + ((int64_t)1 << (nshift - 1))) >> nshift;
    return z * ((int64_t)1 << (-nshift));
}

__host__ __device__ inline int32_t secret_bit(uint64_t seed, uint32_t dom, uint32_t i) {
    return (int32_t)(philox4x32_10(seed, i, 0, 0, dom).x & 1u);
}

// ---------------------------------------------------------------------------
// Signed gadget decomposition on a 32-bit torus: centre, drop the low
// 32 - l*logB bits with rounding, then l signed digits in [-B/2, B/2), least
// significant first.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int32_t sext32(int32_t v, int bits) {
    return (v << (32 - bits)) >> (32 - bits);
}

template <int LOGB, int L>
__device__ __forceinline__ void decompose32(uint32_t t, int32_t (&dig)[L]) {
    constexpr int IGN = 32 - L * LOGB;
    int32_t d = (int32_t)t;
    if constexpr (IGN > 0) {
        int32_t r = sext32(d, IGN);
        d = (d >> IGN) + (r < 0 ? 1 : 0);
    }
#pragma unroll
    for (int j = 0; j < L; ++j) {
        int32_t r = sext32(d, LOGB);
        dig[j] = r;
        d = (d >> LOGB) + (r < 0 ? 1 : 0);
    }
}

// ---------------------------------------------------------------------------
// Complex FP64 helpers
// ---------------------------------------------------------------------------
struct cplx { double re, im; };

__device__ __forceinline__ cplx cadd(cplx a, cplx b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cplx csub(cplx a, cplx b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ cplx cmul(cplx a, cplx b) {
    return {fma(a.re, b.re, -a.im * b.im), fma(a.re, b.im, a.im * b.re)};
}
__device__ __forceinline__ cplx cmulc(cplx a, cplx b) {  // a * conj(b)
    return {fma(a.re, b.re, a.im * b.im), fma(a.im, b.re, -a.re * b.im)};
}
__device__ __forceinline__ void cmac(cplx& acc, cplx a, cplx b) {
    acc.re = fma(a.re, b.re, acc.re);
    acc.re = fma(-a.im, b.im, acc.re);
    acc.im = fma(a.re, b.im, acc.im);
    acc.im = fma(a.im, b.re, acc.im);
}

// multiply by -i (forward) or +i (inverse)
template <bool INV>
__device__ __forceinline__ cplx mul_mi(cplx a) {
    if constexpr (INV) return {-a.im, a.re};
    else return {a.im, -a.re};
}

// Complex FP32 (the split-precision product): the same operations, single rounding
// per fused multiply-add.  Scalar code only: the compiler is kept from packed FP32
// forms by the build flags.
struct cplxf { float re, im; };

__device__ __forceinline__ cplxf cadd(cplxf a, cplxf b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cplxf csub(cplxf a, cplxf b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ cplxf cmul(cplxf a, cplxf b) {
    return {fmaf(a.re, b.re, -a.im * b.im), fmaf(a.re, b.im, a.im * b.re)};
}
__device__ __forceinline__ cplxf cmulc(cplxf a, cplxf b) {
    return {fmaf(a.re, b.re, a.im * b.im), fmaf(a.im, b.re, -a.re * b.im)};
}
template <bool INV>
__device__ __forceinline__ cplxf mul_mi(cplxf a) {
    if constexpr (INV) return {-a.im, a.re};
    else return {a.im, -a.re};
}
__device__ __forceinline__ double sfma(double a, double b, double c) { return fma(a, b, c); }
__device__ __forceinline__ float sfma(float a, float b, float c) { return fmaf(a, b, c); }

// 8-point DFT in registers, natural order in and out; forward uses exp(-2 pi i/8).
template <bool INV>
__device__ __forceinline__ void dft8(cplx (&x)[8]) {
    constexpr double R = 0.70710678118654752440084436210485;
    // DFT4 on even and odd halves
    cplx e0 = cadd(x[0], x[4]), e1 = csub(x[0], x[4]);
    cplx e2 = cadd(x[2], x[6]), e3 = mul_mi<INV>(csub(x[2], x[6]));
    cplx E0 = cadd(e0, e2), E2 = csub(e0, e2), E1 = cadd(e1, e3), E3 = csub(e1, e3);
    cplx o0 = cadd(x[1], x[5]), o1 = csub(x[1], x[5]);
    cplx o2 = cadd(x[3], x[7]), o3 = mul_mi<INV>(csub(x[3], x[7]));
    cplx O0 = cadd(o0, o2), O2 = csub(o0, o2), O1 = cadd(o1, o3), O3 = csub(o1, o3);
    // twiddles W8^k on the odd half
    cplx T1, T3;
    if constexpr (INV) {
        T1 = {R * (O1.re - O1.im), R * (O1.re + O1.im)};
        T3 = {-R * (O3.re + O3.im), R * (O3.re - O3.im)};
    } else {
        T1 = {R * (O1.re + O1.im), R * (O1.im - O1.re)};
        T3 = {R * (O3.im - O3.re), -R * (O3.re + O3.im)};
    }
    cplx T2 = mul_mi<INV>(O2);
    x[0] = cadd(E0, O0); x[4] = csub(E0, O0);
    x[1] = cadd(E1, T1); x[5] = csub(E1, T1);
    x[2] = cadd(E2, T2); x[6] = csub(E2, T2);
    x[3] = cadd(E3, T3); x[7] = csub(E3, T3);
}

// ---------------------------------------------------------------------------
// Wavefront-resident negacyclic FFT for N = 1024 (M = 512 complex points).
//
// Data distribution (lane l of a 64-wide wavefront, register slot s):
//   coefficient domain : lane b, slot a (a < 8) holds the folded pair
//                        (c[64a + b], c[64a + b + 512]) as (re, im).
//   frequency domain   : lane 8*k0 + e, slot f holds X[k0 + 8e + 64f].
// Both LDS transposes use rows of 72 entries so that neighbouring 8-lane groups
// land on different bank halves.
// The negacyclic twist w^(64a+b), w = exp(i pi / N), splits into the
// compile-time factor w^(64a) (applied per slot before the first radix-8) and
// the per-lane factor w^b, which commutes with the in-register radix-8 and is
// folded into the pass-1 twiddle table.  Forward = radix-8 over slots, twiddle,
// LDS transpose, radix-8, twiddle, LDS transpose, radix-8 (decimation in
// frequency, digit-reversed output).  The inverse runs the exact reverse and
// returns 512 x the folded pairs.  The frequency order never leaves the kernel
// family (Fourier keys are stored in the same order).
//
// Twiddle tables (global memory, [slot][lane] so every load is coalesced):
//   p1t[k0][b] = w^b * exp(-2 pi i b k0 / 512)
//   p2[e][l]   = exp(-2 pi i (l%8) e / 64)
// LDS: one exchange buffer of 8*72 cplx (9216 B) per wavefront.
// ---------------------------------------------------------------------------
constexpr int FFT512_LDS_CPLX = 8 * 72;

struct FFTTables {
    const cplx* p1t;  // [8][64]
    const cplx* p2;   // [8][64]
    const cplx* rot;  // [2N] w^j, w = exp(i pi / N): X[k] of the transform is p(w^(1 - 4k))
};

// w^(64a) = exp(i pi a / 16), a = 0..7
__device__ __forceinline__ cplx twist_const(int a) {
    constexpr double C[8][2] = {
        {1.0, 0.0},
        {0.98078528040323043, 0.19509032201612825},
        {0.92387953251128674, 0.38268343236508978},
        {0.83146961230254524, 0.55557023301960218},
        {0.70710678118654757, 0.70710678118654746},
        {0.55557023301960229, 0.83146961230254524},
        {0.38268343236508984, 0.92387953251128674},
        {0.19509032201612833, 0.98078528040323043}};
    return {C[a][0], C[a][1]};
}

// Orders LDS traffic between the lanes of one wavefront: the LDS-only fences
// stop the compiler from moving LDS accesses across the point (global loads
// stay free to be hoisted), the wave barrier keeps the lanes converged.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

// forward: x = folded coefficient pairs (untwisted), out in frequency distribution
__device__ __forceinline__ void fft512_fwd(cplx (&x)[8], cplx* lds, const FFTTables& T, int lane) {
#pragma unroll
    for (int a = 1; a < 8; ++a) x[a] = cmul(x[a], twist_const(a));
    dft8<false>(x);
#pragma unroll
    for (int k = 0; k < 8; ++k) x[k] = cmul(x[k], T.p1t[k * 64 + lane]);
    // exchange 1: lane b writes y[b][k0] at k0*72 + b (72: bank-staggered rows)
#pragma unroll
    for (int k = 0; k < 8; ++k) lds[k * 72 + lane] = x[k];
    wave_sync();
    const int k0 = lane >> 3, d = lane & 7;
#pragma unroll
    for (int c = 0; c < 8; ++c) x[c] = lds[k0 * 72 + 8 * c + d];
    wave_sync();
    dft8<false>(x);
#pragma unroll
    for (int e = 1; e < 8; ++e) x[e] = cmul(x[e], T.p2[e * 64 + lane]);
    // exchange 2: lane (k0,d) writes u[e] at k0*72 + 9e + d
#pragma unroll
    for (int e = 0; e < 8; ++e) lds[k0 * 72 + 9 * e + d] = x[e];
    wave_sync();
    const int e2 = lane & 7;
#pragma unroll
    for (int dd = 0; dd < 8; ++dd) x[dd] = lds[k0 * 72 + 9 * e2 + dd];
    wave_sync();
    dft8<false>(x);
}

// inverse: x in frequency distribution, out = 512 x folded coefficient pairs
__device__ __forceinline__ void fft512_inv(cplx (&x)[8], cplx* lds, const FFTTables& T, int lane) {
    dft8<true>(x);
    const int k0 = lane >> 3, e = lane & 7;
#pragma unroll
    for (int dd = 0; dd < 8; ++dd) lds[k0 * 72 + 9 * e + dd] = x[dd];
    wave_sync();
    const int d = lane & 7;
#pragma unroll
    for (int ee = 0; ee < 8; ++ee) x[ee] = lds[k0 * 72 + 9 * ee + d];
    wave_sync();
#pragma unroll
    for (int ee = 1; ee < 8; ++ee) x[ee] = cmulc(x[ee], T.p2[ee * 64 + lane]);
    dft8<true>(x);
#pragma unroll
    for (int c = 0; c < 8; ++c) lds[k0 * 72 + 8 * c + d] = x[c];
    wave_sync();
#pragma unroll
    for (int k = 0; k < 8; ++k) x[k] = lds[k * 72 + lane];
    wave_sync();
#pragma unroll
    for (int k = 0; k < 8; ++k) x[k] = cmulc(x[k], T.p1t[k * 64 + lane]);
    dft8<true>(x);
#pragma unroll
    for (int a = 1; a < 8; ++a) x[a] = cmulc(x[a], twist_const(a));
}

// round(x * scale) mod 2^32 for |x * scale| < 2^51: adding 1.5 * 2^52 leaves the
// rounded integer (ties to even, as rint) in the low mantissa bits.
__device__ __forceinline__ uint32_t round_mod32(double x, double scale) {
    const double y = fma(x, scale, 6755399441055744.0);
    return (uint32_t)__double_as_longlong(y);
}

// Signed digits by the offset method: with off = 2^(IGN-1) + sum_j (B/2) 2^(IGN + j logB),
// digit_j = ((t + off) >> (IGN + j logB) & (B-1)) - B/2.  This is the same balanced
// representation as decompose32 (round half up, digits in [-B/2, B/2)).
template <int LOGB, int L>
__host__ __device__ constexpr uint32_t digit_offset32() {
    constexpr int IGN = 32 - L * LOGB;
    uint32_t off = IGN > 0 ? (1u << (IGN - 1)) : 0u;
    for (int j = 0; j < L; ++j) off += (1u << (LOGB - 1)) << (IGN + j * LOGB);
    return off;
}

template <int LOGB, int L>
__device__ __forceinline__ int32_t digit32(uint32_t t_plus_off, int j) {
    constexpr int IGN = 32 - L * LOGB;
    return (int32_t)((t_plus_off >> (IGN + j * LOGB)) & ((1u << LOGB) - 1)) - (1 << (LOGB - 1));
}

}  // namespace tfhe
