// kernels.hpp -- host launchers of the gfx950 kernels (defined in the .hip files).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace tfhe {
struct cplx;
struct FFTTables;

// gate level (N = 1024, torus-32)
void launch_gen_bsk32(uint64_t seed, int n, int N, int logB, int l, int nshift, const int32_t* s,
                      const int32_t* K, uint32_t* bsk, hipStream_t st);
void launch_gen_ksk32(uint64_t seed, int n, int N, int logB, int l, int nshift, const int32_t* s,
                      const int32_t* K, uint32_t* ksk, hipStream_t st, uint32_t dom = 5, int linear = 0);
void launch_to_fourier1024(const uint32_t* coef, size_t npoly, const FFTTables& T, cplx* out, hipStream_t st);
void launch_gate_blind_rotate(const uint32_t* ct1, const uint32_t* ct2, uint32_t batch, int n, int doubled,
                              uint32_t q1, const cplx* bskf, const FFTTables& T, uint32_t* lweN,
                              hipStream_t st);
// programmable bootstrap: ct over Z_(2N >> logf), tv [2N >> logf] torus-32 test-vector entries
void launch_func_blind_rotate(const uint32_t* ct, uint32_t batch, int n, int logf, const uint32_t* tv,
                              const cplx* bskf, const FFTTables& T, uint32_t* lweN, hipStream_t st);
// out = ((x & mx) - (y & my) + [b] c) & mo on [batch][n+1]; y may be NULL
void launch_lwe_axpb(const uint32_t* x, uint32_t mx, const uint32_t* y, uint32_t my, uint32_t c, uint32_t mo, int n,
                     uint32_t batch, uint32_t* out, hipStream_t st);
// out = round(x 2^-shift) & mo on [batch][n+1]
void launch_lwe_modswitch(const uint32_t* x, int shift, uint32_t mo, int n, uint32_t batch, uint32_t* out,
                          hipStream_t st);
void launch_plain_init(uint32_t* acc, uint32_t q1, uint32_t bt, hipStream_t st);
void launch_plain_step(uint32_t* acc, cplx* digf, uint32_t* t, const cplx* key, uint32_t m,
                       const FFTTables& T, hipStream_t st);
void launch_plain_extract(const uint32_t* acc, uint32_t* lweN, hipStream_t st);
void launch_polymul1024(const int32_t* x, const uint32_t* y, uint32_t batch, const FFTTables& T, uint32_t* out,
                        hipStream_t st);
void launch_external_product1024(uint32_t* acc, cplx* digf, uint32_t* t, const cplx* key, const FFTTables& T,
                                 hipStream_t st);
// key switch N -> n; at: scratch of N * ks_bpad(batch) u32 (rounded, transposed inputs)
uint32_t ks_bpad(uint32_t batch);
// compute units of the current device (cached per device), for launch shapes sized
// to the number of workgroups that run at once
unsigned device_cus();
// slices <= 0: the resident-workgroup model (ks_model_slices); > 0: that many i-slices
int ks_model_slices(uint32_t batch, int N, int n);
int launch_keyswitch32(const uint32_t* lweN, uint32_t batch, int N, int n, int logB, int l, const uint32_t* ksk,
                       uint32_t* at, uint32_t* out, hipStream_t st, int slices = 0);
// circuit-bootstrapping level (N = 2048; PBS torus-64, PFKS / RGSW / LUT torus-32)
struct cplxf;
template <class C>
struct FFT1024TablesT;
using FFT1024Tables = FFT1024TablesT<cplx>;
using FFT1024TablesF = FFT1024TablesT<cplxf>;
void launch_to_fourier2048(const uint64_t* coef, size_t npoly, const FFT1024Tables& T, cplx* out, hipStream_t st);
// torus-32 product x * y mod (X^2048 + 1, 2^32): form 0 = FP32, limbs of 11/11/10 bits;
// 1 = FP32, limbs of 11/6/5/5/5 bits; 2 = one FP64 product.  Returns -1 for another form.
int launch_polymul2048_split32(const int32_t* x, const uint32_t* y, uint32_t batch, const FFT1024TablesF& TF,
                               const FFT1024Tables& T, uint32_t* out, hipStream_t st, int form);
// limbs = 2: exact product; 1: the one-product form of the PBS (rounded once)
void launch_polymul2048(const int32_t* x, const uint64_t* y, uint32_t batch, const FFT1024Tables& T, uint64_t* out,
                        hipStream_t st, int limbs = 2);
// mb = 1: [n][8][2][N] rows of RGSW(s_i); mb = 2: [n/2][3][8][2][N] rows of RGSW(pattern indicators)
void launch_gen_bsk64(uint64_t seed, int n, int nshift, const int32_t* s, const int32_t* K2, const cplx* K2f,
                      const FFT1024Tables& T, uint64_t* bsk, hipStream_t st, int mb = 1, int logB = 9);
// key-switching keys of HomTrace / SchemeSwitch (gadget 2^8 x 5): npoly x 5 RLWE_K2 rows [2][N]
// u64, row q * 5 + d = RLWE(msgpoly[q] 2^(24 + 8 d)) in Philox domain dom
void launch_gen_ks64(uint64_t seed, uint32_t dom, uint32_t npoly, int nshift, const int64_t* msgpoly, const cplx* K2f,
                     const FFT1024Tables& T, uint64_t* key, hipStream_t st);
void launch_gen_pfks(uint64_t seed, int nshift, const int32_t* K2, const cplx* K2f, const FFT1024Tables& T,
                     uint32_t* key, hipStream_t st);
void launch_to_fourier2048_32(const uint32_t* coef, size_t npoly, const FFT1024Tables& T, cplx* out, hipStream_t st);
void launch_to_fourier2048_t64(const uint64_t* coef, size_t npoly, const FFT1024Tables& T, cplx* out, hipStream_t st);
// limbs = 2: exact two-limb BSK2 rows (launch_to_fourier2048); 1: one torus-scaled row (launch_to_fourier2048_t64)
// mb = 2 (limbs = 1, n even): two key bits per step with the pattern keys of launch_gen_bsk64(mb = 2)
// form: 1 = branch-free cross-wave FFT step, 0 = branchy (two key bits per step: 1 = wide key
// batch at one wave per SIMD, 0 = two waves per SIMD, 2 = split over 8 waves), < 0 =
// pbs_model_form's choice
int pbs_model_form(uint32_t batch, int limbs, int mb);
// Test polynomial and output of the multi-value PBS.  lut == NULL (default): the circuit
// bootstrap's two-level test vector, b + 2^30 switched with bw = 1, LWE outputs of
// coefficients 0 and 1 into out.  lut != NULL (the reference's gpu_BootstrapManyLUT):
// modulus switch with bw, acc starts as (0, lut X^-b~) (lut [N] torus-64), and the whole
// RLWE [batch][2][N] goes to rlwe (Gpow_j / 2 added to body coefficient j < 2); out is not
// written.
struct PbsLut {
    const uint64_t* lut = nullptr;
    uint64_t* rlwe = nullptr;
    uint32_t bw = 1;
    uint32_t b_add = 0;    // added to b before the modulus switch
    int offsets = 1;       // 0: no Gpow_j / 2 added (the HomTrace + SchemeSwitch form adds its own)
    int logB = 9;          // PBS gadget base log: 9, or 10 (exact form only) for HomTrace + SchemeSwitch
};
void launch_pbs2(const uint32_t* lwe, uint32_t batch, int n, const cplx* bskf, int limbs, int mb,
                 const FFT1024Tables& T, uint64_t* out, hipStream_t st, int form = -1, const PbsLut& io = PbsLut{});
// PFKS: inputs rounded and transposed to at [N+1][pfks_bpad(B)] u32, then the key switch for
// functions u0 .. u0+nu-1 added into zero-filled outputs (out + (u-u0)*out_u_offset, rows out_stride apart)
uint32_t pfks_bpad(uint32_t B);
void launch_pfks_digits(const uint64_t* lwe64, uint32_t B, uint32_t* at, hipStream_t st);
void launch_zero_rows32(uint32_t* out, uint32_t count, size_t stride, uint32_t width, hipStream_t st);
// pseudo-random words (an integer hash of the index) for timing inputs
void launch_fill_hash32(uint32_t* p, size_t count, uint32_t seed, hipStream_t st);
// Linear-key PFKS (int8 GEMM on the matrix cores): key KT [2][5] x (2N x pfkl_kp()) int8
// (zero the buffer, then launch_gen_pfkl40), digit scratch pfkl_digit_bytes(B) bytes;
// launch_pfkl writes count B torus-32 RLWE outputs (2N words) for functions u0 .. u0+nu-1
// at out + L * out_stride + (u - u0) * out_u_offset (no zero fill needed).
size_t pfkl_key_bytes();
// byte offset of key byte (function u, slice, column, k = 8 i + t) in the fragment-ordered key
size_t pfkl_key_offset(int u, int slice, int col, uint32_t k);
size_t pfkl_digit_bytes(uint32_t rows);
uint32_t pfkl_kp();
void launch_gen_pfkl40(uint64_t seed, const int32_t* K2, int8_t* KT, hipStream_t st);
void launch_pfkl(const uint64_t* lwe64, uint32_t B, int8_t* D, const int8_t* KT, int u0, int nu, uint32_t* out,
                 size_t out_stride, size_t out_u_offset, hipStream_t st);
// Linear-key N -> n key switch (int8 GEMM on the matrix cores): key KT of kskl_key_bytes()
// int8 (zero the buffer, then launch_kskl_pack from the key words [N][8][n+1] u32);
// launch_ks_linear switches count torus-32 LWE_K2 [count][N+1] into out [count][n+1] in
// chunks of at most `chunk` inputs, with digit scratch D of pfkl_digit_bytes(chunk) bytes.
// Both return -1 for n + 1 > 640.
size_t kskl_key_bytes();
int launch_kskl_pack(const uint32_t* kl, int n, int8_t* KT, hipStream_t st);
int launch_ks_linear(const uint32_t* lwe32, uint32_t count, int n, const int8_t* KT, int8_t* D, uint32_t chunk,
                     uint32_t* out, hipStream_t st);
// LWEs as a u64 mask array (rows a_stride apart) and a u64 body array -> [batch][n+1] u32
void launch_pack_lwe64(const uint64_t* a, const uint64_t* b, uint32_t batch, int n, uint32_t a_stride, uint32_t* out,
                       hipStream_t st);
// slices <= 0: the resident-workgroup cost model (pfks_model_slices); > 0: that many i-slices
int pfks_model_slices(uint32_t B, int nu);
void launch_pfks32(const uint32_t* at, uint32_t B, const uint32_t* key, size_t key_u_stride, int u0, int nu,
                   uint32_t* out, size_t out_stride, size_t out_u_offset, hipStream_t st, int slices = 0);
// parts = 1: LUT k reads RGSW group k and table k; parts > 1: `parts` table polynomials
// (table_stride apart) share each RGSW group (LUT k: group k / parts, table k % parts, of
// table set (group mod tperiod) when tperiod > 0, sets parts polynomials apart)
void launch_lut_vertical(const cplx* rgsw_f, uint32_t nlut, const uint32_t* table, size_t table_stride, int stride,
                         int out_bits, const FFT1024Tables& T, uint32_t* out, hipStream_t st, int parts = 1,
                         int tperiod = 0);
// HomTrace + SchemeSwitch of the PBS accumulators acc [batch][2][N] (LUT form, no offsets)
// into torus-32 RGSW rows [batch][4][2][N] (row 2j SS, 2j + 1 HT); two-limb Fourier keys
void launch_htss(const uint64_t* acc, uint32_t batch, const cplx* htkey, const cplx* sskey, const FFT1024Tables& T,
                 uint32_t* rgsw, hipStream_t st);
// T-table round: XOR of the four T-table words per column at dimension N, lutw
// [blocks 16][32][N+1] -> out [blocks][128][N+1]; dec = 1: the InvShiftRows word map
void launch_aes_ttable_xor(const uint32_t* lutw, uint32_t blocks, int dec, uint32_t* out, hipStream_t st);
void launch_aes_init(const uint32_t* rk0, const uint8_t* ctr, uint32_t blocks, int n, uint32_t* state, hipStream_t st);
void launch_aes_linear(const uint32_t* in, const uint32_t* rk, const int16_t* src, int nsrc, uint32_t blocks, int n,
                       uint32_t* out, hipStream_t st);
void launch_xor_public(uint32_t* ct, const uint8_t* pub, uint32_t nbits, int n, hipStream_t st);
}  // namespace tfhe
