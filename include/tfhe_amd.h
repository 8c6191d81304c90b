/*
 * tfhe_amd.h -- C ABI of the MI355X-native TFHE bootstrapping library
 * (libtfhe_amd.so, gfx950).
 *
 * Plain pointers and sizes only.  Arguments named d_* are device pointers; the
 * stream argument is a hipStream_t passed as void* (NULL = the default stream).
 * Every entry point returns TFHE_OK (0) or an error code; tfhe_last_error()
 * returns a thread-local message for the last failure.  No entry point exits
 * the process (the reference's CUDA_CHECK_AND_EXIT, include/cuda_wrapper.cuh:45-54,
 * has no counterpart here).
 *
 * Torus encodings: level-0 LWE ciphertexts are uint32 vectors (a[0..n-1], b)
 * over Z/2^32 with phase b - <a, s>.  Gate ciphertexts encode a bit m as
 * m * 2^30 (OpenFHE plaintext modulus p = 4).
 *
 * Threading: one host thread per context; contexts are not thread-safe
 * (reference bitwise/include/cirbts/cirbts.cuh:303-385).  All device work is
 * asynchronous on the given stream; the caller synchronises.
 */
#ifndef TFHE_AMD_H
#define TFHE_AMD_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    TFHE_OK = 0,
    TFHE_E_PARAM = 1,    /* invalid parameter / unsupported gate (reference OPENFHE_THROW config_error) */
    TFHE_E_SHAPE = 2,    /* shape mismatch */
    TFHE_E_HIP = 3,      /* HIP runtime failure (reference PHANTOM_CHECK_CUDA) */
    TFHE_E_CAPACITY = 4, /* batch above the context's max_batch (reference cirbts-batch.cu:97-100) */
    TFHE_E_STATE = 5     /* keys missing */
};

/* OpenFHE BINGATE numbering (the enum the reference's gate API takes). */
enum {
    TFHE_OR = 0, TFHE_AND = 1, TFHE_NOR = 2, TFHE_NAND = 3, TFHE_XOR_FAST = 4, TFHE_XNOR_FAST = 5,
    TFHE_XOR = 11, TFHE_XNOR = 12
};

const char* tfhe_last_error(void);

/* ------------------------------------------------------------------------- */
/* Gate bootstrapping context.  Replaces phantom::bitwise::GPUBinFHEContext   */
/* (reference bitwise/include/binfhecontext.cuh:8-119).                       */
/* ------------------------------------------------------------------------- */
typedef struct tfhe_gate_params {
    int n;           /* LWE dimension (636) */
    int N;           /* ring dimension (1024) */
    int bk_logB;     /* blind-rotation gadget base log (7) */
    int bk_l;        /* blind-rotation gadget digits (3) */
    int ks_logB;     /* key-switching base log (2) */
    int ks_l;        /* key-switching digits (8) */
    int lwe_nshift;  /* LWE noise: Irwin-Hall sigma = 2^-lwe_nshift of the torus (15) */
    int rlwe_nshift; /* RLWE noise of the bootstrapping key (25) */
} tfhe_gate_params;

typedef struct tfhe_gate_ctx tfhe_gate_ctx;

/* Constructor GPUBinFHEContext(cc) (binfhecontext.cuh:18-29): allocates device
 * keys and scratch for up to max_batch ciphertexts per call. */
int tfhe_gate_ctx_create(const tfhe_gate_params* params, uint32_t max_batch, tfhe_gate_ctx** out);
int tfhe_gate_ctx_destroy(tfhe_gate_ctx* ctx);

/* GPUBTKeyGen(sk, SYM_ENCRYPT) (binfhecontext.cuh:31-37, binfhe-base-scheme.cu:18-97):
 * lwe_sk is a host array of n binary coefficients; the RLWE secret, the
 * bootstrapping key and the key-switching key are derived from seed on the GPU. */
int tfhe_gate_bt_keygen(tfhe_gate_ctx* ctx, const int32_t* lwe_sk, uint64_t seed);

/* Upload externally generated coefficient-domain keys (host pointers):
 * bsk [n][2*bk_l][2][N] uint32, ksk [N][ks_l][2^ks_logB][n+1] uint32.
 * The BSK is converted to the Fourier domain on the GPU. */
int tfhe_gate_set_keys(tfhe_gate_ctx* ctx, const uint32_t* bsk, const uint32_t* ksk);

/* Copy the RLWE secret (N int32) used by the context to the host. */
int tfhe_gate_get_rlwe_sk(const tfhe_gate_ctx* ctx, int32_t* out);

/* BatchGPUEvalBinGate(gate, v1, v2) (binfhecontext.cuh:53-57,
 * binfhe-base-scheme-batch.cu:81-210): batch independent gate bootstraps in one
 * persistent blind-rotation launch plus one batched key switch.
 * d_ct1, d_ct2, d_out: [batch][n+1] uint32.  d_ct1 == d_ct2 is rejected as the
 * reference rejects ct1 == ct2 (binfhe-base-scheme.cu:158-159). */
int tfhe_gate_eval_batch(tfhe_gate_ctx* ctx, int gate, const uint32_t* d_ct1, const uint32_t* d_ct2,
                         uint32_t batch, uint32_t* d_out, void* stream);

/* tfhe_gate_eval_batch that allocates nothing.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments and the same output words as
 * tfhe_gate_eval_batch; its scratch (the blind rotation's [batch][N+1] outputs and the key
 * switch's [N][batch] digits, 8.0 MiB) is a fixed device array of the library, one per GPU,
 * held until the process ends, instead of the context's buffers.  batch is capped by
 * TFHE_NOALLOC_MAX_GATES, not by the context's max_batch (TFHE_E_CAPACITY beyond it).
 * Every call on one GPU shares that GPU's array: calls that overlap on one GPU (two streams
 * or two host threads) race and must be serialised by the caller. */
#define TFHE_NOALLOC_MAX_GATES 1024u
int tfhe_gate_eval_batch_noalloc(tfhe_gate_ctx* ctx, int gate, const uint32_t* d_ct1, const uint32_t* d_ct2,
                                 uint32_t batch, uint32_t* d_out, void* stream);

/* GPUEvalBinGate(gate, ct1, ct2) (binfhecontext.cuh:48-51, binfhe-base-scheme.cu:151-252)
 * in the reference's plain per-step form: one ciphertext, three kernel launches
 * per blind-rotation step.  Synchronises the stream (the rotation indices are
 * read on the host, as the reference reads ct->GetA() on the host). */
int tfhe_gate_eval_plain(tfhe_gate_ctx* ctx, int gate, const uint32_t* d_ct1, const uint32_t* d_ct2,
                         uint32_t* d_out, void* stream);

/* The two halves of tfhe_gate_eval_batch, exposed for parity tests:
 * blind rotation + sample extraction into d_lweN [batch][N+1] (the + Q/8 + 1
 * offset applied), then the N -> n key switch into d_out [batch][n+1]. */
int tfhe_gate_blind_rotate_batch(tfhe_gate_ctx* ctx, int gate, const uint32_t* d_ct1,
                                 const uint32_t* d_ct2, uint32_t batch, uint32_t* d_lweN, void* stream);
int tfhe_gate_keyswitch_batch(tfhe_gate_ctx* ctx, const uint32_t* d_lweN, uint32_t batch, uint32_t* d_out,
                              void* stream);

/* ------------------------------------------------------------------------- */
/* Programmable bootstrap and large-precision functions on the gate context  */
/* (Liu-Micciancio-Polyakov, eprint 2021/1337).  LWE ciphertexts are         */
/* [batch][n+1] uint32 over a power-of-two modulus (2 .. 2^32): residues     */
/* under the same LWE key, a message m over Z_mod encrypted as the 2^32-torus */
/* phase m 2^32 / mod rounded to Z_mod.  Every flow is a fixed sequence of   */
/* launches on `stream` (nothing is read back); the flows share the context's */
/* scratch, so calls on one context must not overlap (one stream at a time). */
/* The table pointers are host memory, read before the call returns.         */
/* ------------------------------------------------------------------------- */
#define TFHE_FUNC_Q 1024u   /* q_small: the modulus EvalFloor/EvalSign/EvalDecomp work at */
#define TFHE_FUNC_BETA 64u  /* beta: the offset against bootstrapping noise; q_small/(2 beta) = 8 levels per floor */
enum { TFHE_FUNC_NEGACYCLIC = 0, TFHE_FUNC_PERIODIC = 1, TFHE_FUNC_ARBITRARY = 2 };

/* The class of a table over Z_q (q a power of two): negacyclic when lut[x + q/2] = -lut[x]
 * mod q, periodic when lut[x + q/2] = lut[x], arbitrary otherwise; -1 for a bad argument.
 * The test EvalFunc makes (OpenFHE BinFHEScheme::checkInputFunction). */
int tfhe_func_property(const uint32_t* lut, uint64_t q);

/* BatchGPUBootstrapFunc(params, EK, v_ct, f, fmod) (binfhe-base-scheme-batch.cu:267-290,
 * single form binfhe-base-scheme.cu:259-378): d_ct over Z_ctmod (ctmod a power of two <= 2N),
 * lut [ctmod] over Z_fmod (fmod a power of two <= 2^32).  Test vector m[j 2N/ctmod] =
 * lut[(b - j) mod ctmod] 2^32/fmod for j < ctmod/2; the output over Z_fmod encrypts
 * lut[phase] when <a,s> mod ctmod < ctmod/2 and -lut[phase + ctmod/2] otherwise. */
int tfhe_gate_bootstrap_func_batch(tfhe_gate_ctx* ctx, const uint32_t* d_ct, uint32_t batch, uint64_t ctmod,
                                   const uint32_t* lut, uint64_t fmod, uint32_t* d_out, void* stream);

/* BatchGPUEvalFunc(v_ct, LUT) (binfhecontext.cuh:70-72, binfhe-base-scheme.cu:380-481): d_ct
 * and d_out over Z_q (q a power of two <= 2N; <= N when the table is arbitrary), lut [q] over
 * Z_q; the output encrypts lut at the input phase plus beta.  Tables built as
 * f(floor(x p / q)) q / p (OpenFHE GenerateLUTviaFunction) absorb that offset and the noise for
 * inputs m q / p + e.  One bootstrap for a negacyclic table, two otherwise. */
int tfhe_gate_eval_func_batch(tfhe_gate_ctx* ctx, const uint32_t* d_ct, uint32_t batch, uint64_t q,
                              const uint32_t* lut, uint32_t* d_out, void* stream);

/* BatchGPUEvalFloor(v_ct, roundbits) (binfhecontext.cuh:86-88, binfhe-base-scheme.cu:484-531):
 * d_ct and d_out over Z_mod; clears the low part of the message modulo the inner modulus
 * q = TFHE_FUNC_Q (roundbits 0) or TFHE_FUNC_BETA 2^(roundbits+1) (q <= min(2N, mod)). */
int tfhe_gate_eval_floor_batch(tfhe_gate_ctx* ctx, const uint32_t* d_ct, uint32_t batch, uint64_t mod,
                               uint32_t roundbits, uint32_t* d_out, void* stream);

/* BatchGPUEvalSign(v_ct, schemeSwitch) (binfhecontext.cuh:102-104, binfhe-base-scheme.cu:533-608):
 * d_ct over Z_mod (TFHE_FUNC_Q < mod <= 2^32), d_out over Z_TFHE_FUNC_Q encrypting 0 for a
 * message below mod/2 and q_small/2 above (scheme_switch: q_small/4 and -q_small/4 instead). */
int tfhe_gate_eval_sign_batch(tfhe_gate_ctx* ctx, const uint32_t* d_ct, uint32_t batch, uint64_t mod,
                              int scheme_switch, uint32_t* d_out, void* stream);

/* BatchGPUEvalDecomp(v_ct) (binfhecontext.cuh:116-118, binfhe-base-scheme.cu:633-698): digits
 * of d_ct (over Z_mod, TFHE_FUNC_Q < mod <= 2^32) into d_out [digits][batch][n+1], digit k over
 * Z_digit_mods[k] (TFHE_FUNC_Q for all but the last); *n_digits and digit_mods (host, optional,
 * max_digits entries) are written before any launch.  TFHE_E_CAPACITY when max_digits is too
 * small (mod 2^32 takes 9). */
int tfhe_gate_eval_decomp_batch(tfhe_gate_ctx* ctx, const uint32_t* d_ct, uint32_t batch, uint64_t mod,
                                uint32_t max_digits, uint32_t* d_out, uint32_t* n_digits, uint64_t* digit_mods,
                                void* stream);

/* Exact negacyclic products through the wavefront FP64 FFT, N = 1024:
 * d_out[k] = d_x[k] * d_y[k] mod (X^N + 1, 2^32) for small-integer d_x (|x| < 2^10)
 * and torus-32 d_y, all [batch][1024].  The FFT of the reference
 * (bitwise/src/fft.cu:3797-3875, cuFFTDxWrapper) in one self-contained call. */
int tfhe_polymul1024_batch(const int32_t* d_x, const uint32_t* d_y, uint32_t batch, uint32_t* d_out, void* stream);

/* One external product t = dec(acc) (*) BSK[key_index] in the reference's
 * per-step form (kernel_SignedDigitDecompose + FFT + kernel_EvalAccCoreCGGI_fft +
 * inverse, bitwise/src/rgsw-acc.cu:614-668): d_acc, d_out [2][N] uint32. */
int tfhe_gate_external_product(tfhe_gate_ctx* ctx, uint32_t key_index, uint32_t* d_acc, uint32_t* d_out,
                               void* stream);

/* Stateless key generators (the GPU side of keygen parity):
 * d_bsk [n][2*bk_l][2][N], d_ksk [N][ks_l][2^ks_logB][n+1]. */
int tfhe_gen_bsk32(const tfhe_gate_params* p, uint64_t seed, const int32_t* d_lwe_sk, const int32_t* d_rlwe_sk,
                   uint32_t* d_bsk, void* stream);
int tfhe_gen_ksk32(const tfhe_gate_params* p, uint64_t seed, const int32_t* d_lwe_sk, const int32_t* d_rlwe_sk,
                   uint32_t* d_ksk, void* stream);

/* ------------------------------------------------------------------------- */
/* Circuit-bootstrapping context (level 2: N = 2048, torus-64).  Replaces     */
/* GPUCirBTSContext (reference bitwise/include/cirbts/cirbts.cuh:31-500).     */
/* LWE inputs are level-0 ciphertexts (uint32, n + 1) encoding bit * 2^31.    */
/* Outputs are RGSW_K2(bit) on the 32-bit torus in the coefficient domain,   */
/* uint32 [2*cb_l rows][2][N]: row 2j + u is the private functional key       */
/* switch f_u of the torus-64 LWE_K2(bit * 2^(48 + 8j)), i.e. of bit * 2^(16+8j)*/
/* on the 32-bit torus, f_0 = -K2 * x, f_1 = x.                              */
/* ------------------------------------------------------------------------- */
typedef struct tfhe_cbs_params {
    int n;            /* level-0 LWE dimension (636) */
    int N;            /* ring dimension (2048) */
    int bk_logB, bk_l;  /* PBS gadget 2^9 x 4 */
    int cb_logB, cb_l;  /* circuit-bootstrapping gadget 2^8 x 2 */
    int pf_logB, pf_l;  /* private functional key switching 2^4 x 8 */
    int ks_logB, ks_l;  /* N -> n key switching 2^2 x 8 */
    int lwe_nshift;     /* level-0 noise 2^-15 */
    int rlwe_nshift;    /* BSK2 noise: sigma = 2^-(32 + rlwe_nshift) of the 64-bit torus (19) */
    int pf_nshift;      /* PFKS key noise: sigma = 2^-pf_nshift of the 32-bit torus (30) */
    /* Precision form of the PBS's digit-polynomial x torus-64 key products (the
     * reference picks limb splits for its FFT backend, cirbts-init.cu:1096-1120):
     * 0 = two signed 32-bit limbs, each product rounded exactly, so the PBS output
     *     is bit-identical to exact integer arithmetic (default);
     * 1 = one FP64 product of the key scaled to the torus: half the inverse
     *     transforms, MACs and key bytes.  The rounding adds about 2^-33 of the
     *     torus per bootstrap to a PBS noise of about 2^-28, so outputs decrypt
     *     as the exact form's do, but they are not the same words (a rounding
     *     that moves a gadget digit brings in another key row). */
    int pbs_approx;
    /* Key bits per blind-rotation step (the reference's multi-bit PBS,
     * main.cu:130, kernel.cu:4585-4771): 0 or 1 = one; 2 = two, with three
     * pattern keys per key-bit pair (RGSW of the indicators of (1,0), (0,1),
     * (1,1)): one decomposition and two inverse transforms per pair, 1.5x the
     * MACs and key bytes.  2 needs pbs_approx = 1 and an even n.  3 = both key
     * forms (417 MB), one or two bits per step chosen per launch: two unless
     * tfhe_cbs_autotune measured one bit faster for that launch size (same
     * conditions as 2). */
    int pbs_multibit;
    /* Private functional key switch form (configs: PFKS base 2^4, 8 digits):
     * 0 = lookup key, every digit multiple v = 1..15 its own torus-32 RLWE row
     *     (2 x 4.03 GB), one LDS read and subtraction per (input, row, column);
     * 1 = linear key, one row per (i, t) on Z/2^40 (sigma 2^-33, 2 x 336 MB as
     *     signed bytes), signed digits, the sum an int8 GEMM on the matrix cores,
     *     rounded to the 32-bit torus.  Output noise 2^-23.8 of the torus against
     *     2^-23.0 for the lookup key; the two forms give different words. */
    int pfks_linear;
    /* LWE -> RGSW conversion after the PBS:
     * 0 = private functional key switching in the pfks_linear form (default);
     * 1 = HomTrace + SchemeSwitch, the reference's own conversion (cirbts-batch.cu:
     *     1612-1790): per CB level the PBS accumulator's level coefficient, prescaled by
     *     1/N, is traced over log2 N = 11 automorphism rounds (each a key switch with
     *     gadget 2^8 x 5) into RLWE(bit g_j) (row 2j + 1), and scheme-switched with the
     *     key RLWE(K2^2 g_d) into RLWE(-K2 bit g_j) (row 2j), exact two-limb FFT products
     *     throughout.  The trace multiplies the PBS noise by N at the constant coefficient,
     *     so this form needs a quieter PBS: exact (pbs_approx = 0, pbs_multibit <= 1) with
     *     a 2^10 x 4 gadget (bk_logB = 10) and, for the stated noise, rlwe_nshift = 22
     *     (BSK2, HT and SS keys at sigma 2^-54).  No PFKS key is held (HT + SS keys:
     *     3.9 MB in Fourier form).  Rows are the same RGSW(bit) the PFKS forms give, other words. */
    int lwe_to_rgsw;
    /* N -> n key switch form (base 2^2 x 8 on the rounded top 16 bits, as the
     * reference's kernel_KeySwitchU16, aes_gpu_ks.cuh:341-370):
     * 0 = lookup key, one LWE row per (i, digit j, digit value v) ([N][8][4][n+1] u32),
     *     one LDS-staged row selected and subtracted per digit (default);
     * 1 = linear key, one row per (i, j) ([N][8][n+1] u32, LWE_s(K2_i 2^(16 + 2 j))),
     *     signed digits in [-2, 2), the digit sum an int8 GEMM on the matrix cores
     *     (each key word as 4 signed bytes).  Same phase, other words; a quarter of
     *     the key bytes; noise variance 1.5 x 8N key noises against 0.75 x 8N. */
    int ks_linear;
} tfhe_cbs_params;

typedef struct tfhe_cbs_ctx tfhe_cbs_ctx;

/* GPUCirBTSContext(cc, maxBatchSize) (cirbts.cuh:420-426): max_blocks AES blocks,
 * i.e. 128 * max_blocks circuit bootstraps per call; at most 4096 blocks
 * (TFHE_E_PARAM beyond). */
int tfhe_cbs_ctx_create(const tfhe_cbs_params* params, uint32_t max_blocks, tfhe_cbs_ctx** out);
int tfhe_cbs_ctx_destroy(tfhe_cbs_ctx* ctx);

/* Key generation on the GPU from seed (replaces the fork's CirBTKeyGen,
 * main.cu:232-235, and gpu_init_cirbsk, cirbts-init.cu:200-865): RLWE secret
 * K2, BSK2 (torus-64, Fourier, 2 limbs or 1 with pbs_approx), the torus-32 PFKS key (2 x 4.03 GB)
 * and KSK2. */
int tfhe_cbs_keygen(tfhe_cbs_ctx* ctx, const int32_t* lwe_sk, uint64_t seed);
int tfhe_cbs_get_rlwe_sk(const tfhe_cbs_ctx* ctx, int32_t* out);
/* host copy of PFKS key entry (u, i, t, v), 2N uint32 (keygen parity) */
int tfhe_cbs_get_pfks_entry(const tfhe_cbs_ctx* ctx, int u, int i, int t, int v, uint32_t* out);
/* host copy of linear-key entry (u, i, t) (pfks_linear = 1): 2N words mod 2^40 (a then b) */
int tfhe_cbs_get_pfkl_entry(const tfhe_cbs_ctx* ctx, int u, int i, int t, uint64_t* out);
/* stateless BSK2 generator into d_bsk [n][8][2][N] uint64, or [n/2][3][8][2][N]
 * pattern keys with pbs_multibit = 2 (keygen parity) */
int tfhe_gen_bsk64(const tfhe_cbs_params* p, uint64_t seed, const int32_t* d_lwe_sk, const int32_t* d_rlwe_sk,
                   uint64_t* d_bsk, void* stream);
/* stateless PFKS key generator into d_pfks [2][N+1][8][15][2N] uint32 (the
 * context's key; tfhe_cbs_get_pfks_entry reads entries of it) */
/* stateless HomTrace + SchemeSwitch key generator (lwe_to_rgsw = 1) into d_keys
 * [12][5][2][N] uint64: rows (k, d), k < 11, = RLWE_K2(-sigma_k(K2) 2^(24 + 8d)) with
 * sigma_k: X -> X^(N/2^k + 1), then rows (11, d) = RLWE_K2(K2^2 2^(24 + 8d)); Philox
 * domains 11 and 12, noise rlwe_nshift; synchronises the stream */
int tfhe_gen_htss_keys(const tfhe_cbs_params* p, uint64_t seed, const int32_t* d_rlwe_sk, uint64_t* d_keys,
                       void* stream);
/* stateless linear PFKS key generator (pfks_linear = 1) into d_pfkl, [2][5][2N][16416]
 * int8 (2 x 336 MB): entry (u, i, t) at k = 8 i + t of every column, as 5 signed bytes
 * of its 40-bit words; synchronises the stream */
int tfhe_gen_pfkl40(const tfhe_cbs_params* p, uint64_t seed, const int32_t* d_rlwe_sk, int8_t* d_pfkl, void* stream);
int tfhe_gen_pfks32(const tfhe_cbs_params* p, uint64_t seed, const int32_t* d_rlwe_sk, uint32_t* d_pfks,
                    void* stream);
/* Keys made elsewhere into the context, as GPUCirBTSContext(cc, maxBatchSize)
 * uploads the CPU context's keys (cirbts.cuh:420-426, gpu_init_cirbsk
 * cirbts-init.cu:200-865): rlwe_sk host [N] int32; bsk coefficient rows as
 * tfhe_gen_bsk64 writes them, pfks as tfhe_gen_pfks32, ksk2 [N][8][4][n+1]
 * uint32 ([N][8][n+1] with ks_linear = 1).  The three key buffers may be host or device memory; the BSK2 is
 * transformed to the context's PBS form.  With pfks_linear = 1, pfks points at the
 * linear key as tfhe_gen_pfkl40 writes it. */
int tfhe_cbs_set_keys(tfhe_cbs_ctx* ctx, const int32_t* rlwe_sk, const uint64_t* bsk, const uint32_t* pfks,
                      const uint32_t* ksk2);

/* Key files.  The reference has no file format: its keys come from the OpenFHE
 * fork's CPU context and gpu_init_cirbsk uploads them (cirbts-init.cu:200-865).
 * A key file holds what tfhe_cbs_set_keys takes, behind a versioned 256-byte header
 * (magic "TFHECBSK", version 1, the tfhe_cbs_params, the key seed or 0, the four
 * section sizes and a 64-bit content hash of each, and a hash of the header):
 * rlwe_sk [N] int32, the BSK2 rows as tfhe_gen_bsk64 writes them, the PFKS key of the
 * parameter set's form (as tfhe_gen_pfks32 or tfhe_gen_pfkl40 writes it) and ksk2
 * [N][8][4][n+1] uint32.
 * tfhe_cbs_write_keys: a file from keys made elsewhere (host or device buffers).
 * tfhe_cbs_save_keys: a context's keys, if tfhe_cbs_keygen made them (the BSK2
 * rows are generated again from the seed; TFHE_E_STATE otherwise).
 * tfhe_cbs_load_keys: checks magic, version, header hash, that the parameter set
 * equals the context's (TFHE_E_PARAM otherwise), section sizes and content hashes
 * (TFHE_E_SHAPE), then uploads as tfhe_cbs_set_keys.
 * tfhe_cbs_key_file_info: the parameter set and seed of a file. */
int tfhe_cbs_write_keys(const char* path, const tfhe_cbs_params* params, uint64_t seed, const int32_t* rlwe_sk,
                        const uint64_t* bsk, const void* pfks, const uint32_t* ksk2);
int tfhe_cbs_save_keys(const tfhe_cbs_ctx* ctx, const char* path);
int tfhe_cbs_load_keys(tfhe_cbs_ctx* ctx, const char* path);
int tfhe_cbs_key_file_info(const char* path, tfhe_cbs_params* params, uint64_t* seed);

/* gpu_CircuitBootstrappingBatchToDeviceFromDeviceLWE (cirbts.cuh:466-474,
 * cirbts-batch.cu:1416-1879): d_lwe [batch][n+1] -> d_rgsw [batch][4][2][N] uint32.
 * Pieces: the multi-value PBS (gpu_BootstrapLUT_inplace_batch_device,
 * cirbts-batch.cu:3293) into d_out [batch][2][N+1] uint64 (LWE_K2 of bit * 2^48
 * and bit * 2^56 on the 64-bit torus), and the private functional key switch of
 * count torus-64 LWE_K2 ciphertexts [count][N+1] into torus-32 RLWE rows of 2N
 * uint32 spaced out_stride uint32 apart. */
int tfhe_cbs_batch(tfhe_cbs_ctx* ctx, const uint32_t* d_lwe, uint32_t batch, uint32_t* d_rgsw, void* stream);
int tfhe_cbs_pbs_batch(tfhe_cbs_ctx* ctx, const uint32_t* d_lwe, uint32_t batch, uint64_t* d_out, void* stream);
int tfhe_cbs_pfks_batch(tfhe_cbs_ctx* ctx, int u, const uint64_t* d_lwe64, uint32_t count, size_t out_stride,
                        uint32_t* d_out, void* stream);

/* Borrowed view of circuit-bootstrap outputs held in the context (the reference's
 * DeviceRGSWBatchView, cirbts.cuh:406-418): per ciphertext digits * 2 rows of 2
 * polynomials of N coefficients, element ((row * 2 + poly) * N + coeff).  Here the
 * elements are torus-32 coefficients (uint32), not NTT-domain words.  Valid until
 * the next call on the context that reuses its RGSW buffer (another circuit
 * bootstrap into the context, an AES round or keystream, tfhe_cbs_autotune). */
typedef struct tfhe_rgsw_batch_view {
    const uint32_t* d_rgsw;
    uint32_t batch;
    uint32_t digits;              /* CB levels (2) */
    uint32_t N;
    size_t per_ciphertext_elems;  /* digits * 4 * N */
    void* stream;
} tfhe_rgsw_batch_view;

/* gpu_CircuitBootstrappingBatchToDeviceFromDeviceLWE with the reference's argument
 * layout (cirbts.cuh:466-474, cirbts-batch.cu:1416-1440): masks d_a (u64 words,
 * ciphertext k at d_a + k * a_stride) and bodies d_b [batch] (u64), values on the
 * 32-bit torus (the low 32 bits are used).  The outputs stay in the context; *view
 * describes them.  As in the reference, NULL inputs or batch 0 give an empty view
 * and no error; a_stride < n is TFHE_E_PARAM and batch above the context's
 * capacity TFHE_E_CAPACITY. */
int tfhe_cbs_batch_from_device_lwe(tfhe_cbs_ctx* ctx, const uint64_t* d_a, const uint64_t* d_b, uint32_t batch,
                                   uint32_t a_stride, tfhe_rgsw_batch_view* view, void* stream);

/* gpu_CircuitBootstrappingBatchToDevice(params, ek, vector<LWECiphertext>)
 * (cirbts.cuh:464-466): host LWE ciphertexts h_lwe [batch][n+1] uint32 are staged
 * into the context and bootstrapped; the outputs stay in the context and *view
 * describes them, with the validity rule of tfhe_cbs_batch_from_device_lwe.  NULL
 * input or batch 0 gives an empty view; batch above the capacity is TFHE_E_CAPACITY. */
int tfhe_cbs_batch_to_device(tfhe_cbs_ctx* ctx, const uint32_t* h_lwe, uint32_t batch, tfhe_rgsw_batch_view* view,
                             void* stream);

/* gpu_BootstrapManyLUT(params, ek, ct, LUT, bitwidth) (cirbts.cuh:478-482,
 * cirbts-bootstrap.cu:753-839) for a batch sharing one test polynomial: each level-0
 * LWE (d_lwe [batch][n+1] uint32) is switched to 2N with the bitwidth rule of
 * gpu_SpecilMS (cirbts-utils.cu:52-61: round(v 2N / (2^32 2^bw)) 2^bw), the
 * accumulator starts as (0, LUT X^-b~) with d_lut [N] uint64 on the 64-bit torus,
 * is blind-rotated by the context's PBS form to LUT X^-phase~ (phase~ = b~ - <a~, s>
 * mod 2N; extracting coefficient 0 gives LUT[phase~], negated for phase~ >= N),
 * and Gpow_j / 2 (2^47, 2^55) is added to
 * body coefficient j < 2 (kernel_ModAddGpowScaled, kernel.cu:7459-7473).  d_out
 * [batch][2][N] uint64: the RLWE_K2 (mask, then body) in the coefficient domain
 * (the reference returns it in NTT evaluation form).  bitwidth <= 11. */
int tfhe_cbs_bootstrap_many_lut(tfhe_cbs_ctx* ctx, const uint32_t* d_lwe, uint32_t batch, const uint64_t* d_lut,
                                uint32_t bitwidth, uint64_t* d_out, void* stream);

/* RGSW (torus-32 coefficients, [count][4][2][N] uint32) -> Fourier, one
 * exact FP64 transform per polynomial: d_rgsw_f [count][8][1024] complex doubles. */
int tfhe_rgsw_to_fourier(tfhe_cbs_ctx* ctx, const uint32_t* d_rgsw, uint32_t count, double* d_rgsw_f, void* stream);

/* 8-input CMux vertical-packing LUT on the 32-bit torus (the S-box LUT; CPU
 * semantics of EvalLUT8x32_RLWE_CPU, rlwe_tables.cuh:93-108, GPU form
 * gpu_lut.cuh:1783): for lut k, control RGSWs d_rgsw_f[8k .. 8k+7] (bit 0
 * first), table polynomial d_table + k * table_stride uint32 (coefficient
 * 8x + j = output bit j of entry x, times 2^31); out [nlut][out_bits][N+1]
 * uint32 = the extracted coefficients. */
int tfhe_lut8_eval_batch(tfhe_cbs_ctx* ctx, const double* d_rgsw_f, uint32_t nlut, const uint32_t* d_table,
                         size_t table_stride, int out_bits, uint32_t* d_out32, void* stream);
/* 8 -> 32-bit LUT (the reference's T-table lookup, EvalLUT8x32_RLWE_GPU_DeviceCtrl_Batch,
 * gpu_lut.cuh:1783-2214): byte k's 8 RGSWs (d_rgsw_f as for tfhe_lut8_eval_batch)
 * select entry x of a 256-entry table of 32-bit words held in 4 vertical-packing
 * polynomials d_tables [4][N] (coefficient 8x + j of polynomial p = bit 8p + j of
 * word x, times 2^31).  Output d_out32 [nbytes][32][N+1]: the LWE of bit b of the
 * word, b = 8p + j.  At most 16 * max_blocks bytes per call. */
int tfhe_lut8x32_eval_batch(tfhe_cbs_ctx* ctx, const double* d_rgsw_f, uint32_t nbytes, const uint32_t* d_tables,
                            uint32_t* d_out32, void* stream);

/* N -> n key switch of [count][N+1] uint32 (32-bit torus) LWE_K2 ciphertexts
 * (reference kernel_KeySwitchU16, aes_gpu_ks.cuh:341-370). */
int tfhe_cbs_keyswitch_batch(tfhe_cbs_ctx* ctx, const uint32_t* d_lwe32, uint32_t count, uint32_t* d_out, void* stream);

/* Launch-shape autotuner (no counterpart in the reference, whose shapes are fixed at
 * compile time): times, with events on `stream`, the candidate shapes of one AES
 * round over `blocks` blocks and keeps the fastest in the context for later calls
 * of those sizes.  choices[5] receives the PBS form (0 branchy, 1 branch-free FFT step;
 * with two key bits per step 1 = the wide key batch at one wave per SIMD, 0 = two
 * waves per SIMD, 2 = split over 8 waves; in a pbs_multibit = 3 context 0 / 1 are one
 * bit per step and 2 / 3 / 4 two bits at two waves per SIMD / wide / split) of the
 * full-round part and of the rest of tfhe_cbs_batch's split,
 * the PFKS i-slices of the same two parts, and the N=2048 key switch's i-slices;
 * -1 where a part does not exist.  A part of 16 or more rounds of resident
 * workgroups (8192 inputs on 256 CUs), or whose model shape takes over 250 ms per
 * launch, keeps the model shape without timing others.  Overwrites the context's AES state, RGSW and LUT
 * buffers; drops captured keystream graphs (the next graph call recaptures).  The
 * shapes change no output word of the exact PBS, the PFKS or the key switch.  Not
 * during stream capture. */
int tfhe_cbs_autotune(tfhe_cbs_ctx* ctx, uint32_t blocks, int* choices, void* stream);

/* Pin the PBS launch form used for launches of `inputs` ciphertexts (the numbering of
 * tfhe_cbs_autotune's choices[0]; -1 returns that size to the size model), e.g. to
 * restore a stored autotune result without timing again.  No counterpart in the
 * reference (its shapes are compile-time).  TFHE_E_PARAM for a form the context's PBS
 * does not have.  Drops captured keystream graphs.  The FFT and key-batch forms of one
 * PBS give the same words, except the split form, whose two wave groups' partial sums
 * are added once (words differ by FP64 rounding); in a pbs_multibit = 3 context, choosing
 * one or two key bits per step changes them too (all decrypt within the PBS noise). */
int tfhe_cbs_set_pbs_form(tfhe_cbs_ctx* ctx, uint32_t inputs, int form);

/* Kernel timing of the multi-value PBS (the reference prints per-block
 * compute_only_ms from events around its launches, main.cu:1694-1758): while
 * enabled, every PBS launch made outside stream capture (tfhe_cbs_pbs_batch and
 * everything built on it: circuit bootstraps, AES rounds, keystreams) is
 * bracketed by a pair of HIP events on the stream it is launched on.
 * tfhe_cbs_timing_read waits for the recorded events and returns the summed
 * launch durations (ms), the number of launches and their PBS inputs since
 * enabling or the last read, then starts a new count.  Enabling resets the
 * count. */
int tfhe_cbs_timing_enable(tfhe_cbs_ctx* ctx, int on);
int tfhe_cbs_timing_read(tfhe_cbs_ctx* ctx, double* pbs_ms, uint32_t* launches, uint64_t* inputs);

/* Exact negacyclic product through the 4-wave FP64 FFT, N = 2048:
 * d_out = d_x * d_y mod (X^N + 1, 2^64), small-integer d_x, torus-64 d_y. */
int tfhe_polymul2048_batch(const int32_t* d_x, const uint64_t* d_y, uint32_t batch, uint64_t* d_out, void* stream);
/* The same product in the one-product form of the PBS (pbs_approx = 1): y scaled
 * to the torus, one FP64 product, rounded once; its difference from
 * tfhe_polymul2048_batch is the per-product rounding error of that form (the
 * "max coefficient error against the exact integer product" of the reference's
 * split-precision FFT, cirbts-fft.cu:159-226). */
int tfhe_polymul2048_approx_batch(const int32_t* d_x, const uint64_t* d_y, uint32_t batch, uint64_t* d_out,
                                  void* stream);
/* Split-precision FP32 external-product arithmetic (the reference's split-FFT
 * backend: limb rule cirbts-init.cu:1096-1120, limb extraction cirbts-fft.cu:345-355,
 * exact integer recombination cirbts-fft.cu:159-226, carried to FP32):
 * d_out = d_x * d_y mod (X^2048 + 1, 2^32) for digit polynomials d_x (|x| <= 256)
 * and torus-32 d_y, all [batch][2048].  d_y is cut into balanced signed limbs, each
 * limb's product taken through the FP32 FFT1024 and rounded to an integer, and the
 * limbs recombined mod 2^32.  form 0: limbs of 11, 11, 10 bits (three limbs of at
 * most 11 bits); form 1: 11, 6, 5, 5, 5 bits, so that only the lowest limb's product
 * can round wrongly (the exact-top-limb form); form 2: one FP64 product (exact), the
 * baseline.  The difference from the exact product is the form's coefficient error. */
int tfhe_polymul2048_split32_batch(const int32_t* d_x, const uint32_t* d_y, uint32_t batch, int form,
                                   uint32_t* d_out, void* stream);

/* ------------------------------------------------------------------------- */
/* AES-128 transciphering (reference bitwise/examples/aes_transcipher/main.cu */
/* RunTranscipherBatch :1316-1619).  State bit 8*byte + j (byte = row + 4 col)*/
/* is a level-0 LWE of that bit * 2^31.                                       */
/* ------------------------------------------------------------------------- */
/* client side: the 11 round keys' bits encrypted under sk, [11][128][n+1]
 * (reference BuildEncryptedKeyMaterial, key_material.cuh:12-47) */
void tfhe_aes_encrypt_key_bits(int n, const int32_t* sk, uint64_t seed, const uint8_t key[16], int nshift,
                               uint32_t* out);
/* plain AES-128 block encryption (the decryption check of the pipeline) */
void tfhe_aes128_encrypt_plain(const uint8_t key[16], const uint8_t in[16], uint8_t out[16]);
int tfhe_aes_set_round_keys(tfhe_cbs_ctx* ctx, const uint32_t* rk_bits);
/* one round (1..10): SubBytes by circuit bootstrap + S-box LUT + key switch,
 * then ShiftRows, MixColumns (rounds 1-9) and AddRoundKey as LWE additions
 * (reference run_round_batch, main.cu:1438-1554); [blocks][128][n+1] */
int tfhe_aes_round_batch(tfhe_cbs_ctx* ctx, int round, const uint32_t* d_in, uint32_t blocks, uint32_t* d_out,
                         void* stream);
/* CTR keystream: out [blocks][128][n+1] = LWE bits of AES_k(h_ctr[blk]) with the
 * public counter blocks h_ctr [blocks][16] (host) */
int tfhe_aes_ctr_keystream_batch(tfhe_cbs_ctx* ctx, const uint8_t* h_ctr, uint32_t blocks, uint32_t* d_out,
                                 void* stream);
/* The same keystream in the per-bootstrap form: every circuit bootstrap, S-box
 * LUT and key switch is its own launch of one (the reference's single-ciphertext
 * latency path gpu_CircuitBootstrapping, cirbts-bootstrap.cu:89-596, and
 * EvalLUT8x32_RLWE_GPU, gpu_lut.cuh:1436).  Results equal
 * tfhe_aes_ctr_keystream_batch bit for bit; it is the plain form the batched
 * ones are measured against. */
int tfhe_aes_ctr_keystream_plain(tfhe_cbs_ctx* ctx, const uint8_t* h_ctr, uint32_t blocks, uint32_t* d_out,
                                 void* stream);
/* The same keystream replayed from a captured hipGraph (the reference captures
 * its LUT and circuit-bootstrap launch sequences as CUDA graphs with thread-local
 * capture, gpu_lut.cuh:1859-2039, cirbts-batch.cu:559-581): the first
 * call for a block count captures state init + the 10 rounds on the context's own
 * stream, writing the context's state buffers; every call uploads the counters on
 * `stream`, replays, and copies the keystream to d_out, ordered after and before
 * `stream` by events.  One graph per block count serves any d_out; the context keeps
 * the 8 most recently used.  Results equal tfhe_aes_ctr_keystream_batch bit for
 * bit.  Not thread-safe (the context caches the graphs). */
int tfhe_aes_ctr_keystream_graph(tfhe_cbs_ctx* ctx, const uint8_t* h_ctr, uint32_t blocks, uint32_t* d_out,
                                 void* stream);
/* AES-128 CTR transciphering (the CTR path of RunTranscipherBatch, main.cu:1555-1580):
 * the keystream of the counter blocks h_ctr [blocks][16], plus the public AES-CTR
 * ciphertext bytes h_cipher [blocks][16] (host) added as public bits, so that d_out
 * [blocks][128][n+1] holds LWE encryptions of the plaintext bits (bit 8 byte + j).
 * mode 0 = batched keystream, 1 = hipGraph replay, 2 = per-bootstrap form. */
int tfhe_aes_ctr_transcipher_batch(tfhe_cbs_ctx* ctx, int mode, const uint8_t* h_ctr, const uint8_t* h_cipher,
                                   uint32_t blocks, uint32_t* d_out, void* stream);
/* AES round form of the context (both directions):
 * 0 (default) = SubBytes by the 8 -> 8 S-box LUT, then ShiftRows / MixColumns /
 *     AddRoundKey as LWE additions after the key switch;
 * 1 = T-table rounds, the reference's data flow (run_round_batch, main.cu:1438-1554;
 *     EvalLUT8x32_RLWE_GPU_DeviceCtrl_Batch, gpu_lut.cuh:1783-2214): rounds 1..9 look
 *     each byte up in its row's T-table (Te_r, or Td_r for decryption; 8 -> 32 bits,
 *     four vertical-packing polynomials on the byte's 8 RGSWs), XOR the four words of
 *     each output column at dimension N (the ShiftRows / InvShiftRows word map), key-
 *     switch each of the 128 bits and add the round key; round 10 is the S-box form.
 *     Each bit then carries the noise of four LUTs, against one LUT plus up to 7 (19
 *     for decryption) key-switched LWEs in form 0.
 * The two forms give different words that decrypt to the same bits.  Changing the
 * form drops the context's captured keystream graphs.  The per-bootstrap keystream
 * (tfhe_aes_ctr_keystream_plain, mode 2) is form 0 only. */
int tfhe_aes_set_round_form(tfhe_cbs_ctx* ctx, int form);
/* AES-128 decryption (the reference's default transciphering direction: the server
 * decrypts public AES ciphertext under the encrypted key with Td tables and
 * InvMixColumns round keys, main.cu:279-317, key_material.cuh:12-47 rk_imc_words).
 * Client side: the decryption schedule of the equivalent inverse cipher (FIPS-197
 * 5.3.5: dk[0] = rk[10], dk[r] = InvMixColumns(rk[10-r]), dk[10] = rk[0]), its bits
 * encrypted under sk, [11][128][n+1]. */
void tfhe_aes_encrypt_dec_key_bits(int n, const int32_t* sk, uint64_t seed, const uint8_t key[16], int nshift,
                                   uint32_t* out);
int tfhe_aes_set_dec_round_keys(tfhe_cbs_ctx* ctx, const uint32_t* dk_bits);
/* one round (1..10) of the equivalent inverse cipher on [blocks][128][n+1] */
int tfhe_aes_dec_round_batch(tfhe_cbs_ctx* ctx, int round, const uint32_t* d_in, uint32_t blocks, uint32_t* d_out,
                             void* stream);
/* d_out [blocks][128][n+1] = LWE bits of AES-128^-1_k(h_cipher[blk]) for public
 * ciphertext blocks h_cipher [blocks][16] (host): the state starts as the public bits
 * plus dk[0], then the 10 rounds.  Uneven batches run as two block groups as the
 * keystream does. */
int tfhe_aes_decrypt_batch(tfhe_cbs_ctx* ctx, const uint8_t* h_cipher, uint32_t blocks, uint32_t* d_out,
                           void* stream);
/* XOR of public bits into LWE bits: d_ct[i] += (0, bit_i * 2^31), d_pub packed bits */
int tfhe_lwe_xor_public(uint32_t* d_ct, const uint8_t* d_pub, uint32_t nbits, int n, void* stream);

/* ------------------------------------------------------------------------- */
/* Host helpers standing in for the OpenFHE fork's KeyGen/Encrypt/Decrypt     */
/* (reference bitwise/examples/aes_transcipher/main.cu:232-235, 626, 1287).   */
/* ------------------------------------------------------------------------- */
void tfhe_lwe_keygen(int n, uint64_t seed, int32_t* sk);
void tfhe_rlwe_keygen(int N, uint64_t seed, uint32_t domain, int32_t* sk);
/* ct = (a, <a,s> + m * 2^32/p + e), a and e from (seed, id) */
void tfhe_lwe_encrypt(int n, const int32_t* sk, uint64_t seed, uint32_t id, uint32_t m, uint32_t p,
                      int nshift, uint32_t* ct);
/* round(phase * p / 2^32) mod p */
uint32_t tfhe_lwe_decrypt(int n, const int32_t* sk, const uint32_t* ct, uint32_t p);

#ifdef __cplusplus
}
#endif
#endif
