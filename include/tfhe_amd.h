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
