/*
 * tfhe_oracle.h -- CPU restatement of the reference's TFHE bootstrapping path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  Nothing in the shipped library links, calls or executes it.
 *
 * Every function cites the reference file:line whose behaviour it restates.
 * The reference's CPU half (the OpenFHE fork) is absent from the reference tree,
 * so the arithmetic domain is the TFHE torus (Z/2^32 for the gate path and the
 * level-0 LWE, Z/2^64 for the N=2048 circuit-bootstrapping level); the
 * reference's formulas (decomposition, CGGI accumulator update, sample
 * extraction, key switching, CMux) are followed in that domain.
 *
 * Each function is pinned by a check that does not retype it (tests/test_oracle_*.py):
 * published vectors (Philox, FIPS-197), schoolbook products, closed forms
 * (modswitch, monomials, blind-rotation phase), decryption of every probed key
 * row to its message, and explicit sums over key tables.  What no check pins is
 * parity with the reference's own ciphertext words: its key generator, noise
 * sampler and PRNG stream live in the absent OpenFHE fork, so the samplers here
 * (Philox domains, Irwin-Hall noise) are this library's own and the ciphertext
 * words are "parity unpinned" against the reference.
 */
#ifndef TFHE_ORACLE_H
#define TFHE_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Philox4x32-10 (Salmon et al., SC'11) and the key-material sampler ---- */
void orc_philox4x32_10(const uint32_t key[2], const uint32_t ctr[4], uint32_t out[4]);
uint32_t orc_rand32(uint64_t seed, uint32_t dom, uint32_t id, uint32_t sub, uint32_t j);
uint64_t orc_rand64(uint64_t seed, uint32_t dom, uint32_t id, uint32_t j);
/* Irwin-Hall(12) integer noise: round(z * 2^-nshift) with z = sum of 12 u32 - 6*2^32. */
int64_t orc_noise(uint64_t seed, uint32_t dom, uint32_t id, uint32_t j, int nshift);
int32_t orc_secret_bit(uint64_t seed, uint32_t dom, uint32_t i);

/* sampler domains: the same numbering is used by the product's key generator */
enum {
    ORC_DOM_LWE_SK = 1, ORC_DOM_RLWE1_SK = 2, ORC_DOM_RLWE2_SK = 3,
    ORC_DOM_BSK1 = 4, ORC_DOM_KSK1 = 5, ORC_DOM_BSK2 = 6, ORC_DOM_PFKS = 7,
    ORC_DOM_KSK2 = 8, ORC_DOM_BSK2MB = 9, ORC_DOM_PFKL = 10, ORC_DOM_HTKEY = 11, ORC_DOM_SSKEY = 12,
    ORC_DOM_KSKL = 13,
    ORC_DOM_ENC = 16
};

/* ---- parameter sets ---- */
typedef struct {
    int n;            /* LWE dimension */
    int N;            /* ring dimension */
    int bk_logB, bk_l;   /* blind-rotation gadget */
    int ks_logB, ks_l;   /* key-switching gadget */
    int lwe_nshift;      /* LWE noise: sigma = 2^-(32 - (32 - nshift)) -> see orc_noise */
    int rlwe_nshift;     /* RLWE (BSK) noise shift */
} orc_gate_params;

/* ---- negacyclic products ---- */
void orc_nc_mul_exact32(const int32_t* x, const uint32_t* y, uint32_t* out, int N);
void orc_nc_mul_exact64(const int32_t* x, const uint64_t* y, uint64_t* out, int N);
double orc_nc_mul_fft(const int64_t* x, const int64_t* y, int64_t* out, int N);
void orc_fft_forward(const double* x /*[N]*/, double* re, double* im, int N);
void orc_fft_inverse(const double* re, const double* im, double* x /*[N]*/, int N);

/* ---- gadget decomposition ---- */
void orc_decompose32(uint32_t t, int logB, int l, int32_t* digits);
void orc_decompose64(uint64_t t, int logB, int l, int32_t* digits);

/* ---- key generation (torus-32 gate level) ---- */
void orc_gen_lwe_sk(uint64_t seed, int n, int32_t* s);
void orc_gen_rlwe_sk(uint64_t seed, uint32_t dom, int N, int32_t* K);
void orc_rlwe_encrypt32(uint64_t seed, uint32_t dom, uint32_t id, int N, const int32_t* K,
                        const uint32_t* msg_a, const uint32_t* msg_b, int nshift, uint32_t* ct);
void orc_lwe_encrypt32(uint64_t seed, uint32_t dom, uint32_t id, int n, const int32_t* s,
                       uint32_t mu, int nshift, uint32_t* ct);
uint32_t orc_lwe_phase32(int n, const int32_t* s, const uint32_t* ct);
void orc_gen_bsk32(uint64_t seed, const orc_gate_params* p, const int32_t* s, const int32_t* K,
                   uint32_t* bsk);
void orc_gen_ksk32(uint64_t seed, const orc_gate_params* p, const int32_t* s, const int32_t* K,
                   uint32_t* ksk);

/* ---- gate bootstrap pieces ---- */
void orc_bsk_to_fourier(const orc_gate_params* p, const uint32_t* bsk, double* bsk_f);
uint32_t orc_modswitch_2N(uint32_t v, int N);
void orc_gate_test_vector(int N, int gate, uint32_t b_tilde, uint32_t* tv);
double orc_external_product32(int N, int logB, int l, const double* rgsw_f, const uint32_t* acc,
                              uint32_t* out);
void orc_monic_update32(int N, uint32_t m, const uint32_t* t, uint32_t* acc);
double orc_blind_rotate32(const orc_gate_params* p, const double* bsk_f, const uint32_t* a_tilde,
                          uint32_t* acc);
void orc_sample_extract32(int N, const uint32_t* acc, uint32_t idx, uint32_t* lwe);
void orc_keyswitch32(const orc_gate_params* p, const uint32_t* ksk, const uint32_t* lweN,
                     uint32_t* out);
int orc_gate_prep(int gate, int n, const uint32_t* ct1, const uint32_t* ct2, uint32_t* prep);
double orc_gate_bootstrap(const orc_gate_params* p, const double* bsk_f, const uint32_t* ksk,
                          int gate, const uint32_t* ct1, const uint32_t* ct2, uint32_t* out);

void orc_gen_ksk32_dom(uint64_t seed, uint32_t dom, const orc_gate_params* p, const int32_t* s,
                       const int32_t* K, uint32_t* ksk);

/* programmable bootstrap and large-precision functions of the gate level, func_oracle.c */
int orc_func_property(const uint32_t* lut, uint64_t q);
int orc_bootstrap_func(const orc_gate_params* p, const double* bsk_f, const uint32_t* ksk, const uint32_t* ct,
                       uint64_t ctmod, const uint32_t* f, uint64_t fmod, uint32_t* out);
int orc_eval_func(const orc_gate_params* p, const double* bsk_f, const uint32_t* ksk, const uint32_t* ct,
                  uint64_t q, const uint32_t* lut, uint32_t beta, uint32_t* out);
int orc_eval_floor(const orc_gate_params* p, const double* bsk_f, const uint32_t* ksk, const uint32_t* ct,
                   uint64_t mod, uint32_t beta, uint32_t roundbits, uint64_t qsmall, uint32_t* out);
int orc_eval_sign(const orc_gate_params* p, const double* bsk_f, const uint32_t* ksk, const uint32_t* ct,
                  uint64_t mod, uint32_t beta, uint64_t qsmall, int scheme_switch, uint32_t* out);
int orc_eval_decomp(const orc_gate_params* p, const double* bsk_f, const uint32_t* ksk, const uint32_t* ct,
                    uint64_t mod, uint32_t beta, uint64_t qsmall, uint32_t* out, int max_digits);

/* ---- circuit-bootstrapping level (N = 2048; PBS torus-64, PFKS/RGSW/LUT torus-32), cbs_oracle.c ---- */
typedef struct {
    int n, N;             /* 636, 2048 */
    int bk_logB, bk_l;    /* PBS gadget 2^9 x 4 */
    int cb_logB, cb_l;    /* circuit-bootstrapping level gadget 2^8 x 2 */
    int pf_logB, pf_l;    /* private functional KS 2^4 x 8 */
    int ks_logB, ks_l;    /* N -> n key switch 2^2 x 8 (torus-32) */
    int lwe_nshift;       /* level-0 LWE noise (torus-32): 15 */
    int rlwe_nshift;      /* level-2 noise (torus-64): 19, i.e. sigma = 2^-51 */
    int pf_nshift;        /* PFKS key noise (torus-32): 30, i.e. sigma = 2^-30 (4 units of 2^-32) */
} orc_cbs_params;

void orc_to_fourier64(const uint64_t* poly, int N, double* out);
void orc_rlwe_encrypt64(uint64_t seed, uint32_t dom, uint32_t id, int N, const int32_t* K,
                        const uint64_t* msg_a, const uint64_t* msg_b, int nshift, uint64_t* ct);
void orc_rgsw_encrypt64(uint64_t seed, uint32_t dom, uint32_t id0, int N, const int32_t* K, int m, int logB,
                        int l, int nshift, uint64_t* rgsw);
void orc_gen_bsk64(uint64_t seed, const orc_cbs_params* p, const int32_t* s, const int32_t* K2, uint64_t* bsk);
void orc_bsk64_to_fourier(const orc_cbs_params* p, const uint64_t* bsk, double* bsk_f);
double orc_external_product64(int N, int logB, int l, const double* rgsw_f, const uint64_t* acc, uint64_t* out);
void orc_monic_update64(int N, uint32_t m, const uint64_t* t, uint64_t* acc);
double orc_blind_rotate64(const orc_cbs_params* p, const double* bsk_f, const uint32_t* a_tilde, uint64_t* acc);
void orc_sample_extract64(int N, const uint64_t* acc, uint32_t idx, uint64_t* lwe);
uint32_t orc_special_ms(uint32_t v, int N, int bw);
double orc_pbs_manylut(const orc_cbs_params* p, const double* bsk_f, const uint32_t* ct32, uint64_t* out);
/* two key bits per step (pattern keys of the indicators of (1,0), (0,1), (1,1)), cbs_oracle.c */
uint32_t orc_mb2_indicator(const int32_t* s, int k, int pat);
void orc_gen_bsk64_mb2_row(uint64_t seed, const orc_cbs_params* p, const int32_t* s, const int32_t* K2,
                           uint32_t q, int r, uint64_t* row);
void orc_gen_bsk64_mb2(uint64_t seed, const orc_cbs_params* p, const int32_t* s, const int32_t* K2, uint64_t* bsk);
void orc_rgsw64_to_fourier(const orc_cbs_params* p, long nrgsw, const uint64_t* bsk, double* bsk_f);
double orc_blind_rotate64_mb2(const orc_cbs_params* p, const double* bsk_f, const uint32_t* a_tilde, uint64_t* acc);
double orc_pbs_manylut_mb(const orc_cbs_params* p, const double* bsk_f, int mb, const uint32_t* ct32,
                          uint64_t* out);
/* gpu_BootstrapManyLUT with test polynomial lut [N] and bitwidth bw: out RLWE [2][N] */
double orc_bootstrap_many_lut(const orc_cbs_params* p, const double* bsk_f, int mb, const uint32_t* ct32,
                              const uint64_t* lut, int bw, uint64_t* out);
double orc_blind_rotate_lut64(const orc_cbs_params* p, const double* bsk_f, int mb, const uint32_t* ct32,
                              const uint64_t* lut, int bw, uint64_t* out);
/* HomTrace + SchemeSwitch (gadget 2^8 x 5) */
void orc_automorphism64(int N, const uint64_t* in, uint32_t k, uint64_t* out);
void orc_gen_ht_key(uint64_t seed, const orc_cbs_params* p, const int32_t* K2, uint64_t* key);
void orc_gen_ss_key(uint64_t seed, const orc_cbs_params* p, const int32_t* K2, uint64_t* key);
void orc_ksk64_to_fourier(const orc_cbs_params* p, int nkeys, const uint64_t* key, double* key_f);
double orc_homtrace64(const orc_cbs_params* p, const double* htkey_f, uint64_t* ct);
double orc_schemeswitch64(const orc_cbs_params* p, const double* sskey_f, const uint64_t* ct, uint64_t* out);
double orc_cbs_htss(const orc_cbs_params* p, const double* bsk_f, const double* htkey_f, const double* sskey_f,
                    const uint32_t* ct32, uint32_t* rgsw);
uint32_t orc_pfks_id(const orc_cbs_params* p, int u, int i, int t, int v);
void orc_gen_pfks_entry(uint64_t seed, const orc_cbs_params* p, const int32_t* K2, int u, int i, int t, int v,
                        uint32_t* ct);
void orc_pfks(uint64_t seed, const orc_cbs_params* p, const int32_t* K2, int u, const uint64_t* lwe, uint32_t* out);
double orc_cbs(uint64_t seed, const orc_cbs_params* p, const int32_t* K2, const double* bsk_f,
               const uint32_t* ct32, uint32_t* rgsw);
/* linear-key PFKS on Z/2^40 (one key row per (i, t), signed digits), see cbs_oracle.c */
enum { ORC_PFKL_BITS = 40, ORC_PFKL_NSHIFT = 25 };
void orc_gen_pfkl_entry(uint64_t seed, const orc_cbs_params* p, const int32_t* K2, int u, int i, int t,
                        uint64_t* ct40);
void orc_pfks_linear(uint64_t seed, const orc_cbs_params* p, const int32_t* K2, int u, const uint64_t* lwe,
                     uint32_t* out);
double orc_cbs_linear(uint64_t seed, const orc_cbs_params* p, const int32_t* K2, const double* bsk_f,
                      const uint32_t* ct32, uint32_t* rgsw);
void orc_rgsw_to_fourier32(int N, int rows, const uint32_t* rgsw, double* rgsw_f);
double orc_cmux32(int N, int logB, int l, const double* rgsw_f, const uint32_t* d0, const uint32_t* d1,
                  uint32_t* out);
double orc_vertical_packing(const orc_cbs_params* p, int stride, const double* rgsw_f, const uint32_t* table_poly,
                            uint32_t* acc);
void orc_ks2(const orc_cbs_params* p, const uint32_t* ksk2, const uint64_t* lwe64, uint32_t* out);
void orc_gen_ksk2(uint64_t seed, const orc_cbs_params* p, const int32_t* s, const int32_t* K2, uint32_t* ksk);
/* the N -> n key switch with a linear key (one row per (i, j)) and signed digits, cbs_oracle.c */
void orc_gen_kskl_entry(uint64_t seed, const orc_cbs_params* p, const int32_t* s, const int32_t* K2, int i, int j,
                        uint32_t* row);
void orc_gen_kskl(uint64_t seed, const orc_cbs_params* p, const int32_t* s, const int32_t* K2, uint32_t* kl);
void orc_ks2_linear32(const orc_cbs_params* p, const uint32_t* kl, const uint32_t* lwe32, uint32_t* out);
void orc_ks2_linear(const orc_cbs_params* p, const uint32_t* kl, const uint64_t* lwe64, uint32_t* out);

/* ---- plain AES-128 (FIPS-197) ---- */
uint8_t orc_aes_sbox(uint8_t x);
void orc_aes128_expand(const uint8_t key[16], uint8_t rk[176]);
void orc_aes128_encrypt(const uint8_t rk[176], const uint8_t in[16], uint8_t out[16]);
uint8_t orc_aes_inv_sbox(uint8_t y);
uint32_t orc_aes_ttable(int dec, int r, uint8_t x);
void orc_aes128_decrypt(const uint8_t rk[176], const uint8_t in[16], uint8_t out[16]);
void orc_aes128_dec_schedule(const uint8_t rk[176], uint8_t dk[176]);
void orc_aes128_decrypt_ttable(const uint8_t dk[176], const uint8_t in[16], uint8_t out[16]);
void orc_aes128_encrypt_ttable(const uint8_t rk[176], const uint8_t in[16], uint8_t out[16]);

#ifdef __cplusplus
}
#endif
#endif
