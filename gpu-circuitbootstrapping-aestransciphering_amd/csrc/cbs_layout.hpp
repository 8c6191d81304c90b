// cbs_layout.hpp -- parameter checks and key / scratch sizes of the circuit-
// bootstrapping context, shared by its C-ABI translation units.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include "../../include/tfhe_amd.h"

namespace tfhe {

constexpr int kN = 2048;

// lwe_to_rgsw = 1: HomTrace + SchemeSwitch, which needs the exact PBS with a 2^10 x 4 gadget
inline bool htss(const tfhe_cbs_params& p) { return p.lwe_to_rgsw == 1; }

inline bool cbs_params_supported(const tfhe_cbs_params* p) {
    return p && p->N == kN && p->bk_l == 4 && p->cb_logB == 8 && p->cb_l == 2 && p->pf_logB == 4 &&
           p->pf_l == 8 && p->ks_logB == 2 && p->ks_l == 8 && p->n > 0 && p->n <= 1024 && p->lwe_nshift >= 1 &&
           p->lwe_nshift <= 31 && p->rlwe_nshift >= -31 && p->rlwe_nshift <= 31 && p->pf_nshift >= 1 &&
           p->pf_nshift <= 31 && (p->pbs_approx == 0 || p->pbs_approx == 1) && p->pbs_multibit >= 0 &&
           p->pbs_multibit <= 3 && (p->pbs_multibit < 2 || (p->pbs_approx == 1 && p->n % 2 == 0)) &&
           (p->pfks_linear == 0 || p->pfks_linear == 1) && (p->lwe_to_rgsw == 0 || p->lwe_to_rgsw == 1) &&
           (p->ks_linear == 0 || (p->ks_linear == 1 && p->n + 1 <= 640)) &&
           (p->lwe_to_rgsw == 0 ? p->bk_logB == 9
                                : p->bk_logB == 10 && p->pbs_approx == 0 && p->pbs_multibit <= 1 && p->pfks_linear == 0);
}

// HomTrace + SchemeSwitch keys: 11 trace rounds and the scheme switch, 5 gadget rows each
constexpr int kHtRounds = 11, kKsL = 5;
inline size_t htss_key_words() { return (size_t)(kHtRounds + 1) * kKsL * 2 * kN; }  // coefficient form, u64

inline int pbs_limbs(const tfhe_cbs_params& p) { return p.pbs_approx ? 1 : 2; }
inline int pbs_mb(const tfhe_cbs_params& p) { return p.pbs_multibit == 2 ? 2 : 1; }
// pbs_multibit = 3: both key forms, one or two key bits per step chosen per launch size
inline bool pbs_auto(const tfhe_cbs_params& p) { return p.pbs_multibit == 3; }
// RGSWs of the BSK2: one per key bit, three pattern keys per key-bit pair, or both
inline size_t bsk2_rgsw(const tfhe_cbs_params& p) {
    return pbs_auto(p) ? (size_t)p.n + (size_t)p.n / 2 * 3 : pbs_mb(p) == 2 ? (size_t)p.n / 2 * 3 : (size_t)p.n;
}
// first two-bit pattern RGSW in the BSK2 of an auto context (after the n one-bit RGSWs)
inline size_t bsk2_mb2_first(const tfhe_cbs_params& p) { return pbs_auto(p) ? (size_t)p.n : 0; }
// capacity bound: keeps every 32-bit count (2 x inputs for the PFKS) and every grid y
// dimension (padded inputs / 64) of the path in range
constexpr uint32_t kMaxBlocks = 4096;
// captured keystream graphs kept per context
constexpr size_t kMaxGraphs = 8;
// inputs per linear-form PFKS launch (digit scratch pfkl_digit_bytes(kPfklChunk) = 134 MB)
constexpr uint32_t kPfklChunk = 8192;
inline size_t bsk2_cplx(const tfhe_cbs_params& p) { return bsk2_rgsw(p) * 8 * 2 * pbs_limbs(p) * 1024; }
// coefficient-form BSK2 words (uint64), as tfhe_gen_bsk64 writes them
inline size_t bsk2_words(const tfhe_cbs_params& p) { return bsk2_rgsw(p) * 8 * 2 * kN; }
inline size_t pfks_words_per_u() { return (size_t)(kN + 1) * 8 * 15 * 2 * kN; }
// N -> n key-switching key words (uint32): [N][8][4][n+1] lookup form, [N][8][n+1] linear form
inline size_t ksk2_words(const tfhe_cbs_params& p) { return (size_t)kN * 8 * (p.ks_linear ? 1 : 4) * (p.n + 1); }

// the context's parameter set (capi_cbs.cpp)
const tfhe_cbs_params* cbs_ctx_params(const tfhe_cbs_ctx* c);
// a key file from the four sections rlwe_sk, bsk2 rows, PFKS key, ksk2 (host or device
// pointers; keyblob.cpp)
int write_key_blob(const char* path, const tfhe_cbs_params& p, uint64_t seed, const void* const (&src)[4]);

}  // namespace tfhe
