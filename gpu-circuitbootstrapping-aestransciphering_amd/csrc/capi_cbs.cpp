// capi_cbs.cpp -- C ABI of the circuit-bootstrapping context and the AES-128
// transciphering pipeline built on it (include/tfhe_amd.h).
#include <hip/hip_runtime.h>

#include <cstring>
#include <algorithm>
#include <map>
#include <string>
#include <vector>

#include "../../include/tfhe_amd.h"
#include "capi_common.hpp"
#include "cbs_layout.hpp"
#include "kernels.hpp"

using namespace tfhe;

namespace {
// ---- AES-128 pieces computed by the library itself (client-side helpers and
// the bit-level linear layer); the S-box is GF(2^8) inversion + affine map ----
uint8_t gmul(uint8_t a, uint8_t b) {
    uint8_t r = 0;
    while (b) {
        if (b & 1) r ^= a;
        a = (uint8_t)((a << 1) ^ ((a & 0x80) ? 0x1b : 0));
        b >>= 1;
    }
    return r;
}
uint8_t sbox(uint8_t x) {
    uint8_t inv = 0;
    if (x) {
        uint8_t r = 1, b = x;
        for (int e = 254; e; e >>= 1) {
            if (e & 1) r = gmul(r, b);
            b = gmul(b, b);
        }
        inv = r;
    }
    uint8_t res = 0x63, s = inv;
    for (int i = 0; i < 5; ++i) {
        res ^= s;
        s = (uint8_t)((s << 1) | (s >> 7));
    }
    return res;
}
void key_expand(const uint8_t key[16], uint8_t rk[176]) {
    std::memcpy(rk, key, 16);
    uint8_t rcon = 1;
    for (int i = 4; i < 44; ++i) {
        uint8_t t[4];
        std::memcpy(t, rk + 4 * (i - 1), 4);
        if (i % 4 == 0) {
            const uint8_t t0 = t[0];
            t[0] = (uint8_t)(sbox(t[1]) ^ rcon);
            t[1] = sbox(t[2]);
            t[2] = sbox(t[3]);
            t[3] = sbox(t0);
            rcon = gmul(rcon, 2);
        }
        for (int k = 0; k < 4; ++k) rk[4 * i + k] = (uint8_t)(rk[4 * (i - 4) + k] ^ t[k]);
    }
}
// ShiftRows then (optionally) MixColumns on 16 bytes, byte index = row + 4 col
void shift_mix(const uint8_t in[16], uint8_t out[16], bool mix) {
    uint8_t t[16];
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) t[r + 4 * c] = in[r + 4 * ((c + r) % 4)];
    if (!mix) {
        std::memcpy(out, t, 16);
        return;
    }
    for (int c = 0; c < 4; ++c) {
        const uint8_t* a = t + 4 * c;
        out[4 * c + 0] = (uint8_t)(gmul(a[0], 2) ^ gmul(a[1], 3) ^ a[2] ^ a[3]);
        out[4 * c + 1] = (uint8_t)(a[0] ^ gmul(a[1], 2) ^ gmul(a[2], 3) ^ a[3]);
        out[4 * c + 2] = (uint8_t)(a[0] ^ a[1] ^ gmul(a[2], 2) ^ gmul(a[3], 3));
        out[4 * c + 3] = (uint8_t)(gmul(a[0], 3) ^ a[1] ^ a[2] ^ gmul(a[3], 2));
    }
}
constexpr int kLinSrc = 8;  // max sources per output bit of ShiftRows+MixColumns

// S^-1(y): GF(2^8) inverse of the inverse affine map of y ^ 0x63 (rotations by 1, 3, 6)
uint8_t inv_sbox(uint8_t y) {
    const uint8_t v = (uint8_t)(y ^ 0x63);
    auto rl = [](uint8_t x, int k) { return (uint8_t)((x << k) | (x >> (8 - k))); };
    const uint8_t x = (uint8_t)(rl(v, 1) ^ rl(v, 3) ^ rl(v, 6));
    uint8_t r = x ? 1 : 0, b = x;
    for (int e = 254; x && e; e >>= 1) {
        if (e & 1) r = gmul(r, b);
        b = gmul(b, b);
    }
    return r;
}
// InvShiftRows then (optionally) InvMixColumns on 16 bytes, byte index = row + 4 col
void inv_shift_mix(const uint8_t in[16], uint8_t out[16], bool mix) {
    uint8_t t[16];
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) t[r + 4 * ((c + r) % 4)] = in[r + 4 * c];
    if (!mix) {
        std::memcpy(out, t, 16);
        return;
    }
    for (int c = 0; c < 4; ++c) {
        const uint8_t* a = t + 4 * c;
        out[4 * c + 0] = (uint8_t)(gmul(a[0], 14) ^ gmul(a[1], 11) ^ gmul(a[2], 13) ^ gmul(a[3], 9));
        out[4 * c + 1] = (uint8_t)(gmul(a[0], 9) ^ gmul(a[1], 14) ^ gmul(a[2], 11) ^ gmul(a[3], 13));
        out[4 * c + 2] = (uint8_t)(gmul(a[0], 13) ^ gmul(a[1], 9) ^ gmul(a[2], 14) ^ gmul(a[3], 11));
        out[4 * c + 3] = (uint8_t)(gmul(a[0], 11) ^ gmul(a[1], 13) ^ gmul(a[2], 9) ^ gmul(a[3], 14));
    }
}
constexpr int kLinSrcDec = 20;  // max sources per output bit of InvShiftRows+InvMixColumns (19)
// T-table word of row r (reference aes128.cuh:227-276): Te_r[x] = rotr(be(2 S, S, S, 3 S), 8 r),
// Td_r[x] = rotr(be(14 S^-1, 9 S^-1, 13 S^-1, 11 S^-1), 8 r); byte p of the word (row p of the
// output column) is bits 24 - 8 p .. 31 - 8 p
uint32_t ttable_word(int dec, int r, uint8_t x) {
    const uint8_t s = dec ? inv_sbox(x) : sbox(x);
    const uint32_t w = dec ? ((uint32_t)gmul(s, 14) << 24) | ((uint32_t)gmul(s, 9) << 16) |
                                 ((uint32_t)gmul(s, 13) << 8) | gmul(s, 11)
                           : ((uint32_t)gmul(s, 2) << 24) | ((uint32_t)s << 16) | ((uint32_t)s << 8) | gmul(s, 3);
    return r ? (w >> (8 * r)) | (w << (32 - 8 * r)) : w;
}
// GF(2)-linear byte map -> source bit lists [128][nsrc] (-1 padded), bit = 8 byte + j
template <class F>
int linear_sources(F map, int nsrc, std::vector<int16_t>& src) {
    src.assign((size_t)128 * nsrc, -1);
    std::vector<int> cnt(128, 0);
    for (int ib = 0; ib < 128; ++ib) {
        uint8_t in[16] = {0}, o[16];
        in[ib / 8] = (uint8_t)(1u << (ib % 8));
        map(in, o);
        for (int ob = 0; ob < 128; ++ob)
            if ((o[ob / 8] >> (ob % 8)) & 1) {
                if (cnt[ob] >= nsrc) return fail(TFHE_E_PARAM, "linear layer: too many sources");
                src[(size_t)ob * nsrc + cnt[ob]++] = (int16_t)ib;
            }
    }
    return TFHE_OK;
}

// Per-call scratch of one group of AES blocks: the context's main set, or the small
// set that the remainder group of an uneven batch uses on the second stream while the
// full-round group runs on the caller's stream.
struct CbsScratch {
    uint64_t* lwe64 = nullptr;  // [cap][2][N+1] PBS outputs
    uint32_t* at = nullptr;     // lookup form: [N+1][pfks_bpad(2 cap) + 256] PFKS inputs, rounded and transposed
    int8_t* pfkl_D = nullptr;   // linear PFKS or key switch: digits of min(2 cap, kPfklChunk) inputs
    uint32_t* ks_at = nullptr;  // [N][ks_bpad(cap)] key-switch digits
    uint32_t* rgsw = nullptr;   // [cap][4][2][N]
    cplx* rgsw_f = nullptr;     // [cap][8][1024]
    uint32_t* lut32 = nullptr;  // [cap][N+1]
    uint32_t* lutw = nullptr;   // T-table rounds: [cap / 8 bytes][32][N+1] word bits, made on first use
    uint32_t* state[2] = {nullptr, nullptr};  // [cap][n+1]
    size_t cap = 0;
};

// the T-table rounds' LUT output buffer of a scratch set (allocated on first use)
int ensure_lutw(CbsScratch& s) {
    return s.lutw ? TFHE_OK : dmalloc((void**)&s.lutw, s.cap / 8 * 32 * (kN + 1) * sizeof(uint32_t));
}

// allocates a scratch set for cap circuit bootstraps
// conv: the context's LWE -> RGSW form (0 lookup PFKS, 1 linear PFKS, 2 HomTrace + SchemeSwitch);
// ks_linear: the N -> n key switch runs as a GEMM (it shares the linear PFKS's digit scratch)
int alloc_scratch(CbsScratch& s, size_t cap, int n, int conv, int ks_linear) {
    int rc = TFHE_OK;
    s.cap = cap;
    const bool linear = conv == 1 || ks_linear;
    rc = rc ? rc : dmalloc((void**)&s.lwe64, cap * 2 * (kN + 1) * sizeof(uint64_t));
    // + 256 columns: a batch split into full rounds and a tail pads each part separately
    if (conv == 0)
        rc = rc ? rc
                : dmalloc((void**)&s.at, (size_t)(kN + 1) * (pfks_bpad(2 * (uint32_t)cap) + 256) * sizeof(uint32_t));
    rc = rc ? rc : dmalloc((void**)&s.ks_at, (size_t)kN * ks_bpad((uint32_t)cap) * sizeof(uint32_t));
    rc = rc ? rc : dmalloc((void**)&s.rgsw, cap * 4 * 2 * kN * sizeof(uint32_t));
    rc = rc ? rc : dmalloc((void**)&s.rgsw_f, cap * 8 * 1024 * sizeof(cplx));
    rc = rc ? rc : dmalloc((void**)&s.lut32, cap * (kN + 1) * sizeof(uint32_t));
    rc = rc ? rc : dmalloc((void**)&s.state[0], cap * (n + 1) * sizeof(uint32_t));
    rc = rc ? rc : dmalloc((void**)&s.state[1], cap * (n + 1) * sizeof(uint32_t));
    if (linear)
        rc = rc ? rc : dmalloc((void**)&s.pfkl_D, pfkl_digit_bytes(std::min<uint32_t>(2 * (uint32_t)cap, kPfklChunk)));
    return rc;
}

void free_scratch(CbsScratch& s) {
    void* bufs[] = {s.lwe64, s.at, s.pfkl_D, s.ks_at, s.rgsw, s.rgsw_f, s.lut32, s.lutw, s.state[0], s.state[1]};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    s = CbsScratch{};
}
}  // namespace

struct tfhe_cbs_ctx {
    tfhe_cbs_params p{};
    uint32_t max_batch = 0;  // LWE inputs per circuit-bootstrapping call
    bool has_keys = false;
    std::vector<int32_t> lwe_sk, K2;
    uint64_t seed = 0;
    int32_t* d_K2 = nullptr;
    cplx* d_bsk2f = nullptr;
    uint32_t* d_pfks = nullptr;  // [2][N+1][8][15][2N] torus-32 (lookup form)
    int8_t* d_pfkl = nullptr;    // [2][5][2N][pfkl_kp()] signed key bytes (linear form)
    uint32_t* d_ksk2 = nullptr;  // key-switching key words (the form of p.ks_linear)
    int8_t* d_kskl = nullptr;    // linear key switch: the key as signed byte slices in fragment order
    cplx* d_htkf = nullptr;       // HomTrace key [11][5][2][2 limbs][1024] (lwe_to_rgsw = 1)
    cplx* d_sskf = nullptr;       // SchemeSwitch key [5][2][2 limbs][1024]
    uint64_t* d_htss_lut = nullptr;  // [N] test polynomial of the HomTrace form's PBS
    CbsScratch sc;                // per-call scratch for max_batch circuit bootstraps
    // scratch of the remainder group of an uneven AES batch (tail_blocks blocks at most),
    // run on cstream beside the full-round group
    CbsScratch tail;
    uint32_t tail_blocks = 0;
    hipEvent_t ev_gfork = nullptr, ev_gjoin = nullptr;
    DeviceTables1024 tables;
    // AES state
    uint32_t max_blocks = 0;
    uint32_t* d_sbox_poly = nullptr;            // [N]
    uint32_t* d_inv_sbox_poly = nullptr;        // [N]
    uint32_t* d_ttab = nullptr;                 // [2 (Te, Td)][4 rows][4 output rows][N] T-table polynomials
    int16_t* d_lin_mix = nullptr;               // [128][kLinSrc]
    int16_t* d_lin_last = nullptr;
    int16_t* d_lin_inv_mix = nullptr;           // [128][kLinSrcDec] InvShiftRows + InvMixColumns
    int16_t* d_lin_inv_last = nullptr;          // [128][kLinSrcDec] InvShiftRows
    int16_t* d_lin_id = nullptr;                // [128] identity (the round key addition of T-table rounds)
    uint32_t* d_rk = nullptr;                   // [11][128][n+1] encrypted round-key bits
    uint32_t* d_dk = nullptr;                   // [11][128][n+1] encrypted decryption round-key bits
    bool has_dec_round_keys = false;
    // AES round form: 0 = S-box LUT then ShiftRows / MixColumns / AddRoundKey as LWE
    // additions; 1 = T-table rounds (rounds 1..9: four 8 -> 32 LUTs per column XORed at
    // dimension N, one key switch per bit, the round key added after it), as the reference
    int round_form = 0;
    uint8_t* d_ctr = nullptr;                   // [max_blocks][16]
    uint8_t* d_pub = nullptr;                   // [max_blocks][16] public AES-CTR ciphertext bytes
    bool has_round_keys = false;
    // captured CTR keystreams (aes_init + 10 rounds), one per block count (at most kMaxGraphs,
    // least recently used evicted), writing the context's state buffers,
    // replayed on the context's own stream, ordered after / before the caller's stream by events
    struct AesGraph {
        uint32_t blocks;
        hipGraphExec_t exec;
    };
    std::vector<AesGraph> graphs;
    hipStream_t gstream = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    // second stream of tfhe_cbs_batch: the PBS of a partial last round runs there while the
    // PFKS of the full rounds' outputs fills the idle CUs (fork / join by events)
    hipStream_t cstream = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // launch shapes chosen by tfhe_cbs_autotune, keyed by the launch's size (absent: the
    // kernels' own models): PBS inputs -> FFT form, PFKS LWEs -> i-slices, key switches -> i-slices
    std::map<uint32_t, int> tune_pbs, tune_pfks, tune_ks;
    // kernel timing (tfhe_cbs_timing_enable / _read): an event pair around every PBS launch
    // made outside stream capture while enabled; the pairs are reused after each read
    bool timing = false;
    std::vector<hipEvent_t> tev;  // [2 k] before, [2 k + 1] after launch k
    size_t tev_used = 0;          // events recorded since the last read
    uint64_t t_inputs = 0;        // PBS inputs of the timed launches
};

extern "C" int tfhe_cbs_ctx_destroy(tfhe_cbs_ctx* c) {
    if (!c) return TFHE_OK;
    void* bufs[] = {c->d_K2,          c->d_bsk2f,       c->d_pfks,          c->d_ksk2,           c->d_sbox_poly,
                    c->d_inv_sbox_poly, c->d_ttab,     c->d_lin_mix,        c->d_lin_last,       c->d_lin_inv_mix,
                    c->d_lin_inv_last,  c->d_lin_id,   c->d_rk,             c->d_dk,             c->d_ctr,
                    c->d_pub,           c->d_pfkl,     c->d_htkf,           c->d_sskf,           c->d_htss_lut,
                    c->d_kskl};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    free_scratch(c->sc);
    free_scratch(c->tail);
    if (c->ev_gfork) (void)hipEventDestroy(c->ev_gfork);
    if (c->ev_gjoin) (void)hipEventDestroy(c->ev_gjoin);
    for (auto& g : c->graphs) (void)hipGraphExecDestroy(g.exec);
    if (c->ev_in) (void)hipEventDestroy(c->ev_in);
    if (c->ev_out) (void)hipEventDestroy(c->ev_out);
    if (c->gstream) (void)hipStreamDestroy(c->gstream);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->cstream) (void)hipStreamDestroy(c->cstream);
    for (hipEvent_t e : c->tev) (void)hipEventDestroy(e);
    c->tables.release();
    delete c;
    return TFHE_OK;
}

extern "C" int tfhe_cbs_ctx_create(const tfhe_cbs_params* params, uint32_t max_blocks, tfhe_cbs_ctx** out) {
    if (!out) return fail(TFHE_E_PARAM, "tfhe_cbs_ctx_create: out is NULL");
    *out = nullptr;
    if (!cbs_params_supported(params))
        return fail(TFHE_E_PARAM, "tfhe_cbs_ctx_create: only N=2048, PBS 2^9x4, CB 2^8x2, PFKS 2^4x8, KS 2^2x8 is "
                                  "built (two key bits per step: one-product form and even n)");
    if (max_blocks == 0) max_blocks = 1;
    if (max_blocks > kMaxBlocks)
        return fail(TFHE_E_PARAM, "tfhe_cbs_ctx_create: at most 4096 AES blocks (524288 circuit bootstraps) per context");
    auto* c = new tfhe_cbs_ctx();
    c->p = *params;
    c->max_blocks = max_blocks;
    c->max_batch = 128 * max_blocks;
    const auto& p = c->p;
    const size_t B = c->max_batch;
    int rc = TFHE_OK;
    rc = rc ? rc : dmalloc((void**)&c->d_K2, kN * sizeof(int32_t));
    rc = rc ? rc : dmalloc((void**)&c->d_bsk2f, bsk2_cplx(p) * sizeof(cplx));
    rc = rc ? rc : dmalloc((void**)&c->d_ksk2, ksk2_words(p) * sizeof(uint32_t));
    if (p.ks_linear) rc = rc ? rc : dmalloc((void**)&c->d_kskl, kskl_key_bytes());
    const int conv = htss(p) ? 2 : p.pfks_linear ? 1 : 0;
    rc = rc ? rc : alloc_scratch(c->sc, B, p.n, conv, p.ks_linear);
    {
        // uneven AES batches of a linear-PFKS context: blocks beyond the last full round of
        // resident PBS workgroups (2 x CUs inputs, rb blocks) run as their own group on the
        // second stream
        const uint32_t slots = 2 * device_cus(), rb = slots / 128;
        if (!rc && conv != 0 && slots % 128 == 0 && rb >= 2 && max_blocks > rb) {
            c->tail_blocks = rb - 1;
            rc = alloc_scratch(c->tail, 128 * (size_t)c->tail_blocks, p.n, conv, p.ks_linear);
        }
        // the second stream and its events exist before any call, so that no call made
        // under stream capture creates them
        if (!rc) rc = hipcheck(hipStreamCreateWithFlags(&c->cstream, hipStreamNonBlocking), "cbs stream");
        if (!rc) rc = hipcheck(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming), "cbs event");
        if (!rc) rc = hipcheck(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming), "cbs event");
        if (!rc) rc = hipcheck(hipEventCreateWithFlags(&c->ev_gfork, hipEventDisableTiming), "cbs event");
        if (!rc) rc = hipcheck(hipEventCreateWithFlags(&c->ev_gjoin, hipEventDisableTiming), "cbs event");
    }
    rc = rc ? rc : dmalloc((void**)&c->d_sbox_poly, kN * sizeof(uint32_t));
    rc = rc ? rc : dmalloc((void**)&c->d_lin_mix, 128 * kLinSrc * sizeof(int16_t));
    rc = rc ? rc : dmalloc((void**)&c->d_lin_last, 128 * kLinSrc * sizeof(int16_t));
    rc = rc ? rc : dmalloc((void**)&c->d_rk, (size_t)11 * 128 * (p.n + 1) * sizeof(uint32_t));
    rc = rc ? rc : dmalloc((void**)&c->d_dk, (size_t)11 * 128 * (p.n + 1) * sizeof(uint32_t));
    rc = rc ? rc : dmalloc((void**)&c->d_inv_sbox_poly, kN * sizeof(uint32_t));
    rc = rc ? rc : dmalloc((void**)&c->d_ttab, (size_t)2 * 4 * 4 * kN * sizeof(uint32_t));
    rc = rc ? rc : dmalloc((void**)&c->d_lin_inv_mix, 128 * kLinSrcDec * sizeof(int16_t));
    rc = rc ? rc : dmalloc((void**)&c->d_lin_inv_last, 128 * kLinSrcDec * sizeof(int16_t));
    rc = rc ? rc : dmalloc((void**)&c->d_lin_id, 128 * sizeof(int16_t));
    if (!rc && htss(p)) {
        // LUT[k] = -(Gpow_(k mod 2) / 2) / N: after the trace, coefficient 0 of level j carries
        // bit Gpow_j (Gpow_j = 2^(48 + 8 j) on the 64-bit torus)
        std::vector<uint64_t> lut(kN);
        for (int k = 0; k < kN; ++k) lut[k] = 0 - (1ull << (47 + 8 * (k & 1) - 11));
        rc = dmalloc((void**)&c->d_htss_lut, kN * sizeof(uint64_t));
        if (!rc)
            rc = hipcheck(hipMemcpy(c->d_htss_lut, lut.data(), kN * sizeof(uint64_t), hipMemcpyHostToDevice), "htss lut");
    }
    rc = rc ? rc : dmalloc((void**)&c->d_ctr, (size_t)max_blocks * 16);
    rc = rc ? rc : dmalloc((void**)&c->d_pub, (size_t)max_blocks * 16);
    rc = rc ? rc : c->tables.init();
    if (!rc) {
        // S-box table packed for vertical packing: coefficient 8x + j = bit j of S(x), times 2^31
        std::vector<uint32_t> poly(kN, 0);
        for (int x = 0; x < 256; ++x)
            for (int j = 0; j < 8; ++j) poly[8 * x + j] = (uint32_t)((sbox((uint8_t)x) >> j) & 1) << 31;
        rc = hipcheck(hipMemcpy(c->d_sbox_poly, poly.data(), kN * sizeof(uint32_t), hipMemcpyHostToDevice), "sbox");
        for (int x = 0; !rc && x < 256; ++x)
            for (int j = 0; j < 8; ++j) poly[8 * x + j] = (uint32_t)((inv_sbox((uint8_t)x) >> j) & 1) << 31;
        if (!rc)
            rc = hipcheck(hipMemcpy(c->d_inv_sbox_poly, poly.data(), kN * sizeof(uint32_t), hipMemcpyHostToDevice),
                          "inverse sbox");
        // T-table polynomials: (dir, row r, output row p) coefficient 8x + j = bit j of byte p
        // of T_r[x], times 2^31
        std::vector<uint32_t> tt((size_t)2 * 4 * 4 * kN, 0);
        for (int d = 0; d < 2; ++d)
            for (int r = 0; r < 4; ++r)
                for (int x = 0; x < 256; ++x) {
                    const uint32_t w = ttable_word(d, r, (uint8_t)x);
                    for (int pr = 0; pr < 4; ++pr)
                        for (int j = 0; j < 8; ++j)
                            tt[(((size_t)d * 4 + r) * 4 + pr) * kN + 8 * x + j] = ((w >> (24 - 8 * pr + j)) & 1u) << 31;
                }
        if (!rc) rc = hipcheck(hipMemcpy(c->d_ttab, tt.data(), tt.size() * sizeof(uint32_t), hipMemcpyHostToDevice), "ttab");
        std::vector<int16_t> src;
        if (!rc) rc = linear_sources([](const uint8_t* i, uint8_t* o) { inv_shift_mix(i, o, true); }, kLinSrcDec, src);
        if (!rc)
            rc = hipcheck(hipMemcpy(c->d_lin_inv_mix, src.data(), src.size() * sizeof(int16_t), hipMemcpyHostToDevice),
                          "inverse linear tables");
        if (!rc) rc = linear_sources([](const uint8_t* i, uint8_t* o) { inv_shift_mix(i, o, false); }, kLinSrcDec, src);
        if (!rc)
            rc = hipcheck(hipMemcpy(c->d_lin_inv_last, src.data(), src.size() * sizeof(int16_t), hipMemcpyHostToDevice),
                          "inverse linear tables");
        if (!rc) rc = linear_sources([](const uint8_t* i, uint8_t* o) { std::memcpy(o, i, 16); }, 1, src);
        if (!rc)
            rc = hipcheck(hipMemcpy(c->d_lin_id, src.data(), src.size() * sizeof(int16_t), hipMemcpyHostToDevice),
                          "identity table");
    }
    if (!rc) {
        // source bit lists of the GF(2)-linear ShiftRows(+MixColumns) map, bit = 8*byte + j
        for (int mix = 0; mix < 2 && !rc; ++mix) {
            std::vector<int16_t> src(128 * kLinSrc, -1);
            std::vector<int> cnt(128, 0);
            for (int ib = 0; ib < 128; ++ib) {
                uint8_t in[16] = {0}, o[16];
                in[ib / 8] = (uint8_t)(1u << (ib % 8));
                shift_mix(in, o, mix == 1);
                for (int ob = 0; ob < 128; ++ob)
                    if ((o[ob / 8] >> (ob % 8)) & 1) {
                        if (cnt[ob] >= kLinSrc) { rc = fail(TFHE_E_PARAM, "linear layer: too many sources"); break; }
                        src[ob * kLinSrc + cnt[ob]++] = (int16_t)ib;
                    }
            }
            if (!rc)
                rc = hipcheck(hipMemcpy(mix ? c->d_lin_mix : c->d_lin_last, src.data(), src.size() * sizeof(int16_t),
                                        hipMemcpyHostToDevice), "linear tables");
        }
    }
    if (rc) {
        tfhe_cbs_ctx_destroy(c);
        return rc;
    }
    *out = c;
    return TFHE_OK;
}

// Fourier transform of the RLWE secret (host, N int32) into a new device buffer
// (2 limbs, k_to_fourier2048 layout), as the key generators take it
static int k2_fourier(const int32_t* h_K2, const FFT1024Tables& T, cplx** d_K2f, hipStream_t st) {
    uint64_t* d_K64 = nullptr;
    *d_K2f = nullptr;
    int rc = dmalloc((void**)d_K2f, 2 * 1024 * sizeof(cplx));
    rc = rc ? rc : dmalloc((void**)&d_K64, kN * sizeof(uint64_t));
    std::vector<uint64_t> hK64(kN);
    for (int i = 0; i < kN; ++i) hK64[i] = (uint64_t)(int64_t)h_K2[i];
    if (!rc) rc = hipcheck(hipMemcpy(d_K64, hK64.data(), kN * sizeof(uint64_t), hipMemcpyHostToDevice), "K64");
    if (!rc) launch_to_fourier2048(d_K64, 1, T, *d_K2f, st);
    if (!rc) rc = check_launch("K2 fourier");
    if (!rc) rc = hipcheck(hipStreamSynchronize(st), "K2 fourier sync");
    (void)hipFree(d_K64);
    return rc;
}

static int stateless_tables(DeviceTables1024*& t) {
    return per_device_tables(t, [](DeviceTables1024& d) { return d.init(); });
}

extern "C" int tfhe_gen_bsk64(const tfhe_cbs_params* p, uint64_t seed, const int32_t* d_lwe_sk, const int32_t* d_rlwe_sk,
                              uint64_t* d_bsk, void* stream) {
    if (!cbs_params_supported(p) || !d_lwe_sk || !d_rlwe_sk || !d_bsk) return fail(TFHE_E_PARAM, "tfhe_gen_bsk64: bad argument");
    DeviceTables1024* tables = nullptr;
    int rc = stateless_tables(tables);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    std::vector<int32_t> hK(kN);
    rc = hipcheck(hipMemcpy(hK.data(), d_rlwe_sk, kN * sizeof(int32_t), hipMemcpyDeviceToHost), "K2");
    cplx* d_K2f = nullptr;
    if (!rc) rc = k2_fourier(hK.data(), tables->T, &d_K2f, st);
    if (!rc) {
        if (pbs_auto(*p)) {  // one-bit RGSWs, then the two-bit pattern RGSWs
            launch_gen_bsk64(seed, p->n, p->rlwe_nshift, d_lwe_sk, d_rlwe_sk, d_K2f, tables->T, d_bsk, st, 1);
            launch_gen_bsk64(seed, p->n, p->rlwe_nshift, d_lwe_sk, d_rlwe_sk, d_K2f, tables->T,
                             d_bsk + bsk2_mb2_first(*p) * 8 * 2 * kN, st, 2);
        } else {
            launch_gen_bsk64(seed, p->n, p->rlwe_nshift, d_lwe_sk, d_rlwe_sk, d_K2f, tables->T, d_bsk, st, pbs_mb(*p),
                             p->bk_logB);
        }
    }
    if (!rc) rc = check_launch("gen_bsk64");
    if (!rc) rc = hipcheck(hipStreamSynchronize(st), "gen_bsk64 sync");
    (void)hipFree(d_K2f);
    return rc;
}

extern "C" int tfhe_gen_pfks32(const tfhe_cbs_params* p, uint64_t seed, const int32_t* d_rlwe_sk, uint32_t* d_pfks,
                               void* stream) {
    if (!cbs_params_supported(p) || !d_rlwe_sk || !d_pfks) return fail(TFHE_E_PARAM, "tfhe_gen_pfks32: bad argument");
    DeviceTables1024* tables = nullptr;
    int rc = stateless_tables(tables);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    std::vector<int32_t> hK(kN);
    rc = hipcheck(hipMemcpy(hK.data(), d_rlwe_sk, kN * sizeof(int32_t), hipMemcpyDeviceToHost), "K2");
    cplx* d_K2f = nullptr;
    if (!rc) rc = k2_fourier(hK.data(), tables->T, &d_K2f, st);
    if (!rc) launch_gen_pfks(seed, p->pf_nshift, d_rlwe_sk, d_K2f, tables->T, d_pfks, st);
    if (!rc) rc = check_launch("gen_pfks32");
    if (!rc) rc = hipcheck(hipStreamSynchronize(st), "gen_pfks32 sync");
    (void)hipFree(d_K2f);
    return rc;
}

// the conversion key of the context's form (allocated once)
static int alloc_pfks_key(tfhe_cbs_ctx* c) {
    if (htss(c->p)) {
        int rc = c->d_htkf ? TFHE_OK : dmalloc((void**)&c->d_htkf, (size_t)kHtRounds * kKsL * 2 * 2 * 1024 * sizeof(cplx));
        return rc ? rc : c->d_sskf ? TFHE_OK : dmalloc((void**)&c->d_sskf, (size_t)kKsL * 2 * 2 * 1024 * sizeof(cplx));
    }
    if (c->p.pfks_linear) return c->d_pfkl ? TFHE_OK : dmalloc((void**)&c->d_pfkl, pfkl_key_bytes());
    return c->d_pfks ? TFHE_OK : dmalloc((void**)&c->d_pfks, 2 * pfks_words_per_u() * sizeof(uint32_t));
}

// HomTrace + SchemeSwitch keys in coefficient form, d_keys [12][5][2][N] u64: message
// polynomials -sigma_k(K2) (k < 11, sigma_k: X -> X^(N/2^k + 1)) and K2^2, each times the
// gadget rows 2^(24 + 8d) (launch_gen_ks64); Philox domains DOM_HTKEY / DOM_SSKEY
static int gen_htss_keys(const tfhe_cbs_params& p, uint64_t seed, const int32_t* hK2, const cplx* d_K2f,
                         const FFT1024Tables& T, uint64_t* d_keys, hipStream_t st) {
    std::vector<int64_t> msg((size_t)(kHtRounds + 1) * kN, 0);
    for (int k = 0; k < kHtRounds; ++k) {
        const uint32_t e = (uint32_t)((kN >> k) + 1);
        for (int i = 0; i < kN; ++i) {
            const uint32_t t = (uint32_t)(((uint64_t)i * e) & (2 * kN - 1));
            if (t < (uint32_t)kN) msg[(size_t)k * kN + t] = -(int64_t)hK2[i];
            else msg[(size_t)k * kN + t - kN] = (int64_t)hK2[i];
        }
    }
    int64_t* sq = msg.data() + (size_t)kHtRounds * kN;  // K2^2 mod X^N + 1
    for (int i = 0; i < kN; ++i) {
        if (!hK2[i]) continue;
        for (int j = 0; j < kN; ++j) {
            const int64_t v = (int64_t)hK2[i] * hK2[j];
            if (i + j < kN) sq[i + j] += v;
            else sq[i + j - kN] -= v;
        }
    }
    int64_t* d_msg = nullptr;
    int rc = dmalloc((void**)&d_msg, msg.size() * sizeof(int64_t));
    if (!rc) rc = hipcheck(hipMemcpy(d_msg, msg.data(), msg.size() * sizeof(int64_t), hipMemcpyHostToDevice), "htss msg");
    if (!rc) {
        launch_gen_ks64(seed, DOM_HTKEY, kHtRounds, p.rlwe_nshift, d_msg, d_K2f, T, d_keys, st);
        launch_gen_ks64(seed, DOM_SSKEY, 1, p.rlwe_nshift, d_msg + (size_t)kHtRounds * kN, d_K2f, T,
                        d_keys + (size_t)kHtRounds * kKsL * 2 * kN, st);
        rc = check_launch("gen htss keys");
    }
    if (!rc) rc = hipcheck(hipStreamSynchronize(st), "gen htss keys sync");
    (void)hipFree(d_msg);
    return rc;
}

// coefficient-form HomTrace + SchemeSwitch keys (device) -> the context's two-limb Fourier keys
static int htss_keys_to_fourier(tfhe_cbs_ctx* c, const uint64_t* d_keys, hipStream_t st) {
    launch_to_fourier2048(d_keys, (size_t)kHtRounds * kKsL * 2, c->tables.T, c->d_htkf, st);
    launch_to_fourier2048(d_keys + (size_t)kHtRounds * kKsL * 2 * kN, (size_t)kKsL * 2, c->tables.T, c->d_sskf, st);
    return check_launch("htss keys fourier");
}

extern "C" int tfhe_gen_htss_keys(const tfhe_cbs_params* p, uint64_t seed, const int32_t* d_rlwe_sk, uint64_t* d_keys,
                                  void* stream) {
    if (!cbs_params_supported(p) || !d_rlwe_sk || !d_keys) return fail(TFHE_E_PARAM, "tfhe_gen_htss_keys: bad argument");
    DeviceTables1024* tables = nullptr;
    int rc = stateless_tables(tables);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    std::vector<int32_t> hK(kN);
    rc = hipcheck(hipMemcpy(hK.data(), d_rlwe_sk, kN * sizeof(int32_t), hipMemcpyDeviceToHost), "K2");
    cplx* d_K2f = nullptr;
    if (!rc) rc = k2_fourier(hK.data(), tables->T, &d_K2f, st);
    if (!rc) rc = gen_htss_keys(*p, seed, hK.data(), d_K2f, tables->T, d_keys, st);
    (void)hipFree(d_K2f);
    return rc;
}

extern "C" int tfhe_gen_pfkl40(const tfhe_cbs_params* p, uint64_t seed, const int32_t* d_rlwe_sk, int8_t* d_pfkl,
                               void* stream) {
    if (!cbs_params_supported(p) || !d_rlwe_sk || !d_pfkl) return fail(TFHE_E_PARAM, "tfhe_gen_pfkl40: bad argument");
    hipStream_t st = (hipStream_t)stream;
    int rc = hipcheck(hipMemsetAsync(d_pfkl, 0, pfkl_key_bytes(), st), "pfkl zero");
    if (!rc) {
        launch_gen_pfkl40(seed, d_rlwe_sk, d_pfkl, st);
        rc = check_launch("gen_pfkl40");
    }
    if (!rc) rc = hipcheck(hipStreamSynchronize(st), "gen_pfkl40 sync");
    return rc;
}

// linear key switch: the key words d_ksk2 as signed byte slices in fragment order (padding zero)
static int pack_kskl(tfhe_cbs_ctx* c) {
    if (!c->p.ks_linear) return TFHE_OK;
    int rc = hipcheck(hipMemset(c->d_kskl, 0, kskl_key_bytes()), "kskl zero");
    if (!rc && launch_kskl_pack(c->d_ksk2, c->p.n, c->d_kskl, nullptr)) rc = fail(TFHE_E_PARAM, "kskl: n too large");
    return rc ? rc : check_launch("kskl pack");
}

extern "C" int tfhe_cbs_set_keys(tfhe_cbs_ctx* c, const int32_t* rlwe_sk, const uint64_t* bsk, const uint32_t* pfks,
                                 const uint32_t* ksk2) {
    if (!c || !rlwe_sk || !bsk || !pfks || !ksk2) return fail(TFHE_E_PARAM, "tfhe_cbs_set_keys: bad argument");
    const auto& p = c->p;
    int rc = alloc_pfks_key(c);
    uint64_t* d_bsk = nullptr;
    const size_t bsk_words = bsk2_rgsw(p) * 8 * 2 * kN;
    rc = rc ? rc : dmalloc((void**)&d_bsk, bsk_words * sizeof(uint64_t));
    c->has_keys = false;
    c->K2.assign(rlwe_sk, rlwe_sk + kN);
    c->lwe_sk.clear();
    // host or device sources: hipMemcpyDefault follows the pointers
    if (!rc) rc = hipcheck(hipMemcpy(c->d_K2, c->K2.data(), kN * sizeof(int32_t), hipMemcpyHostToDevice), "K2");
    if (!rc) rc = hipcheck(hipMemcpy(d_bsk, bsk, bsk_words * sizeof(uint64_t), hipMemcpyDefault), "bsk2");
    if (!rc && htss(p)) {  // pfks: the HomTrace + SchemeSwitch keys as tfhe_gen_htss_keys writes them
        uint64_t* d_keys = nullptr;
        rc = dmalloc((void**)&d_keys, htss_key_words() * sizeof(uint64_t));
        if (!rc) rc = hipcheck(hipMemcpy(d_keys, pfks, htss_key_words() * sizeof(uint64_t), hipMemcpyDefault), "htss keys");
        if (!rc) rc = htss_keys_to_fourier(c, d_keys, nullptr);
        if (!rc) rc = hipcheck(hipDeviceSynchronize(), "htss keys sync");
        (void)hipFree(d_keys);
    } else if (!rc) {
        rc = p.pfks_linear ? hipcheck(hipMemcpy(c->d_pfkl, pfks, pfkl_key_bytes(), hipMemcpyDefault), "pfkl")
                           : hipcheck(hipMemcpy(c->d_pfks, pfks, 2 * pfks_words_per_u() * sizeof(uint32_t),
                                                hipMemcpyDefault), "pfks");
    }
    if (!rc) rc = hipcheck(hipMemcpy(c->d_ksk2, ksk2, ksk2_words(p) * sizeof(uint32_t), hipMemcpyDefault), "ksk2");
    if (!rc) rc = pack_kskl(c);
    if (!rc) {
        if (pbs_limbs(p) == 2)
            launch_to_fourier2048(d_bsk, bsk2_rgsw(p) * 8 * 2, c->tables.T, c->d_bsk2f, nullptr);
        else
            launch_to_fourier2048_t64(d_bsk, bsk2_rgsw(p) * 8 * 2, c->tables.T, c->d_bsk2f, nullptr);
        rc = check_launch("set_keys fourier");
    }
    if (!rc) rc = hipcheck(hipDeviceSynchronize(), "set_keys sync");
    (void)hipFree(d_bsk);
    if (!rc) c->has_keys = true;
    return rc;
}

extern "C" int tfhe_cbs_keygen(tfhe_cbs_ctx* c, const int32_t* lwe_sk, uint64_t seed) {
    if (!c || !lwe_sk) return fail(TFHE_E_PARAM, "tfhe_cbs_keygen: bad argument");
    const auto& p = c->p;
    c->seed = seed;
    c->lwe_sk.assign(lwe_sk, lwe_sk + p.n);
    c->K2.resize(kN);
    tfhe_rlwe_keygen(kN, seed, DOM_RLWE2_SK, c->K2.data());
    int rc = alloc_pfks_key(c);
    int32_t* d_s = nullptr;
    uint64_t *d_bsk = nullptr, *d_K64 = nullptr;
    cplx* d_K2f = nullptr;
    rc = rc ? rc : dmalloc((void**)&d_s, p.n * sizeof(int32_t));
    rc = rc ? rc : dmalloc((void**)&d_bsk, bsk2_rgsw(p) * 8 * 2 * kN * sizeof(uint64_t));
    rc = rc ? rc : dmalloc((void**)&d_K64, kN * sizeof(uint64_t));
    rc = rc ? rc : dmalloc((void**)&d_K2f, 2 * 1024 * sizeof(cplx));
    std::vector<uint64_t> hK64(kN);
    for (int i = 0; i < kN; ++i) hK64[i] = (uint64_t)(int64_t)c->K2[i];
    if (!rc) rc = hipcheck(hipMemcpy(d_s, c->lwe_sk.data(), p.n * sizeof(int32_t), hipMemcpyHostToDevice), "sk");
    if (!rc) rc = hipcheck(hipMemcpy(c->d_K2, c->K2.data(), kN * sizeof(int32_t), hipMemcpyHostToDevice), "K2");
    if (!rc) rc = hipcheck(hipMemcpy(d_K64, hK64.data(), kN * sizeof(uint64_t), hipMemcpyHostToDevice), "K64");
    if (!rc) {
        launch_to_fourier2048(d_K64, 1, c->tables.T, d_K2f, nullptr);
        if (pbs_auto(p)) {
            launch_gen_bsk64(seed, p.n, p.rlwe_nshift, d_s, c->d_K2, d_K2f, c->tables.T, d_bsk, nullptr, 1);
            launch_gen_bsk64(seed, p.n, p.rlwe_nshift, d_s, c->d_K2, d_K2f, c->tables.T,
                             d_bsk + bsk2_mb2_first(p) * 8 * 2 * kN, nullptr, 2);
        } else {
            launch_gen_bsk64(seed, p.n, p.rlwe_nshift, d_s, c->d_K2, d_K2f, c->tables.T, d_bsk, nullptr, pbs_mb(p),
                             p.bk_logB);
        }
        if (pbs_limbs(p) == 2)
            launch_to_fourier2048(d_bsk, bsk2_rgsw(p) * 8 * 2, c->tables.T, c->d_bsk2f, nullptr);
        else
            launch_to_fourier2048_t64(d_bsk, bsk2_rgsw(p) * 8 * 2, c->tables.T, c->d_bsk2f, nullptr);
        if (htss(p)) {
            uint64_t* d_keys = nullptr;
            rc = dmalloc((void**)&d_keys, htss_key_words() * sizeof(uint64_t));
            if (!rc) rc = gen_htss_keys(p, seed, c->K2.data(), d_K2f, c->tables.T, d_keys, nullptr);
            if (!rc) rc = htss_keys_to_fourier(c, d_keys, nullptr);
            if (!rc) rc = hipcheck(hipDeviceSynchronize(), "htss keys sync");
            (void)hipFree(d_keys);
        } else if (p.pfks_linear) {
            rc = hipcheck(hipMemset(c->d_pfkl, 0, pfkl_key_bytes()), "pfkl zero");
            launch_gen_pfkl40(seed, c->d_K2, c->d_pfkl, nullptr);
        } else {
            launch_gen_pfks(seed, p.pf_nshift, c->d_K2, d_K2f, c->tables.T, c->d_pfks, nullptr);
        }
        launch_gen_ksk32(seed, p.n, kN, p.ks_logB, p.ks_l, p.lwe_nshift, d_s, c->d_K2, c->d_ksk2, nullptr,
                         p.ks_linear ? DOM_KSKL : DOM_KSK2, p.ks_linear);
        if (!rc) rc = check_launch("cbs keygen");
        if (!rc) rc = pack_kskl(c);
    }
    if (!rc) rc = hipcheck(hipDeviceSynchronize(), "cbs keygen sync");
    (void)hipFree(d_s);
    (void)hipFree(d_bsk);
    (void)hipFree(d_K64);
    (void)hipFree(d_K2f);
    if (!rc) c->has_keys = true;
    return rc;
}

const tfhe_cbs_params* tfhe::cbs_ctx_params(const tfhe_cbs_ctx* c) { return &c->p; }

extern "C" int tfhe_cbs_save_keys(const tfhe_cbs_ctx* c, const char* path) {
    if (!c || !path) return fail(TFHE_E_PARAM, "tfhe_cbs_save_keys: bad argument");
    if (!c->has_keys || c->lwe_sk.empty())
        return fail(TFHE_E_STATE, "tfhe_cbs_save_keys: the context's keys were not made by tfhe_cbs_keygen "
                                  "(uploaded keys are written from their sources with tfhe_cbs_write_keys)");
    // the context holds the BSK2 in its Fourier form: its coefficient rows are made again
    // from the seed; the PFKS key and the KSK2 are written as the context holds them
    const auto& p = c->p;
    int32_t* d_s = nullptr;
    uint64_t* d_bsk = nullptr;
    int rc = dmalloc((void**)&d_s, p.n * sizeof(int32_t));
    rc = rc ? rc : dmalloc((void**)&d_bsk, bsk2_words(p) * sizeof(uint64_t));
    if (!rc) rc = hipcheck(hipMemcpy(d_s, c->lwe_sk.data(), p.n * sizeof(int32_t), hipMemcpyHostToDevice), "sk");
    if (!rc) rc = tfhe_gen_bsk64(&p, c->seed, d_s, c->d_K2, d_bsk, nullptr);
    uint64_t* d_keys = nullptr;  // HomTrace + SchemeSwitch keys, also held in Fourier form only
    if (!rc && htss(p)) {
        rc = dmalloc((void**)&d_keys, htss_key_words() * sizeof(uint64_t));
        if (!rc) rc = tfhe_gen_htss_keys(&p, c->seed, c->d_K2, d_keys, nullptr);
    }
    if (!rc) {
        const void* conv = htss(p) ? (const void*)d_keys : p.pfks_linear ? (const void*)c->d_pfkl : (const void*)c->d_pfks;
        const void* const src[4] = {c->K2.data(), d_bsk, conv, c->d_ksk2};
        rc = write_key_blob(path, p, c->seed, src);
    }
    (void)hipFree(d_s);
    (void)hipFree(d_bsk);
    if (d_keys) (void)hipFree(d_keys);
    return rc;
}

extern "C" int tfhe_cbs_get_rlwe_sk(const tfhe_cbs_ctx* c, int32_t* out) {
    if (!c || !out || c->K2.empty()) return fail(TFHE_E_STATE, "tfhe_cbs_get_rlwe_sk: no keys");
    std::memcpy(out, c->K2.data(), kN * sizeof(int32_t));
    return TFHE_OK;
}

extern "C" int tfhe_cbs_get_pfkl_entry(const tfhe_cbs_ctx* c, int u, int i, int t, uint64_t* out) {
    if (!c || !out || !c->has_keys || !c->d_pfkl) return fail(TFHE_E_STATE, "tfhe_cbs_get_pfkl_entry: no linear key");
    if (u < 0 || u > 1 || i < 0 || i > kN || t < 0 || t > 7)
        return fail(TFHE_E_PARAM, "tfhe_cbs_get_pfkl_entry: index out of range");
    // the entry's 2N columns of each signed byte slice (one byte per 32-column fragment
    // row: columns 32 apart are one fragment block, pfkl_kp() x 32 bytes, apart), then the
    // words mod 2^40
    const size_t cols = 2 * kN, kp = pfkl_kp();
    const uint32_t k = 8u * (uint32_t)i + (uint32_t)t;
    std::vector<int8_t> sl((size_t)5 * cols);
    int rc = TFHE_OK;
    for (int b = 0; b < 5 && !rc; ++b)
        for (int cl = 0; cl < 32 && !rc; ++cl)
            rc = hipcheck(hipMemcpy2D(sl.data() + (size_t)b * cols + cl, 32, c->d_pfkl + pfkl_key_offset(u, b, cl, k),
                                      kp * 32, 1, cols / 32, hipMemcpyDeviceToHost), "pfkl entry");
    for (size_t col = 0; col < cols && !rc; ++col) {
        int64_t v = 0;
        for (int b = 0; b < 5; ++b) v += (int64_t)sl[(size_t)b * cols + col] * ((int64_t)1 << (8 * b));
        out[col] = (uint64_t)v & ((1ull << 40) - 1);
    }
    return rc;
}

extern "C" int tfhe_cbs_get_pfks_entry(const tfhe_cbs_ctx* c, int u, int i, int t, int v, uint32_t* out) {
    if (!c || !out || !c->has_keys || !c->d_pfks) return fail(TFHE_E_STATE, "tfhe_cbs_get_pfks_entry: no keys");
    if (u < 0 || u > 1 || i < 0 || i > kN || t < 0 || t > 7 || v < 1 || v > 15)
        return fail(TFHE_E_PARAM, "tfhe_cbs_get_pfks_entry: index out of range");
    const size_t row = (((size_t)u * (kN + 1) + i) * 8 + t) * 15 + (v - 1);
    return hipcheck(hipMemcpy(out, c->d_pfks + row * 2 * kN, 2 * kN * sizeof(uint32_t), hipMemcpyDeviceToHost),
                    "pfks entry");
}

static int tuned(const std::map<uint32_t, int>& m, uint32_t key, int none) {
    const auto it = m.find(key);
    return it == m.end() ? none : it->second;
}

static int check_cbs(const tfhe_cbs_ctx* c, uint32_t batch, uint32_t cap) {
    if (!c) return fail(TFHE_E_PARAM, "cbs call: ctx is NULL");
    if (!c->has_keys) return fail(TFHE_E_STATE, "cbs call: no keys (call tfhe_cbs_keygen)");
    if (batch > cap) return fail(TFHE_E_CAPACITY, "cbs call: batch exceeds the context capacity");
    return TFHE_OK;
}

// pbs_multibit = 3 without a tuned choice: two key bits per step at every size (with the key
// loads batched it measured faster than one bit at 128, 2048 and 8192 inputs: 5.9 / 34.4 /
// 136 ms against 8.1 / 44.4 / 178), split over 8 waves (4) for launches of at most one
// workgroup per CU, else at two waves per SIMD (2); 3 is the wide key batch.  Forms 0 and 1
// are one bit per step with that FFT form.
static int auto_pbs_model(uint32_t batch) { return 2 + pbs_model_form(batch, 1, 2); }

// The HomTrace form's PBS: the accumulator (0, LUT X^-b~) with its test polynomial, b + 2^30
// switched with bw = 1, the 2^10 gadget, written whole (no offsets) into rlwe [batch][2][N]
static PbsLut htss_pbs_io(const tfhe_cbs_ctx* c, uint64_t* rlwe) {
    PbsLut io;
    io.lut = c->d_htss_lut;
    io.rlwe = rlwe;
    io.bw = 1;
    io.b_add = 1u << 30;
    io.offsets = 0;
    io.logB = c->p.bk_logB;
    return io;
}

// Event pair around one PBS launch while timing is on and the stream is not capturing:
// begin records the first event and returns the second (or null), end records it.
static int pbs_timing_begin(tfhe_cbs_ctx* c, hipStream_t st, hipEvent_t* e1) {
    *e1 = nullptr;
    if (!c->timing) return TFHE_OK;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    int rc = hipcheck(hipStreamIsCapturing(st, &cs), "pbs timing");
    if (rc || cs != hipStreamCaptureStatusNone) return rc;
    while (c->tev.size() < c->tev_used + 2) {
        hipEvent_t e = nullptr;
        rc = hipcheck(hipEventCreate(&e), "pbs timing event");
        if (rc) return rc;
        c->tev.push_back(e);
    }
    rc = hipcheck(hipEventRecord(c->tev[c->tev_used], st), "pbs timing");
    if (!rc) *e1 = c->tev[c->tev_used + 1];
    return rc;
}
static int pbs_timing_end(tfhe_cbs_ctx* c, hipStream_t st, hipEvent_t e1, uint32_t batch) {
    if (!e1) return TFHE_OK;
    int rc = hipcheck(hipEventRecord(e1, st), "pbs timing");
    c->tev_used += 2;
    c->t_inputs += batch;
    return rc;
}

extern "C" int tfhe_cbs_pbs_batch(tfhe_cbs_ctx* c, const uint32_t* d_lwe, uint32_t batch, uint64_t* d_out, void* stream) {
    int rc = check_cbs(c, batch, c ? c->max_batch : 0);
    if (rc) return rc;
    if (!d_lwe || !d_out) return fail(TFHE_E_PARAM, "pbs: NULL buffer");
    if (htss(c->p))
        return fail(TFHE_E_PARAM, "pbs: the HomTrace + SchemeSwitch form's PBS keeps whole accumulators "
                                  "(tfhe_cbs_bootstrap_many_lut gives them)");
    if (!batch) return TFHE_OK;
    int mb = pbs_mb(c->p), form = tuned(c->tune_pbs, batch, -1);
    if (pbs_auto(c->p)) {  // forms 2 and 3 = two key bits per step, else one bit with that FFT form
        if (form < 0) form = auto_pbs_model(batch);
        mb = form >= 2 ? 2 : 1;
        if (mb == 2) form -= 2;
    }
    const cplx* bsk = c->d_bsk2f + (mb == 2 ? bsk2_mb2_first(c->p) * 8 * 2 * 1024 : 0);
    hipStream_t st = (hipStream_t)stream;
    hipEvent_t e1 = nullptr;
    rc = pbs_timing_begin(c, st, &e1);
    if (rc) return rc;
    launch_pbs2(d_lwe, batch, c->p.n, bsk, pbs_limbs(c->p), mb, c->tables.T, d_out, st, form);
    rc = check_launch("pbs2");
    return rc ? rc : pbs_timing_end(c, st, e1, batch);
}

extern "C" int tfhe_cbs_bootstrap_many_lut(tfhe_cbs_ctx* c, const uint32_t* d_lwe, uint32_t batch,
                                           const uint64_t* d_lut, uint32_t bitwidth, uint64_t* d_out, void* stream) {
    int rc = check_cbs(c, batch, c ? c->max_batch : 0);
    if (rc) return rc;
    if (!d_lwe || !d_lut || !d_out) return fail(TFHE_E_PARAM, "bootstrap_many_lut: NULL buffer");
    if (bitwidth > 11) return fail(TFHE_E_PARAM, "bootstrap_many_lut: bitwidth must be at most 11 (2N = 4096)");
    if (!batch) return TFHE_OK;
    int mb = pbs_mb(c->p), form = tuned(c->tune_pbs, batch, -1);
    if (pbs_auto(c->p)) {
        if (form < 0) form = auto_pbs_model(batch);
        mb = form >= 2 ? 2 : 1;
        if (mb == 2) form -= 2;
    }
    const cplx* bsk = c->d_bsk2f + (mb == 2 ? bsk2_mb2_first(c->p) * 8 * 2 * 1024 : 0);
    PbsLut io;
    io.lut = d_lut;
    io.rlwe = d_out;
    io.bw = bitwidth;
    io.logB = c->p.bk_logB;
    launch_pbs2(d_lwe, batch, c->p.n, bsk, pbs_limbs(c->p), mb, c->tables.T, nullptr, (hipStream_t)stream, form, io);
    return check_launch("bootstrap_many_lut");
}

extern "C" int tfhe_cbs_timing_enable(tfhe_cbs_ctx* c, int on) {
    if (!c) return fail(TFHE_E_PARAM, "timing: NULL context");
    c->timing = on != 0;
    c->tev_used = 0;
    c->t_inputs = 0;
    return TFHE_OK;
}

extern "C" int tfhe_cbs_timing_read(tfhe_cbs_ctx* c, double* pbs_ms, uint32_t* launches, uint64_t* inputs) {
    if (!c || !pbs_ms || !launches || !inputs) return fail(TFHE_E_PARAM, "timing: bad argument");
    double total = 0.0;
    for (size_t k = 0; k < c->tev_used; k += 2) {
        int rc = hipcheck(hipEventSynchronize(c->tev[k + 1]), "timing read");
        float ms = 0.0f;
        if (!rc) rc = hipcheck(hipEventElapsedTime(&ms, c->tev[k], c->tev[k + 1]), "timing read");
        if (rc) return rc;
        total += ms;
    }
    *pbs_ms = total;
    *launches = (uint32_t)(c->tev_used / 2);
    *inputs = c->t_inputs;
    c->tev_used = 0;
    c->t_inputs = 0;
    return TFHE_OK;
}

// linear-form PFKS in launches of at most kPfklChunk inputs (the digit scratch)
static void pfkl_chunks(tfhe_cbs_ctx* c, int8_t* D, const uint64_t* lwe64, uint32_t count, int u0, int nu,
                        uint32_t* out, size_t out_stride, size_t out_u_offset, hipStream_t st) {
    for (uint32_t c0 = 0; c0 < count; c0 += kPfklChunk) {
        const uint32_t cnt = std::min(kPfklChunk, count - c0);
        launch_pfkl(lwe64 + (size_t)c0 * (kN + 1), cnt, D, c->d_pfkl, u0, nu, out + (size_t)c0 * out_stride, out_stride,
                    out_u_offset, st);
    }
}

extern "C" int tfhe_cbs_pfks_batch(tfhe_cbs_ctx* c, int u, const uint64_t* d_lwe64, uint32_t count, size_t out_stride,
                                   uint32_t* d_out, void* stream) {
    int rc = check_cbs(c, count, c ? 2 * c->max_batch : 0);
    if (rc) return rc;
    if (u < 0 || u > 1 || !d_lwe64 || !d_out || out_stride < 2 * (size_t)kN)
        return fail(TFHE_E_PARAM, "pfks: bad argument");
    if (htss(c->p)) return fail(TFHE_E_PARAM, "pfks: the context converts by HomTrace + SchemeSwitch");
    if (!count) return TFHE_OK;
    hipStream_t st = (hipStream_t)stream;
    if (c->p.pfks_linear) {
        pfkl_chunks(c, c->sc.pfkl_D, d_lwe64, count, u, 1, d_out, out_stride, 0, st);
        return check_launch("pfks linear");
    }
    launch_pfks_digits(d_lwe64, count, c->sc.at, st);
    launch_zero_rows32(d_out, count, out_stride, 2 * kN, st);
    launch_pfks32(c->sc.at, count, c->d_pfks, pfks_words_per_u(), u, 1, d_out, out_stride, 0, st);
    return check_launch("pfks");
}

// The PBS runs in rounds of 2 x CUs resident workgroups: a batch with full rounds and a
// partial last one is split there (full, tail); otherwise full = 0.
static void cbs_split(uint32_t batch, uint32_t& full, uint32_t& tail) {
    const uint32_t slots = 2 * device_cus();
    full = batch / slots * slots;
    tail = batch - full;
    if (!full || !tail) full = 0, tail = batch;
}

// PFKS of count PBS outputs (LWE L = 2 input + j -> RGSW row 2j + u of that input,
// at L * 4N + u * 2N): both functions in one launch over the rounded, transposed inputs
static void cbs_pfks(tfhe_cbs_ctx* c, const uint64_t* lwe64, uint32_t count, uint32_t* at, int8_t* D, uint32_t* rows,
                     hipStream_t st) {
    if (c->p.pfks_linear) {
        pfkl_chunks(c, D, lwe64, count, 0, 2, rows, 4 * (size_t)kN, 2 * (size_t)kN, st);
        return;
    }
    launch_pfks_digits(lwe64, count, at, st);
    launch_pfks32(at, count, c->d_pfks, pfks_words_per_u(), 0, 2, rows, 4 * (size_t)kN, 2 * (size_t)kN, st,
                  tuned(c->tune_pfks, count, 0));
}

// circuit bootstraps with scratch set sc (checked arguments)
static int cbs_batch_impl(tfhe_cbs_ctx* c, const CbsScratch& sc, const uint32_t* d_lwe, uint32_t batch,
                          uint32_t* d_rgsw, hipStream_t st) {
    if (htss(c->p)) {  // exact 2^10 x 4 PBS accumulators, then HomTrace + SchemeSwitch per CB level
        hipEvent_t e1 = nullptr;
        int rc = pbs_timing_begin(c, st, &e1);
        if (rc) return rc;
        launch_pbs2(d_lwe, batch, c->p.n, c->d_bsk2f, 2, 1, c->tables.T, nullptr, st, tuned(c->tune_pbs, batch, -1),
                    htss_pbs_io(c, sc.lwe64));
        rc = check_launch("htss pbs");
        if (!rc) rc = pbs_timing_end(c, st, e1, batch);
        if (rc) return rc;
        launch_htss(sc.lwe64, batch, c->d_htkf, c->d_sskf, c->tables.T, d_rgsw, st);
        return check_launch("htss");
    }
    const size_t lwe64_words = 2 * (size_t)(kN + 1), rgsw_words = 4 * 2 * (size_t)kN;
    if (!c->p.pfks_linear) launch_zero_rows32(d_rgsw, 2 * batch, 4 * (size_t)kN, 4 * kN, st);  // lookup form adds
    // The PBS runs in rounds of as many workgroups as fit at once (two per CU).  A partial
    // last round leaves CUs idle, so it runs on the second stream while the full rounds'
    // PFKS starts on this one.
    uint32_t full = 0, tail = 0;
    cbs_split(batch, full, tail);
    int rc = TFHE_OK;
    if (full && tail) {
        rc = tfhe_cbs_pbs_batch(c, d_lwe, full, sc.lwe64, st);
        if (!rc) rc = hipcheck(hipEventRecord(c->ev_fork, st), "fork");
        if (!rc) rc = hipcheck(hipStreamWaitEvent(c->cstream, c->ev_fork, 0), "fork wait");
        if (!rc) rc = tfhe_cbs_pbs_batch(c, d_lwe + (size_t)full * (c->p.n + 1), tail, sc.lwe64 + full * lwe64_words,
                                         c->cstream);
        if (!rc) rc = hipcheck(hipEventRecord(c->ev_join, c->cstream), "join");
        if (rc) return rc;
        cbs_pfks(c, sc.lwe64, 2 * full, sc.at, sc.pfkl_D, d_rgsw, st);
        rc = hipcheck(hipStreamWaitEvent(st, c->ev_join, 0), "join wait");
        if (rc) return rc;
        uint32_t* at_tail = sc.at ? sc.at + (size_t)(kN + 1) * pfks_bpad(2 * full) : nullptr;  // lookup form only
        cbs_pfks(c, sc.lwe64 + full * lwe64_words, 2 * tail, at_tail, sc.pfkl_D, d_rgsw + full * rgsw_words, st);
    } else {
        rc = tfhe_cbs_pbs_batch(c, d_lwe, batch, sc.lwe64, st);
        if (rc) return rc;
        cbs_pfks(c, sc.lwe64, 2 * batch, sc.at, sc.pfkl_D, d_rgsw, st);
    }
    return check_launch("cbs pfks");
}

extern "C" int tfhe_cbs_batch(tfhe_cbs_ctx* c, const uint32_t* d_lwe, uint32_t batch, uint32_t* d_rgsw, void* stream) {
    int rc = check_cbs(c, batch, c ? c->max_batch : 0);
    if (rc) return rc;
    if (!d_lwe || !d_rgsw) return fail(TFHE_E_PARAM, "cbs: NULL buffer");
    if (!batch) return TFHE_OK;
    return cbs_batch_impl(c, c->sc, d_lwe, batch, d_rgsw, (hipStream_t)stream);
}

extern "C" int tfhe_cbs_batch_from_device_lwe(tfhe_cbs_ctx* c, const uint64_t* d_a, const uint64_t* d_b,
                                              uint32_t batch, uint32_t a_stride, tfhe_rgsw_batch_view* view,
                                              void* stream) {
    if (!c || !view) return fail(TFHE_E_PARAM, "cbs_from_device_lwe: bad argument");
    *view = tfhe_rgsw_batch_view{};
    if (!d_a || !d_b || batch == 0) return TFHE_OK;  // the reference returns an empty view
    if (a_stride < (uint32_t)c->p.n) return fail(TFHE_E_PARAM, "cbs_from_device_lwe: a_stride < n");
    int rc = check_cbs(c, batch, c->max_batch);
    if (rc) return rc;
    // packed inputs in the spare AES state buffer (context scratch)
    launch_pack_lwe64(d_a, d_b, batch, c->p.n, a_stride, c->sc.state[1], (hipStream_t)stream);
    rc = check_launch("pack_lwe64");
    if (!rc) rc = tfhe_cbs_batch(c, c->sc.state[1], batch, c->sc.rgsw, stream);
    if (rc) return rc;
    view->d_rgsw = c->sc.rgsw;
    view->batch = batch;
    view->digits = 2;
    view->N = kN;
    view->per_ciphertext_elems = (size_t)2 * 4 * kN;
    view->stream = stream;
    return TFHE_OK;
}

extern "C" int tfhe_cbs_batch_to_device(tfhe_cbs_ctx* c, const uint32_t* h_lwe, uint32_t batch,
                                        tfhe_rgsw_batch_view* view, void* stream) {
    if (!c || !view) return fail(TFHE_E_PARAM, "cbs_batch_to_device: bad argument");
    *view = tfhe_rgsw_batch_view{};
    if (!h_lwe || batch == 0) return TFHE_OK;  // the reference returns an empty view
    int rc = check_cbs(c, batch, c->max_batch);
    if (rc) return rc;
    // staged in the spare AES state buffer (context scratch), as the device-LWE form packs there
    rc = hipcheck(hipMemcpyAsync(c->sc.state[1], h_lwe, (size_t)batch * (c->p.n + 1) * sizeof(uint32_t),
                                 hipMemcpyHostToDevice, (hipStream_t)stream),
                  "cbs_batch_to_device upload");
    if (!rc) rc = tfhe_cbs_batch(c, c->sc.state[1], batch, c->sc.rgsw, stream);
    if (rc) return rc;
    view->d_rgsw = c->sc.rgsw;
    view->batch = batch;
    view->digits = 2;
    view->N = kN;
    view->per_ciphertext_elems = (size_t)2 * 4 * kN;
    view->stream = stream;
    return TFHE_OK;
}

extern "C" int tfhe_rgsw_to_fourier(tfhe_cbs_ctx* c, const uint32_t* d_rgsw, uint32_t count, double* d_rgsw_f,
                                    void* stream) {
    if (!c || !d_rgsw || !d_rgsw_f) return fail(TFHE_E_PARAM, "rgsw_to_fourier: bad argument");
    if (!count) return TFHE_OK;
    launch_to_fourier2048_32(d_rgsw, (size_t)count * 8, c->tables.T, reinterpret_cast<cplx*>(d_rgsw_f),
                             (hipStream_t)stream);
    return check_launch("to_fourier2048_32");
}

extern "C" int tfhe_lut8_eval_batch(tfhe_cbs_ctx* c, const double* d_rgsw_f, uint32_t nlut, const uint32_t* d_table,
                                    size_t table_stride, int out_bits, uint32_t* d_out32, void* stream) {
    if (!c || !d_rgsw_f || !d_table || !d_out32 || out_bits < 1 || out_bits > 8)
        return fail(TFHE_E_PARAM, "lut8: bad argument");
    if (!nlut) return TFHE_OK;
    launch_lut_vertical(reinterpret_cast<const cplx*>(d_rgsw_f), nlut, d_table, table_stride, 8, out_bits, c->tables.T,
                        d_out32, (hipStream_t)stream);
    return check_launch("lut_vertical");
}

extern "C" int tfhe_lut8x32_eval_batch(tfhe_cbs_ctx* c, const double* d_rgsw_f, uint32_t nbytes, const uint32_t* d_tables,
                                      uint32_t* d_out32, void* stream) {
    if (!c || !d_rgsw_f || !d_tables || !d_out32) return fail(TFHE_E_PARAM, "lut8x32: bad argument");
    if (nbytes > 16 * c->max_blocks) return fail(TFHE_E_CAPACITY, "lut8x32: bytes exceed the context capacity");
    if (!nbytes) return TFHE_OK;
    launch_lut_vertical(reinterpret_cast<const cplx*>(d_rgsw_f), 4 * nbytes, d_tables, kN, 8, 8, c->tables.T, d_out32,
                        (hipStream_t)stream, 4);
    return check_launch("lut8x32");
}

static int cbs_keyswitch(tfhe_cbs_ctx* c, const CbsScratch& sc, const uint32_t* d_lwe32, uint32_t count,
                         uint32_t* d_out, hipStream_t st) {
    if (c->p.ks_linear) {  // the digit sum as an int8 GEMM, in chunks the scratch holds
        const uint32_t chunk = std::min<uint32_t>(2 * (uint32_t)sc.cap, kPfklChunk);
        if (launch_ks_linear(d_lwe32, count, c->p.n, c->d_kskl, sc.pfkl_D, chunk, d_out, st))
            return fail(TFHE_E_PARAM, "cbs keyswitch: unsupported shape");
        return check_launch("linear keyswitch N=2048");
    }
    if (launch_keyswitch32(d_lwe32, count, kN, c->p.n, c->p.ks_logB, c->p.ks_l, c->d_ksk2, sc.ks_at, d_out, st,
                           tuned(c->tune_ks, count, 0)))
        return fail(TFHE_E_PARAM, "cbs keyswitch: unsupported shape");
    return check_launch("keyswitch32 N=2048");
}

extern "C" int tfhe_cbs_keyswitch_batch(tfhe_cbs_ctx* c, const uint32_t* d_lwe32, uint32_t count, uint32_t* d_out,
                                        void* stream) {
    if (!c || !c->has_keys || !d_lwe32 || !d_out) return fail(TFHE_E_PARAM, "cbs keyswitch: bad argument");
    if (count > c->max_batch) return fail(TFHE_E_CAPACITY, "cbs keyswitch: count exceeds the context capacity");
    if (!count) return TFHE_OK;
    return cbs_keyswitch(c, c->sc, d_lwe32, count, d_out, (hipStream_t)stream);
}

namespace {
// Times launch(choice) for each candidate with events on st (the first candidate is the
// model's own choice, run once untimed to warm up) and returns the fastest.  The model's
// choice is kept unless another is at least 2% faster, and alone when one run of it
// takes longer than budget_ms (a large launch: its shape is whole rounds anyway).
template <class Launch>
int tune_one(hipStream_t st, hipEvent_t e0, hipEvent_t e1, const std::vector<int>& cand, float budget_ms,
             Launch&& launch, int& best) {
    auto timed = [&](int choice, int reps, float& ms) -> int {
        ms = 1e30f;
        for (int r = 0; r < reps; ++r) {
            int rc = hipcheck(hipEventRecord(e0, st), "autotune event");
            if (!rc) launch(choice);
            if (!rc) rc = check_launch("autotune");
            if (!rc) rc = hipcheck(hipEventRecord(e1, st), "autotune event");
            if (!rc) rc = hipcheck(hipEventSynchronize(e1), "autotune sync");
            float t = 0.f;
            if (!rc) rc = hipcheck(hipEventElapsedTime(&t, e0, e1), "autotune time");
            if (rc) return rc;
            if (t < ms) ms = t;
        }
        return TFHE_OK;
    };
    best = cand[0];
    launch(cand[0]);
    float t0 = 0.f;
    int rc = timed(cand[0], 1, t0);
    if (rc || t0 > budget_ms) return rc;
    const int reps = t0 < 20.f ? 3 : 1;
    if (reps > 1 && (rc = timed(cand[0], reps, t0))) return rc;
    float tbest = t0;
    for (size_t k = 1; k < cand.size(); ++k) {
        if (cand[k] == cand[0]) continue;
        float t = 0.f;
        if ((rc = timed(cand[k], reps, t))) return rc;
        if (t < tbest && t < 0.98f * t0) tbest = t, best = cand[k];
    }
    return TFHE_OK;
}
}  // namespace

// Times the launch shapes of one AES round over `blocks` blocks (128 circuit bootstraps
// per block) and keeps the fastest in the context for later calls of that size: the PBS
// FFT form of each part of tfhe_cbs_batch's split, the PFKS i-slices of each part and
// the N=2048 key switch's i-slices.  Inputs are hash-filled context buffers; the round
// state, RGSW and LUT buffers of the context are overwritten.  The choices change
// launch shapes only: the exact PBS, the PFKS and the key switch give the same words.
// Captured keystreams hold the launch shapes they were captured with: wait for the
// context's streams and drop them, so that the next graph call captures the current shapes.
static int drop_graphs(tfhe_cbs_ctx* c, const char* what) {
    int rc = TFHE_OK;
    if (c->gstream) rc = hipcheck(hipStreamSynchronize(c->gstream), what);
    if (!rc && c->cstream) rc = hipcheck(hipStreamSynchronize(c->cstream), what);
    if (rc) return rc;
    for (auto& g : c->graphs) (void)hipGraphExecDestroy(g.exec);
    c->graphs.clear();
    return TFHE_OK;
}

extern "C" int tfhe_cbs_set_pbs_form(tfhe_cbs_ctx* c, uint32_t inputs, int form) {
    if (!c) return fail(TFHE_E_PARAM, "set_pbs_form: NULL context");
    const int nforms = pbs_auto(c->p) ? 5 : pbs_mb(c->p) == 2 ? 3 : 2;
    if (form >= nforms) return fail(TFHE_E_PARAM, "set_pbs_form: no such PBS form for this context");
    const int rc = drop_graphs(c, "set_pbs_form sync");
    if (rc) return rc;
    if (form < 0)
        c->tune_pbs.erase(inputs);
    else
        c->tune_pbs[inputs] = form;
    return TFHE_OK;
}

extern "C" int tfhe_cbs_autotune(tfhe_cbs_ctx* c, uint32_t blocks, int* choices, void* stream) {
    if (!c || !choices) return fail(TFHE_E_PARAM, "autotune: bad argument");
    if (!c->has_keys) return fail(TFHE_E_STATE, "autotune: no keys");
    if (blocks > c->max_blocks) return fail(TFHE_E_CAPACITY, "autotune: blocks exceed the context capacity");
    for (int k = 0; k < 5; ++k) choices[k] = -1;
    if (!blocks) return TFHE_OK;
    hipStream_t st = (hipStream_t)stream;
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    int rc = hipcheck(hipStreamIsCapturing(st, &cap), "autotune capture status");
    if (rc) return rc;
    if (cap != hipStreamCaptureStatusNone) return fail(TFHE_E_STATE, "autotune: the stream is capturing");
    // captured keystreams keep the shapes they were captured with
    rc = drop_graphs(c, "autotune sync");
    if (rc) return rc;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    rc = hipcheck(hipEventCreate(&e0), "autotune event");
    if (!rc) rc = hipcheck(hipEventCreate(&e1), "autotune event");
    const uint32_t nb = 128 * blocks;
    const int n = c->p.n, limbs = pbs_limbs(c->p), mb = pbs_mb(c->p);
    const float budget_ms = 250.f;
    // the HomTrace form's PBS writes whole accumulators with its own test polynomial
    const PbsLut io = htss(c->p) ? htss_pbs_io(c, c->sc.lwe64) : PbsLut{};
    if (!rc) {
        launch_fill_hash32(c->sc.state[0], (size_t)nb * (n + 1), 0x9e3779b9u, st);
        launch_fill_hash32(c->sc.lut32, (size_t)nb * (kN + 1), 0x7f4a7c15u, st);
        rc = check_launch("autotune inputs");
    }
    uint32_t full = 0, tail = 0;
    cbs_split(nb, full, tail);
    const uint32_t parts[2] = {full, tail};
    // 16 or more rounds of resident workgroups: the shape is whole rounds whatever the
    // choice, and one timed launch takes seconds; the models' choices stand
    const uint32_t many = 16 * 2 * device_cus();
    for (int part = 0; part < 2 && !rc; ++part) {
        const uint32_t batch = parts[part];
        if (!batch) continue;
        if (batch >= many) {
            choices[part] = pbs_auto(c->p) ? auto_pbs_model(batch) : pbs_model_form(batch, limbs, mb);
            // the linear-form PFKS is one GEMM shape, HomTrace has none: no i-slices to report
            choices[2 + part] = (c->p.pfks_linear || htss(c->p)) ? -1 : pfks_model_slices(2 * batch, 2);
            continue;
        }
        const int model_form = pbs_auto(c->p) ? auto_pbs_model(batch) : pbs_model_form(batch, limbs, mb);
        int form = model_form;
        if (model_form >= 0) {
            // auto contexts: 2 / 3 / 4 = two key bits per step (two waves per SIMD / wide /
            // split over 8 waves), 0 / 1 = one bit with that FFT form
            std::vector<int> cand{model_form};
            for (int f : pbs_auto(c->p) ? std::vector<int>{4, 3, 2, 1, 0}
                         : mb == 2        ? std::vector<int>{2, 1, 0}
                                          : std::vector<int>{1 - model_form})
                if (f != model_form) cand.push_back(f);
            const cplx* bsk2 = c->d_bsk2f + bsk2_mb2_first(c->p) * 8 * 2 * 1024;
            rc = tune_one(st, e0, e1, cand, budget_ms,
                          [&](int f) {
                              if (f >= 2 && pbs_auto(c->p))
                                  launch_pbs2(c->sc.state[0], batch, n, bsk2, limbs, 2, c->tables.T, c->sc.lwe64, st,
                                              f - 2);
                              else
                                  launch_pbs2(c->sc.state[0], batch, n, c->d_bsk2f, limbs, mb, c->tables.T,
                                              c->sc.lwe64, st, f, io);
                          },
                          form);
            if (!rc) c->tune_pbs[batch] = form;
        } else {
            // one form: its output still feeds the PFKS timing below
            launch_pbs2(c->sc.state[0], batch, n, c->d_bsk2f, limbs, mb, c->tables.T, c->sc.lwe64, st, -1, io);
            rc = check_launch("autotune pbs");
        }
        choices[part] = form;
        if (rc) break;
        const uint32_t count = 2 * batch;
        if (c->p.pfks_linear || htss(c->p)) continue;  // one GEMM shape or no PFKS: nothing to choose
        launch_pfks_digits(c->sc.lwe64, count, c->sc.at, st);
        int sl = 0;
        rc = tune_one(st, e0, e1, {pfks_model_slices(count, 2), 4, 6, 8, 12, 16, 24, 32}, budget_ms,
                      [&](int s) {
                          launch_pfks32(c->sc.at, count, c->d_pfks, pfks_words_per_u(), 0, 2, c->sc.rgsw,
                                        4 * (size_t)kN, 2 * (size_t)kN, st, s);
                      },
                      sl);
        if (rc) break;
        c->tune_pfks[count] = sl;
        choices[2 + part] = sl;
    }
    if (!rc && c->p.ks_linear) {
        choices[4] = -1;  // one GEMM shape: no i-slices
    } else if (!rc && nb >= many) {
        choices[4] = ks_model_slices(nb, kN, n);
    } else if (!rc) {
        int sl = 0;
        rc = tune_one(st, e0, e1, {ks_model_slices(nb, kN, n), 1, 2, 4, 6, 8, 12, 16, 24, 32, 48, 64}, budget_ms,
                      [&](int s) {
                          launch_keyswitch32(c->sc.lut32, nb, kN, n, c->p.ks_logB, c->p.ks_l, c->d_ksk2, c->sc.ks_at,
                                             c->sc.state[1], st, s);
                      },
                      sl);
        if (!rc) c->tune_ks[nb] = sl, choices[4] = sl;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    return rc;
}

static int polymul2048(const int32_t* d_x, const uint64_t* d_y, uint32_t batch, uint64_t* d_out, void* stream,
                       int limbs) {
    if (!d_x || !d_y || !d_out) return fail(TFHE_E_PARAM, "polymul2048: NULL argument");
    if (!batch) return TFHE_OK;
    DeviceTables1024* tables = nullptr;
    int rc = stateless_tables(tables);
    if (rc) return rc;
    launch_polymul2048(d_x, d_y, batch, tables->T, d_out, (hipStream_t)stream, limbs);
    return check_launch("polymul2048");
}

extern "C" int tfhe_polymul2048_batch(const int32_t* d_x, const uint64_t* d_y, uint32_t batch, uint64_t* d_out,
                                      void* stream) {
    return polymul2048(d_x, d_y, batch, d_out, stream, 2);
}

extern "C" int tfhe_polymul2048_approx_batch(const int32_t* d_x, const uint64_t* d_y, uint32_t batch,
                                             uint64_t* d_out, void* stream) {
    return polymul2048(d_x, d_y, batch, d_out, stream, 1);
}

extern "C" int tfhe_polymul2048_split32_batch(const int32_t* d_x, const uint32_t* d_y, uint32_t batch, int form,
                                              uint32_t* d_out, void* stream) {
    if (!d_x || !d_y || !d_out) return fail(TFHE_E_PARAM, "polymul2048_split32: NULL argument");
    if (form < 0 || form > 2) return fail(TFHE_E_PARAM, "polymul2048_split32: form must be 0, 1 or 2");
    if (!batch) return TFHE_OK;
    DeviceTables1024* tables = nullptr;
    int rc = stateless_tables(tables);
    if (rc) return rc;
    DeviceTables1024F* ftables = nullptr;
    rc = per_device_tables(ftables, [](DeviceTables1024F& t) { return t.init(); });
    if (rc) return rc;
    if (launch_polymul2048_split32(d_x, d_y, batch, ftables->T, tables->T, d_out, (hipStream_t)stream, form))
        return fail(TFHE_E_PARAM, "polymul2048_split32: unsupported form");
    return check_launch("polymul2048_split32");
}

// ---------------------------------------------------------------------------
// AES-128 transciphering
// ---------------------------------------------------------------------------
extern "C" void tfhe_aes_encrypt_key_bits(int n, const int32_t* sk, uint64_t seed, const uint8_t key[16], int nshift,
                                          uint32_t* out) {
    uint8_t rk[176];
    key_expand(key, rk);
    for (int r = 0; r < 11; ++r)
        for (int b = 0; b < 128; ++b) {
            const uint32_t bit = (rk[16 * r + b / 8] >> (b % 8)) & 1u;
            tfhe_lwe_encrypt(n, sk, seed, (uint32_t)(r * 128 + b), bit, 2, nshift, out + ((size_t)r * 128 + b) * (n + 1));
        }
}

extern "C" void tfhe_aes128_encrypt_plain(const uint8_t key[16], const uint8_t in[16], uint8_t out[16]) {
    uint8_t rk[176], s[16];
    key_expand(key, rk);
    for (int i = 0; i < 16; ++i) s[i] = (uint8_t)(in[i] ^ rk[i]);
    for (int round = 1; round <= 10; ++round) {
        for (int i = 0; i < 16; ++i) s[i] = sbox(s[i]);
        uint8_t t[16];
        shift_mix(s, t, round != 10);
        for (int i = 0; i < 16; ++i) s[i] = (uint8_t)(t[i] ^ rk[16 * round + i]);
    }
    std::memcpy(out, s, 16);
}

extern "C" int tfhe_aes_set_round_keys(tfhe_cbs_ctx* c, const uint32_t* rk_bits) {
    if (!c || !rk_bits) return fail(TFHE_E_PARAM, "aes_set_round_keys: bad argument");
    int rc = hipcheck(hipMemcpy(c->d_rk, rk_bits, (size_t)11 * 128 * (c->p.n + 1) * sizeof(uint32_t),
                                hipMemcpyHostToDevice), "round keys");
    if (!rc) c->has_round_keys = true;
    return rc;
}

// one AES round of `blocks` blocks with scratch set sc (sc.state[0 or 1] may be d_in or
// d_out; the other is the key switch's output); dir 0 = the cipher (round keys d_rk),
// 1 = the equivalent inverse cipher (decryption round keys d_dk)
static int aes_round_impl(tfhe_cbs_ctx* c, CbsScratch& sc, int dir, int round, const uint32_t* d_in,
                          uint32_t blocks, uint32_t* d_out, hipStream_t st) {
    const uint32_t nb = 128 * blocks;
    const int n = c->p.n;
    const uint32_t* rk = (dir ? c->d_dk : c->d_rk) + (size_t)round * 128 * (n + 1);
    int rc = cbs_batch_impl(c, sc, d_in, nb, sc.rgsw, st);
    if (!rc) rc = tfhe_rgsw_to_fourier(c, sc.rgsw, nb, reinterpret_cast<double*>(sc.rgsw_f), st);
    uint32_t* sb = (d_out == sc.state[0]) ? sc.state[1] : sc.state[0];
    if (c->round_form == 1 && round < 10) {
        // T-table round: per byte four table polynomials of its row's T-table (Te or Td),
        // the four words of a column XORed at dimension N, one key switch per bit, then the
        // round key (the reference's run_round_batch, main.cu:1438-1554)
        if (!rc) rc = ensure_lutw(sc);
        if (!rc) {
            launch_lut_vertical(reinterpret_cast<const cplx*>(sc.rgsw_f), 64 * blocks,
                                c->d_ttab + (size_t)dir * 16 * kN, kN, 8, 8, c->tables.T, sc.lutw, st, 4, 4);
            launch_aes_ttable_xor(sc.lutw, blocks, dir, sc.lut32, st);
            rc = check_launch("aes ttable");
        }
        if (!rc) rc = cbs_keyswitch(c, sc, sc.lut32, nb, sb, st);
        if (!rc) {
            launch_aes_linear(sb, rk, c->d_lin_id, 1, blocks, n, d_out, st);
            rc = check_launch("aes round key");
        }
        return rc;
    }
    if (!rc)
        rc = tfhe_lut8_eval_batch(c, reinterpret_cast<const double*>(sc.rgsw_f), 16 * blocks,
                                  dir ? c->d_inv_sbox_poly : c->d_sbox_poly, 0, 8, sc.lut32, st);
    // SubBytes output into the spare state buffer, then the linear layer into d_out
    if (!rc) rc = cbs_keyswitch(c, sc, sc.lut32, nb, sb, st);
    if (!rc) {
        if (dir)
            launch_aes_linear(sb, rk, round == 10 ? c->d_lin_inv_last : c->d_lin_inv_mix, kLinSrcDec, blocks, n, d_out,
                              st);
        else
            launch_aes_linear(sb, rk, round == 10 ? c->d_lin_last : c->d_lin_mix, kLinSrc, blocks, n, d_out, st);
        rc = check_launch("aes_linear");
    }
    return rc;
}

// One AES round (1..10) on device LWE states [blocks][128][n+1]:
// SubBytes (circuit bootstrap of every bit, S-box vertical packing, key switch),
// then ShiftRows, MixColumns (rounds 1..9) and AddRoundKey as LWE additions.
extern "C" int tfhe_aes_round_batch(tfhe_cbs_ctx* c, int round, const uint32_t* d_in, uint32_t blocks, uint32_t* d_out,
                                    void* stream) {
    if (!c || !d_in || !d_out || round < 1 || round > 10) return fail(TFHE_E_PARAM, "aes_round: bad argument");
    if (!c->has_round_keys) return fail(TFHE_E_STATE, "aes_round: no round keys");
    if (blocks > c->max_blocks) return fail(TFHE_E_CAPACITY, "aes_round: blocks exceed the context capacity");
    if (!blocks) return TFHE_OK;
    return aes_round_impl(c, c->sc, 0, round, d_in, blocks, d_out, (hipStream_t)stream);
}

// One round (1..10) of the equivalent inverse cipher (FIPS-197 5.3.5; the reference's
// Td-table decryption rounds): InvSubBytes, InvShiftRows, InvMixColumns (rounds 1..9) and
// the decryption round key, in the context's round form.
extern "C" int tfhe_aes_dec_round_batch(tfhe_cbs_ctx* c, int round, const uint32_t* d_in, uint32_t blocks,
                                        uint32_t* d_out, void* stream) {
    if (!c || !d_in || !d_out || round < 1 || round > 10) return fail(TFHE_E_PARAM, "aes_dec_round: bad argument");
    if (!c->has_dec_round_keys) return fail(TFHE_E_STATE, "aes_dec_round: no decryption round keys");
    if (blocks > c->max_blocks) return fail(TFHE_E_CAPACITY, "aes_dec_round: blocks exceed the context capacity");
    if (!blocks) return TFHE_OK;
    return aes_round_impl(c, c->sc, 1, round, d_in, blocks, d_out, (hipStream_t)stream);
}

extern "C" int tfhe_aes_set_round_form(tfhe_cbs_ctx* c, int form) {
    if (!c || form < 0 || form > 1) return fail(TFHE_E_PARAM, "aes_set_round_form: form must be 0 or 1");
    if (form == c->round_form) return TFHE_OK;
    int rc = TFHE_OK;
    // captured keystreams hold the launches of the old form
    if (c->gstream) rc = hipcheck(hipStreamSynchronize(c->gstream), "round form sync");
    if (rc) return rc;
    for (auto& g : c->graphs) (void)hipGraphExecDestroy(g.exec);
    c->graphs.clear();
    if (form == 1) {  // the LUT word buffers before any call that may be captured
        rc = ensure_lutw(c->sc);
        if (!rc && c->tail.cap) rc = ensure_lutw(c->tail);
        if (rc) return rc;
    }
    c->round_form = form;
    return TFHE_OK;
}

// the keystream's kernels after the counter upload: state init from round key 0
// and the counter blocks in d_ctr, then rounds 1..10 ping-ponging the state buffers
// One round with every circuit bootstrap, LUT and key switch launched on its own
// (the per-bootstrap form): the same kernels and buffers as tfhe_aes_round_batch,
// so the output is the same bit for bit.
static int aes_round_plain(tfhe_cbs_ctx* c, int round, const uint32_t* d_in, uint32_t blocks, uint32_t* d_out,
                           hipStream_t st) {
    const uint32_t nb = 128 * blocks;
    const size_t rs = (size_t)c->p.n + 1, rgsw_words = (size_t)4 * 2 * kN, rgsw_f = (size_t)8 * 1024;
    int rc = TFHE_OK;
    for (uint32_t b = 0; b < nb && !rc; ++b)
        rc = tfhe_cbs_batch(c, d_in + b * rs, 1, c->sc.rgsw + b * rgsw_words, st);
    for (uint32_t b = 0; b < nb && !rc; ++b)
        rc = tfhe_rgsw_to_fourier(c, c->sc.rgsw + b * rgsw_words, 1,
                                  reinterpret_cast<double*>(c->sc.rgsw_f + b * rgsw_f), st);
    for (uint32_t lut = 0; lut < 16 * blocks && !rc; ++lut)
        rc = tfhe_lut8_eval_batch(c, reinterpret_cast<const double*>(c->sc.rgsw_f + (size_t)8 * lut * rgsw_f), 1,
                                  c->d_sbox_poly, 0, 8, c->sc.lut32 + (size_t)8 * lut * (kN + 1), st);
    uint32_t* sb = (d_out == c->sc.state[0]) ? c->sc.state[1] : c->sc.state[0];
    for (uint32_t b = 0; b < nb && !rc; ++b)
        rc = tfhe_cbs_keyswitch_batch(c, c->sc.lut32 + (size_t)b * (kN + 1), 1, sb + b * rs, st);
    if (!rc) {
        launch_aes_linear(sb, c->d_rk + (size_t)round * 128 * rs, round == 10 ? c->d_lin_last : c->d_lin_mix, kLinSrc,
                          blocks, c->p.n, d_out, st);
        rc = check_launch("aes_linear");
    }
    return rc;
}

// state init from round key 0 and the counter blocks d_ctr, then rounds 1..10 ping-ponging
// the scratch set's state buffers
static int aes_group(tfhe_cbs_ctx* c, CbsScratch& sc, const uint8_t* d_ctr, uint32_t blocks, uint32_t* d_out,
                     hipStream_t st, bool plain, int dir = 0) {
    launch_aes_init(dir ? c->d_dk : c->d_rk, d_ctr, blocks, c->p.n, sc.state[0], st);
    int rc = check_launch("aes_init");
    int cur = 0;
    for (int round = 1; round <= 10 && !rc; ++round) {
        // d_out == NULL: the last round writes the scratch set's state[0] (where the
        // ping-pong of ten rounds lands anyway), for a captured graph that must not name
        // the caller's buffer
        uint32_t* dst = (round == 10 && d_out) ? d_out : sc.state[cur ^ 1];
        rc = plain ? aes_round_plain(c, round, sc.state[cur], blocks, dst, st)
                   : aes_round_impl(c, sc, dir, round, sc.state[cur], blocks, dst, st);
        cur ^= 1;
    }
    return rc;
}

// The keystream's kernels after the counter upload.  An uneven batch (full rounds of
// resident PBS workgroups plus a remainder of fewer blocks than one round holds) runs as
// two groups: the full-round blocks on st, the remainder on the second stream with its own
// scratch, so that the remainder's bootstraps fill the full group's idle slots instead of
// a partial round of their own every AES round.  Blocks are independent, so both groups
// give the words of the single-group keystream.
// Blocks of the full-round group of an uneven batch (0: one group).
// (the lookup-form PFKS is long enough to hide the partial PBS round behind it inside
// tfhe_cbs_batch; grouping measured slower there: 17 blocks 15.3 against 16.1 blocks/s)
static uint32_t aes_full_group_blocks(const tfhe_cbs_ctx* c, uint32_t blocks, bool plain) {
    uint32_t b1 = 0;
    if (!plain && c->tail_blocks) {  // only linear-PFKS contexts hold the second scratch set
        const uint32_t rb = 2 * device_cus() / 128;
        b1 = blocks / rb * rb;
        if (b1 == blocks || blocks - b1 > c->tail_blocks) b1 = 0;
    }
    return b1;
}

// d_out == NULL: each group's keystream stays in its scratch set's state[0]
// d_in: the public input bytes of each block (counter blocks, or ciphertext blocks for
// dir = 1, the decryption), already on the device
static int aes_keystream_kernels(tfhe_cbs_ctx* c, uint32_t blocks, uint32_t* d_out, hipStream_t st,
                                 bool plain = false, int dir = 0, const uint8_t* d_in = nullptr) {
    if (!d_in) d_in = c->d_ctr;
    const uint32_t b1 = aes_full_group_blocks(c, blocks, plain);
    if (!b1) return aes_group(c, c->sc, d_in, blocks, d_out, st, plain, dir);
    int rc = hipcheck(hipEventRecord(c->ev_gfork, st), "group fork");
    if (!rc) rc = hipcheck(hipStreamWaitEvent(c->cstream, c->ev_gfork, 0), "group fork wait");
    if (!rc) rc = aes_group(c, c->sc, d_in, b1, d_out, st, false, dir);
    if (!rc)
        rc = aes_group(c, c->tail, d_in + (size_t)16 * b1, blocks - b1,
                       d_out ? d_out + (size_t)b1 * 128 * (c->p.n + 1) : nullptr, c->cstream, false, dir);
    if (!rc) rc = hipcheck(hipEventRecord(c->ev_gjoin, c->cstream), "group join");
    if (!rc) rc = hipcheck(hipStreamWaitEvent(st, c->ev_gjoin, 0), "group join wait");
    return rc;
}

static int aes_ctr_check(const tfhe_cbs_ctx* c, const uint8_t* h_ctr, uint32_t blocks, const uint32_t* d_out) {
    if (!c || !h_ctr || !d_out) return fail(TFHE_E_PARAM, "aes_ctr: bad argument");
    if (!c->has_round_keys) return fail(TFHE_E_STATE, "aes_ctr: no round keys");
    if (blocks > c->max_blocks) return fail(TFHE_E_CAPACITY, "aes_ctr: blocks exceed the context capacity");
    return TFHE_OK;
}

extern "C" int tfhe_aes_ctr_keystream_batch(tfhe_cbs_ctx* c, const uint8_t* h_ctr, uint32_t blocks, uint32_t* d_out,
                                            void* stream) {
    int rc = aes_ctr_check(c, h_ctr, blocks, d_out);
    if (rc || !blocks) return rc;
    hipStream_t st = (hipStream_t)stream;
    rc = hipcheck(hipMemcpyAsync(c->d_ctr, h_ctr, (size_t)blocks * 16, hipMemcpyHostToDevice, st), "ctr");
    if (rc) return rc;
    return aes_keystream_kernels(c, blocks, d_out, st);
}

extern "C" int tfhe_aes_ctr_keystream_plain(tfhe_cbs_ctx* c, const uint8_t* h_ctr, uint32_t blocks, uint32_t* d_out,
                                            void* stream) {
    int rc = aes_ctr_check(c, h_ctr, blocks, d_out);
    if (rc || !blocks) return rc;
    if (c->round_form != 0) return fail(TFHE_E_PARAM, "aes_ctr_keystream_plain: the per-bootstrap form is the S-box round form");
    hipStream_t st = (hipStream_t)stream;
    rc = hipcheck(hipMemcpyAsync(c->d_ctr, h_ctr, (size_t)blocks * 16, hipMemcpyHostToDevice, st), "ctr");
    if (rc) return rc;
    return aes_keystream_kernels(c, blocks, d_out, st, true);
}

extern "C" int tfhe_aes_ctr_keystream_graph(tfhe_cbs_ctx* c, const uint8_t* h_ctr, uint32_t blocks, uint32_t* d_out,
                                            void* stream) {
    int rc = aes_ctr_check(c, h_ctr, blocks, d_out);
    if (rc || !blocks) return rc;
    hipStream_t st = (hipStream_t)stream;
    if (!c->gstream) {
        rc = hipcheck(hipStreamCreateWithFlags(&c->gstream, hipStreamNonBlocking), "graph stream");
        if (!rc) rc = hipcheck(hipEventCreateWithFlags(&c->ev_in, hipEventDisableTiming), "graph event");
        if (!rc) rc = hipcheck(hipEventCreateWithFlags(&c->ev_out, hipEventDisableTiming), "graph event");
        if (rc) return rc;
    }
    // the counter upload stays on the caller's stream, outside the graph
    rc = hipcheck(hipMemcpyAsync(c->d_ctr, h_ctr, (size_t)blocks * 16, hipMemcpyHostToDevice, st), "ctr");
    if (!rc) rc = hipcheck(hipEventRecord(c->ev_in, st), "event");
    if (!rc) rc = hipcheck(hipStreamWaitEvent(c->gstream, c->ev_in, 0), "event wait");
    if (rc) return rc;
    // the graph writes the keystream into the context's scratch, so one graph per block
    // count serves every output buffer; the copy to d_out runs after the replay, outside it
    hipGraphExec_t exec = nullptr;
    for (size_t k = 0; k < c->graphs.size(); ++k)
        if (c->graphs[k].blocks == blocks) {
            exec = c->graphs[k].exec;
            std::rotate(c->graphs.begin() + k, c->graphs.begin() + k + 1, c->graphs.end());  // most recent last
            break;
        }
    if (!exec) {
        rc = hipcheck(hipStreamBeginCapture(c->gstream, hipStreamCaptureModeThreadLocal), "begin capture");
        if (rc) return rc;
        rc = aes_keystream_kernels(c, blocks, nullptr, c->gstream);
        hipGraph_t graph = nullptr;
        const int rc_end = hipcheck(hipStreamEndCapture(c->gstream, &graph), "end capture");
        if (!rc) rc = rc_end;
        if (!rc) rc = hipcheck(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "graph instantiate");
        if (graph) (void)hipGraphDestroy(graph);
        if (rc) return rc;
        if (c->graphs.size() >= kMaxGraphs) {  // evict the least recently used
            (void)hipGraphExecDestroy(c->graphs.front().exec);
            c->graphs.erase(c->graphs.begin());
        }
        c->graphs.push_back({blocks, exec});
    }
    rc = hipcheck(hipGraphLaunch(exec, c->gstream), "graph launch");
    const size_t row_bytes = (size_t)128 * (c->p.n + 1) * sizeof(uint32_t);
    const uint32_t b1 = aes_full_group_blocks(c, blocks, false);
    if (!rc)
        rc = hipcheck(hipMemcpyAsync(d_out, c->sc.state[0], (b1 ? b1 : blocks) * row_bytes, hipMemcpyDeviceToDevice,
                                     c->gstream), "keystream copy");
    if (!rc && b1)
        rc = hipcheck(hipMemcpyAsync(reinterpret_cast<uint8_t*>(d_out) + b1 * row_bytes, c->tail.state[0],
                                     (blocks - b1) * row_bytes, hipMemcpyDeviceToDevice, c->gstream),
                      "keystream copy");
    if (!rc) rc = hipcheck(hipEventRecord(c->ev_out, c->gstream), "event");
    if (!rc) rc = hipcheck(hipStreamWaitEvent(st, c->ev_out, 0), "event wait");
    return rc;
}

extern "C" int tfhe_aes_ctr_transcipher_batch(tfhe_cbs_ctx* c, int mode, const uint8_t* h_ctr, const uint8_t* h_cipher,
                                              uint32_t blocks, uint32_t* d_out, void* stream) {
    int rc = aes_ctr_check(c, h_ctr, blocks, d_out);
    if (rc) return rc;
    if (!h_cipher || mode < 0 || mode > 2) return fail(TFHE_E_PARAM, "aes_transcipher: bad argument");
    if (!blocks) return TFHE_OK;
    rc = hipcheck(hipMemcpyAsync(c->d_pub, h_cipher, (size_t)blocks * 16, hipMemcpyHostToDevice, (hipStream_t)stream),
                  "public ciphertext");
    if (rc) return rc;
    rc = mode == 1   ? tfhe_aes_ctr_keystream_graph(c, h_ctr, blocks, d_out, stream)
         : mode == 2 ? tfhe_aes_ctr_keystream_plain(c, h_ctr, blocks, d_out, stream)
                     : tfhe_aes_ctr_keystream_batch(c, h_ctr, blocks, d_out, stream);
    if (rc) return rc;
    return tfhe_lwe_xor_public(d_out, c->d_pub, 128 * blocks, c->p.n, stream);
}

extern "C" void tfhe_aes_encrypt_dec_key_bits(int n, const int32_t* sk, uint64_t seed, const uint8_t key[16],
                                              int nshift, uint32_t* out) {
    // decryption schedule of the equivalent inverse cipher: dk[0] = rk[10],
    // dk[r] = InvMixColumns(rk[10 - r]) (r = 1..9), dk[10] = rk[0]
    uint8_t rk[176], dk[176];
    key_expand(key, rk);
    std::memcpy(dk, rk + 160, 16);
    std::memcpy(dk + 160, rk, 16);
    for (int r = 1; r <= 9; ++r) {
        uint8_t t[16];  // InvMixColumns alone: undo the InvShiftRows of inv_shift_mix first
        for (int cc = 0; cc < 4; ++cc)
            for (int row = 0; row < 4; ++row) t[row + 4 * cc] = rk[16 * (10 - r) + row + 4 * ((cc + row) % 4)];
        inv_shift_mix(t, dk + 16 * r, true);
    }
    for (int r = 0; r < 11; ++r)
        for (int b = 0; b < 128; ++b) {
            const uint32_t bit = (dk[16 * r + b / 8] >> (b % 8)) & 1u;
            tfhe_lwe_encrypt(n, sk, seed, (uint32_t)(11 * 128 + r * 128 + b), bit, 2, nshift,
                             out + ((size_t)r * 128 + b) * (n + 1));
        }
}

extern "C" int tfhe_aes_set_dec_round_keys(tfhe_cbs_ctx* c, const uint32_t* dk_bits) {
    if (!c || !dk_bits) return fail(TFHE_E_PARAM, "aes_set_dec_round_keys: bad argument");
    int rc = hipcheck(hipMemcpy(c->d_dk, dk_bits, (size_t)11 * 128 * (c->p.n + 1) * sizeof(uint32_t),
                                hipMemcpyHostToDevice),
                      "dec round keys");
    if (!rc) c->has_dec_round_keys = true;
    return rc;
}

// Homomorphic AES-128 decryption of public ciphertext blocks under the encrypted key (the
// reference's default transciphering path: Td tables and InvMixColumns round keys,
// main.cu:279-317, key_material.cuh:12-47): the state starts as the ciphertext bits plus
// dk[0], then 10 rounds of the equivalent inverse cipher in the context's round form.
extern "C" int tfhe_aes_decrypt_batch(tfhe_cbs_ctx* c, const uint8_t* h_cipher, uint32_t blocks, uint32_t* d_out,
                                      void* stream) {
    if (!c || !h_cipher || !d_out) return fail(TFHE_E_PARAM, "aes_decrypt: bad argument");
    if (!c->has_dec_round_keys) return fail(TFHE_E_STATE, "aes_decrypt: no decryption round keys");
    if (blocks > c->max_blocks) return fail(TFHE_E_CAPACITY, "aes_decrypt: blocks exceed the context capacity");
    if (!blocks) return TFHE_OK;
    hipStream_t st = (hipStream_t)stream;
    int rc = hipcheck(hipMemcpyAsync(c->d_pub, h_cipher, (size_t)blocks * 16, hipMemcpyHostToDevice, st), "cipher");
    if (rc) return rc;
    return aes_keystream_kernels(c, blocks, d_out, st, false, 1, c->d_pub);
}

extern "C" int tfhe_lwe_xor_public(uint32_t* d_ct, const uint8_t* d_pub, uint32_t nbits, int n, void* stream) {
    if (!d_ct || !d_pub || n <= 0) return fail(TFHE_E_PARAM, "xor_public: bad argument");
    if (!nbits) return TFHE_OK;
    launch_xor_public(d_ct, d_pub, nbits, n, (hipStream_t)stream);
    return check_launch("xor_public");
}
