// gate.hip -- gfx950 kernels of the TFHE gate-bootstrapping path (N = 1024, torus-32):
// key generation, BSK Fourier conversion, the persistent batched blind rotation,
// the plain per-step form, sample extraction and the batched LWE key switch.
#include <atomic>

#include "device_common.hpp"
#include "kernels.hpp"

namespace tfhe {

// ===========================================================================
// Key generation (replaces the OpenFHE fork's KeySwitchGen and the reference's
// GPUKeyGenAcc, bitwise/src/binfhe-base-scheme.cu:18-97, rgsw-acc.cu:13-122).
// ===========================================================================

// BSK row (i, r): RLWE_K(msg) with msg = s_i * g_j placed in component u = r & 1,
// j = r >> 1, g_j = 2^(32 - l*logB + j*logB).  Coefficient layout [n][2l][2][N].
__global__ __launch_bounds__(256) void k_gen_bsk32(uint64_t seed, int n, int N, int logB, int l,
                                                   int nshift, const int32_t* __restrict__ s,
                                                   const int32_t* __restrict__ K,
                                                   uint32_t* __restrict__ bsk) {
    extern __shared__ uint32_t sh[];
    uint32_t* a = sh;             // [N]
    int32_t* Ks = (int32_t*)(sh + N);  // [N]
    const int rows = 2 * l;
    const int row = blockIdx.x;   // i*rows + r
    const int i = row / rows, r = row % rows;
    const int j = r >> 1, u = r & 1;
    for (int c = threadIdx.x; c < N; c += blockDim.x) {
        a[c] = rand32(seed, DOM_BSK1, (uint32_t)row, (uint32_t)c);
        Ks[c] = K[c];
    }
    __syncthreads();
    const uint32_t g = 1u << (32 - l * logB + j * logB);
    const uint32_t msg0 = (uint32_t)s[i] * g;
    uint32_t* outa = bsk + (size_t)row * 2 * N;
    uint32_t* outb = outa + N;
    for (int c = threadIdx.x; c < N; c += blockDim.x) {
        uint32_t acc = 0;
        for (int t = 0; t < N; ++t) {
            if (!Ks[t]) continue;
            acc += (t <= c) ? a[c - t] : (0u - a[N + c - t]);
        }
        acc += (uint32_t)noise(seed, DOM_BSK1, (uint32_t)row, (uint32_t)c, nshift);
        uint32_t av = a[c];
        if (c == 0) {
            if (u == 0) av += msg0; else acc += msg0;
        }
        outa[c] = av;
        outb[c] = acc;
    }
}

// KSK entry (i, j, v) = LWE_s(v * K_i * 2^(32 - l*logB + j*logB)), layout [N][l][B][n+1];
// linear = 1: one entry per (i, j) with v = 1 (the linear key of the signed-digit switch),
// layout [N][l][n+1].
__global__ __launch_bounds__(256) void k_gen_ksk32(uint64_t seed, uint32_t dom, int n, int N, int logB, int l,
                                                   int nshift, int linear, const int32_t* __restrict__ s,
                                                   const int32_t* __restrict__ K,
                                                   uint32_t* __restrict__ ksk) {
    __shared__ uint32_t red[256];
    const int B = linear ? 1 : 1 << logB;  // entries per (i, j)
    const uint32_t id = blockIdx.x;  // ((i*l + j)*B + v)
    const int v = linear ? 1 : id % B;
    const int j = (id / B) % l;
    const int i = id / (B * l);
    uint32_t* ct = ksk + (size_t)id * (n + 1);
    uint32_t part = 0;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        uint32_t ak = rand32(seed, dom, id, (uint32_t)k);
        ct[k] = ak;
        part += ak * (uint32_t)s[k];
    }
    red[threadIdx.x] = part;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const int kbits = l * logB;
        uint32_t mu = (uint32_t)v * (uint32_t)K[i] * (1u << (32 - kbits + j * logB));
        ct[n] = red[0] + (uint32_t)noise(seed, dom, id, 0, nshift) + mu;
    }
}

// ===========================================================================
// Fourier conversion of N=1024 torus-32 polynomials into the kernel frequency
// layout [poly][slot 8][lane 64] (one wavefront per polynomial).
// ===========================================================================
__global__ __launch_bounds__(256) void k_to_fourier1024(const uint32_t* __restrict__ coef, size_t npoly,
                                                       FFTTables T, cplx* __restrict__ out) {
    __shared__ cplx lds_all[4][FFT512_LDS_CPLX];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const size_t q = (size_t)blockIdx.x * 4 + w;
    if (q >= npoly) return;
    cplx* lds = lds_all[w];
    const uint32_t* p = coef + q * 1024;
    cplx x[8];
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        x[a] = {(double)(int32_t)p[64 * a + lane], (double)(int32_t)p[64 * a + lane + 512]};
    }
    fft512_fwd(x, lds, T, lane);
    cplx* o = out + q * 512;
#pragma unroll
    for (int f = 0; f < 8; ++f) o[f * 64 + lane] = x[f];
}

// ===========================================================================
// Gate preparation and modulus switch (reference binfhe-base-scheme.cu:151-176,
// rgsw-acc.cu:597-598): prep = ct1 + ct2 (doubled for XOR/XNOR), then
// round(v * 2N / 2^32) mod 2N on every coefficient.
// ===========================================================================
__device__ __forceinline__ uint32_t modswitch_2N(uint32_t v, int log2N2) {
    return (uint32_t)((((uint64_t)v + (1ull << (31 - log2N2))) >> (32 - log2N2)) & ((1u << log2N2) - 1));
}

// Test vector value at coefficient c (reference binfhe-base-scheme.cu:103-138, q = 2N).
__device__ __forceinline__ uint32_t gate_tv(uint32_t q1, uint32_t q2, uint32_t bt, uint32_t c, uint32_t q) {
    const uint32_t Q2p = (1u << 29) + 1u, Q2pNeg = 0u - Q2p;
    uint32_t temp = (bt + q - c) & (q - 1);
    if (q1 < q2) return (temp >= q1 && temp < q2) ? Q2pNeg : Q2p;
    return (temp >= q2 && temp < q1) ? Q2p : Q2pNeg;
}

// ===========================================================================
// Persistent batched gate bootstrap: one wavefront owns one ciphertext for all
// n blind-rotation steps.  Per step: signed decomposition of the register-
// resident accumulator, 6 forward FFTs with the Fourier-domain MAC against the
// streamed BSK row block, 2 inverse FFTs, exact rounding mod 2^32 and the CGGI
// monic update acc += X^m t - t with m = -a~_i mod 2N (reference
// rgsw-acc.cu:606-708 fused into one launch).  Ends with sample extraction
// (coefficient 0) and the + Q/8 + 1 offset; the key switch runs separately.
// ===========================================================================
constexpr int GATE_N = 1024, GATE_LOGB = 7, GATE_L = 3, GATE_ROWS = 6;

// FUNC = true is the programmable bootstrap (reference binfhe-base-scheme.cu:259-378): one input
// ciphertext over Z_ctmod (ctmod = 2N >> logf, a power of two), rotations a_i 2^logf (exact, no
// rounding), test vector m[j 2^logf] = tv[(b - j) mod ctmod] for j < ctmod / 2 (tv already scaled
// to the torus by the host) and zero elsewhere, no offset after extraction.
template <bool FUNC>
__global__ __launch_bounds__(256) void k_gate_blind_rotate(
    const uint32_t* __restrict__ ct1, const uint32_t* __restrict__ ct2, uint32_t batch, int n,
    int doubled, uint32_t q1, const cplx* __restrict__ bskf, FFTTables Tg,
    uint32_t* __restrict__ lweN, int logf, const uint32_t* __restrict__ tv) {
    constexpr int N = GATE_N, M = N / 2;
    __shared__ cplx lds_all[4][FFT512_LDS_CPLX];
    __shared__ uint16_t at_all[4][1024];
    __shared__ cplx tw_lds[1024];
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t cidx = blockIdx.x * 4 + w;
    // twiddle tables staged in LDS: per-lane table reads through the vector memory
    // path stall the FFT passes far more than LDS reads do
    for (int k = threadIdx.x; k < 512; k += blockDim.x) {
        tw_lds[k] = Tg.p1t[k];
        tw_lds[512 + k] = Tg.p2[k];
    }
    __syncthreads();
    const FFTTables T{tw_lds, tw_lds + 512, Tg.rot};
    if (cidx >= batch) return;
    cplx* lds = lds_all[w];
    uint16_t* at = at_all[w];
    const uint32_t* c1 = ct1 + (size_t)cidx * (n + 1);
    constexpr int LOG2N2 = 11;
    const uint32_t q = 2u * N;
    uint32_t acc[2][16];
    if constexpr (FUNC) {
        const uint32_t cmask = (q >> logf) - 1u;
        for (int k = lane; k < n; k += 64) at[k] = (uint16_t)((c1[k] & cmask) << logf);
        const uint32_t b = c1[n] & cmask;
        const uint32_t fmask = (1u << logf) - 1u;
#pragma unroll
        for (int s = 0; s < 16; ++s) {
            const uint32_t c = (uint32_t)(64 * s + lane);  // < N, so j = c >> logf < ctmod / 2
            acc[0][s] = 0u;
            acc[1][s] = (c & fmask) ? 0u : tv[(b - (c >> logf)) & cmask];
        }
    } else {
        const uint32_t* c2 = ct2 + (size_t)cidx * (n + 1);
        for (int k = lane; k < n; k += 64) {
            uint32_t v = c1[k] + c2[k];
            if (doubled) v *= 2u;
            at[k] = (uint16_t)modswitch_2N(v, LOG2N2);
        }
        uint32_t vb = c1[n] + c2[n];
        if (doubled) vb *= 2u;
        const uint32_t bt = modswitch_2N(vb, LOG2N2);
        const uint32_t q2 = (q1 + N) & (q - 1);
#pragma unroll
        for (int s = 0; s < 16; ++s) {
            acc[0][s] = 0u;
            acc[1][s] = gate_tv(q1, q2, bt, (uint32_t)(64 * s + lane), q);
        }
    }
    wave_sync();

    for (int i = 0; i < n; ++i) {
        const uint32_t mrot = (q - (uint32_t)at[i]) & (q - 1);
        if (mrot == 0) continue;
        const cplx* key = bskf + (size_t)i * GATE_ROWS * 2 * M;
        cplx macc[2][8];
#pragma unroll
        for (int o = 0; o < 2; ++o)
#pragma unroll
            for (int f = 0; f < 8; ++f) macc[o][f] = {0.0, 0.0};

        constexpr uint32_t OFF = digit_offset32<GATE_LOGB, GATE_L>();
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            uint32_t tp[16];
#pragma unroll
            for (int s = 0; s < 16; ++s) tp[s] = acc[u][s] + OFF;
#pragma unroll
            for (int j = 0; j < GATE_L; ++j) {
                const int r = 2 * j + u;
                cplx x[8];
#pragma unroll
                for (int a = 0; a < 8; ++a)
                    x[a] = {(double)digit32<GATE_LOGB, GATE_L>(tp[a], j),
                            (double)digit32<GATE_LOGB, GATE_L>(tp[a + 8], j)};
                fft512_fwd(x, lds, T, lane);
                const cplx* kr = key + (size_t)r * 2 * M;
#pragma unroll
                for (int o = 0; o < 2; ++o)
#pragma unroll
                    for (int f = 0; f < 8; ++f) cmac(macc[o][f], x[f], kr[(o * 8 + f) * 64 + lane]);
            }
        }

        // CGGI monic update acc += (X^m - 1) t in the Fourier domain: X[k] is the
        // polynomial at zeta_k = w^(1 - 4k), so X^m multiplies it by w^(m (1 - 4k));
        // one rounding of (X^m - 1) t gives the same integers as rotating the rounded t
        cplx rf[8];
        {
            const uint32_t kl = (uint32_t)((lane >> 3) + 8 * (lane & 7));  // k = kl + 64 f
#pragma unroll
            for (int f = 0; f < 8; ++f) {
                const uint32_t k = kl + 64u * f;
                const uint32_t ex = (mrot * ((2u * N + 1u - 4u * k) & (2u * N - 1u))) & (2u * N - 1u);
                const cplx r = T.rot[ex];
                rf[f] = {r.re - 1.0, r.im};
            }
        }
#pragma unroll
        for (int o = 0; o < 2; ++o) {
            cplx x[8];
#pragma unroll
            for (int f = 0; f < 8; ++f) x[f] = cmul(macc[o][f], rf[f]);
            fft512_inv(x, lds, T, lane);
#pragma unroll
            for (int a = 0; a < 8; ++a) {
                acc[o][a] += round_mod32(x[a].re, 1.0 / M);
                acc[o][a + 8] += round_mod32(x[a].im, 1.0 / M);
            }
        }
    }

    // sample extraction at coefficient 0 (X -> X^{2N-1} automorphism, bitwise/src/kernel.cu:241-250)
    uint32_t* out = lweN + (size_t)cidx * (N + 1);
#pragma unroll
    for (int s = 0; s < 16; ++s) {
        const int c = 64 * s + lane;
        const uint32_t v = acc[0][s];
        out[(N - c) & (N - 1)] = c ? 0u - v : v;
    }
    if (lane == 0) out[N] = FUNC ? acc[1][0] : acc[1][0] + (1u << 29) + 1u;
}

// ===========================================================================
// Plain per-step form (the reference's launch structure, rgsw-acc.cu:606-708):
// per blind-rotation step a decompose+FFT launch (one wavefront per digit row),
// a MAC+inverse-FFT launch (one wavefront per output polynomial) and a monic
// update launch, for one ciphertext at a time.
// ===========================================================================
__global__ __launch_bounds__(64) void k_plain_decompose_fft(const uint32_t* __restrict__ acc, FFTTables T,
                                                           cplx* __restrict__ digf) {
    __shared__ cplx lds[FFT512_LDS_CPLX];
    const int lane = threadIdx.x;
    const int r = blockIdx.x, j = r >> 1, u = r & 1;
    const uint32_t* src = acc + u * GATE_N;
    cplx x[8];
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        int32_t d0[GATE_L], d1[GATE_L];
        decompose32<GATE_LOGB, GATE_L>(src[64 * a + lane], d0);
        decompose32<GATE_LOGB, GATE_L>(src[64 * a + lane + 512], d1);
        x[a] = {(double)d0[j], (double)d1[j]};
    }
    fft512_fwd(x, lds, T, lane);
#pragma unroll
    for (int f = 0; f < 8; ++f) digf[(r * 8 + f) * 64 + lane] = x[f];
}

__global__ __launch_bounds__(64) void k_plain_mac_ifft(const cplx* __restrict__ digf, const cplx* __restrict__ key,
                                                      FFTTables T, uint32_t* __restrict__ t) {
    __shared__ cplx lds[FFT512_LDS_CPLX];
    const int lane = threadIdx.x, o = blockIdx.x;
    cplx x[8];
#pragma unroll
    for (int f = 0; f < 8; ++f) x[f] = {0.0, 0.0};
    for (int r = 0; r < GATE_ROWS; ++r)
#pragma unroll
        for (int f = 0; f < 8; ++f)
            cmac(x[f], digf[(r * 8 + f) * 64 + lane], key[((r * 2 + o) * 8 + f) * 64 + lane]);
    fft512_inv(x, lds, T, lane);
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        t[o * GATE_N + 64 * a + lane] = round_mod32(x[a].re, 1.0 / 512);
        t[o * GATE_N + 64 * a + lane + 512] = round_mod32(x[a].im, 1.0 / 512);
    }
}

// acc += X^m t - t (reference kernels_fft.cu:90-113)
__global__ __launch_bounds__(256) void k_plain_monic(uint32_t* __restrict__ acc, const uint32_t* __restrict__ t,
                                                    uint32_t m) {
    const int o = blockIdx.x;
    const int k = (int)(m & (GATE_N - 1));
    const bool hi = m >= (uint32_t)GATE_N;
    for (int c = threadIdx.x; c < GATE_N; c += blockDim.x) {
        int idx = c - k;
        const bool wrapped = idx < 0;
        if (wrapped) idx += GATE_N;
        uint32_t v = t[o * GATE_N + idx];
        if (hi ^ wrapped) v = 0u - v;
        acc[o * GATE_N + c] += v - t[o * GATE_N + c];
    }
}

__global__ void k_plain_init(uint32_t* __restrict__ acc, uint32_t q1, uint32_t bt) {
    const uint32_t q = 2u * GATE_N, q2 = (q1 + GATE_N) & (q - 1);
    for (int c = threadIdx.x; c < GATE_N; c += blockDim.x) {
        acc[c] = 0u;
        acc[GATE_N + c] = gate_tv(q1, q2, bt, (uint32_t)c, q);
    }
}

__global__ void k_plain_extract(const uint32_t* __restrict__ acc, uint32_t* __restrict__ lweN) {
    for (int c = threadIdx.x; c < GATE_N; c += blockDim.x) {
        const uint32_t v = acc[c];
        lweN[(GATE_N - c) & (GATE_N - 1)] = c ? 0u - v : v;
    }
    if (threadIdx.x == 0) lweN[GATE_N] = acc[GATE_N] + (1u << 29) + 1u;
}

// ===========================================================================
// LWE key switching N -> n over Z/2^32 (reference kernel_LWEKeySwitch_modSwitch,
// bitwise/src/kernel.cu:259-279, with the host half of binfhe-base-scheme.cu:225-246
// folded in): out = (0, b) - sum_{i<N} sum_{j<L} KSK[i][j][digit_j(a~_i)], where
// a~_i = round(a_i / 2^(32 - L LOGB)) and the digits are unsigned base 2^LOGB,
// least significant first (digit 0 selects the key's encryption of zero).
// k_ks_digits rounds and transposes the inputs to [i][Bpad], so that one
// coefficient of a wave's KS_CT ciphertexts is a uniform (scalar) read.
// k_keyswitch32: a workgroup of KS_WAVES waves covers 128 output columns; per
// coefficient i the L * 2^LOGB candidate rows of those columns are staged once
// in a double-buffered LDS tile (one barrier per i) and each wave subtracts,
// for each of its ciphertexts and digits, the row the digit names: one LDS read
// and one subtraction per (ciphertext, digit, column).  blockIdx.x = column
// tile + cols * i-slice and blockIdx.y = ciphertext group, so the groups that
// read one key tile have linear workgroup ids cols * slices apart; with that a
// multiple of 8 they run on one XCD and share its L2.  Slices add into the
// zero-filled output with 32-bit atomics (order-free mod 2^32).
// ===========================================================================
constexpr int KS_WAVES = 8, KS_CT = 16, KS_GROUP = KS_WAVES * KS_CT;

template <int KBITS>
__global__ __launch_bounds__(256) void k_ks_digits(const uint32_t* __restrict__ lweN, uint32_t batch, int N,
                                                  uint32_t Bpad, uint32_t* __restrict__ at) {
    __shared__ uint32_t tile[64][65];
    const int i0 = blockIdx.x * 64, k0 = blockIdx.y * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int r = ty; r < 64; r += 4) {  // r = ciphertext offset, tx = coefficient offset
        const uint32_t k = k0 + r;
        const int i = i0 + tx;
        uint32_t v = 0;
        if (k < batch && i < N)
            v = (uint32_t)(((uint64_t)lweN[(size_t)k * (N + 1) + i] + (1ull << (31 - KBITS))) >> (32 - KBITS));
        tile[r][tx] = v;
    }
    __syncthreads();
    for (int r = ty; r < 64; r += 4) {  // r = coefficient offset, tx = ciphertext offset
        const int i = i0 + r;
        const uint32_t k = k0 + tx;
        if (i < N && k < Bpad) at[(size_t)i * Bpad + k] = tile[tx][r];
    }
}

template <int LOGB, int L>
__global__ __launch_bounds__(64 * KS_WAVES) void k_keyswitch32(const uint32_t* __restrict__ at, uint32_t batch,
                                                              uint32_t Bpad, int N, int n,
                                                              const uint32_t* __restrict__ lweN,
                                                              const uint32_t* __restrict__ ksk, int cols,
                                                              int slices, uint32_t* __restrict__ out) {
    // two adjacent columns per lane (128 per workgroup): one LDS address and one 8-byte
    // read feed two subtractions.  Key rows are n + 1 = 637 words long, so the global
    // loads stay 32-bit (odd rows are not 8-byte aligned).
    constexpr int B = 1 << LOGB, ROWS = L * B, PER = ROWS / KS_WAVES;
    static_assert(ROWS % KS_WAVES == 0, "tile rows must split evenly over the waves");
    __shared__ uint2 tab[2][ROWS][64];
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int tile = (int)blockIdx.x % cols, slice = (int)blockIdx.x / cols;
    const int col0 = tile * 128 + 2 * lane;
    const bool ok0 = col0 <= n, ok1 = col0 + 1 <= n;
    const int islice = (N + slices - 1) / slices;
    const int ibeg = slice * islice, iend = ibeg + islice < N ? ibeg + islice : N;
    if (ibeg >= iend) return;  // empty last slice (uniform over the workgroup)
    const uint32_t cbase = blockIdx.y * KS_GROUP + w * KS_CT;
    const bool active = cbase < batch;  // idle waves still take part in the barriers
    const size_t rs = (size_t)(n + 1);
    uint2 kv[PER];
    auto load_i = [&](int i) {  // this thread's rows w + KS_WAVES q of coefficient i
        uint64_t off = (uint64_t)i * ROWS * rs;
        asm volatile("" : "+s"(off));  // keep the prefetch where it is written
        const uint32_t* base = ksk + off;
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const uint32_t* r = base + (size_t)(w + KS_WAVES * q) * rs;
            kv[q] = make_uint2(ok0 ? r[col0] : 0u, ok1 ? r[col0 + 1] : 0u);
        }
    };
    auto store = [&](int b) {
#pragma unroll
        for (int q = 0; q < PER; ++q) tab[b][w + KS_WAVES * q][lane] = kv[q];
    };
    uint32_t acc[KS_CT][2];
#pragma unroll
    for (int k = 0; k < KS_CT; ++k) acc[k][0] = acc[k][1] = 0u;
    load_i(ibeg);
    store(0);
    if (ibeg + 1 < iend) load_i(ibeg + 1);
    __syncthreads();
    for (int i = ibeg; i < iend; ++i) {
        const int b = (i - ibeg) & 1;
        if (i + 1 < iend) store(b ^ 1);
        if (i + 2 < iend) load_i(i + 2);
        if (active) {
            uint32_t a[KS_CT];
            const uint32_t* ar = at + (size_t)i * Bpad + cbase;
#pragma unroll
            for (int k = 0; k < KS_CT; ++k) a[k] = ar[k];
#pragma unroll
            for (int j = 0; j < L; ++j)
#pragma unroll
                for (int k = 0; k < KS_CT; ++k) {
                    const uint2 v = tab[b][j * B + ((a[k] >> (j * LOGB)) & (B - 1))][lane];
                    acc[k][0] -= v.x;
                    acc[k][1] -= v.y;
                }
        }
#pragma unroll
        for (int k = 0; k < KS_CT; ++k) asm volatile("" : "+v"(acc[k][0]), "+v"(acc[k][1]));  // before the barrier
        __syncthreads();
    }
    if (!active) return;
#pragma unroll
    for (int k = 0; k < KS_CT; ++k) {
        if (cbase + k >= batch) break;
        uint32_t* o = out + (size_t)(cbase + k) * rs;
        const uint32_t bterm = (slice == 0) ? lweN[(size_t)(cbase + k) * (N + 1) + N] : 0u;
        if (ok0) atomicAdd(&o[col0], acc[k][0] + (col0 == n ? bterm : 0u));
        if (ok1) atomicAdd(&o[col0 + 1], acc[k][1] + (col0 + 1 == n ? bterm : 0u));
    }
}

__global__ void k_fill_zero(uint32_t* __restrict__ p, size_t count) {
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < count; k += (size_t)gridDim.x * blockDim.x)
        p[k] = 0u;
}

// ===========================================================================
// Exact negacyclic product of small-integer polynomials x by torus-32
// polynomials y through the wavefront FFT (one wavefront per product).
// ===========================================================================
__global__ __launch_bounds__(256) void k_polymul1024(const int32_t* __restrict__ x, const uint32_t* __restrict__ y,
                                                    uint32_t batch, FFTTables T, uint32_t* __restrict__ out) {
    __shared__ cplx lds_all[4][FFT512_LDS_CPLX];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t q = blockIdx.x * 4 + w;
    if (q >= batch) return;
    cplx* lds = lds_all[w];
    const int32_t* xp = x + (size_t)q * 1024;
    const uint32_t* yp = y + (size_t)q * 1024;
    cplx X[8], Y[8];
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        X[a] = {(double)xp[64 * a + lane], (double)xp[64 * a + lane + 512]};
        Y[a] = {(double)(int32_t)yp[64 * a + lane], (double)(int32_t)yp[64 * a + lane + 512]};
    }
    fft512_fwd(X, lds, T, lane);
    fft512_fwd(Y, lds, T, lane);
#pragma unroll
    for (int f = 0; f < 8; ++f) X[f] = cmul(X[f], Y[f]);
    fft512_inv(X, lds, T, lane);
    uint32_t* op = out + (size_t)q * 1024;
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        op[64 * a + lane] = round_mod32(X[a].re, 1.0 / 512);
        op[64 * a + lane + 512] = round_mod32(X[a].im, 1.0 / 512);
    }
}

// ===========================================================================
// Host launchers
// ===========================================================================
void launch_gen_bsk32(uint64_t seed, int n, int N, int logB, int l, int nshift, const int32_t* s,
                      const int32_t* K, uint32_t* bsk, hipStream_t st) {
    k_gen_bsk32<<<n * 2 * l, 256, 2 * N * sizeof(uint32_t), st>>>(seed, n, N, logB, l, nshift, s, K, bsk);
}

void launch_gen_ksk32(uint64_t seed, int n, int N, int logB, int l, int nshift, const int32_t* s,
                      const int32_t* K, uint32_t* ksk, hipStream_t st, uint32_t dom, int linear) {
    k_gen_ksk32<<<N * l * (linear ? 1 : 1 << logB), 256, 0, st>>>(seed, dom, n, N, logB, l, nshift, linear, s, K,
                                                                   ksk);
}

void launch_to_fourier1024(const uint32_t* coef, size_t npoly, const FFTTables& T, cplx* out, hipStream_t st) {
    k_to_fourier1024<<<(unsigned)((npoly + 3) / 4), 256, 0, st>>>(coef, npoly, T, out);
}

void launch_gate_blind_rotate(const uint32_t* ct1, const uint32_t* ct2, uint32_t batch, int n, int doubled,
                              uint32_t q1, const cplx* bskf, const FFTTables& T, uint32_t* lweN,
                              hipStream_t st) {
    k_gate_blind_rotate<false><<<(batch + 3) / 4, 256, 0, st>>>(ct1, ct2, batch, n, doubled, q1, bskf, T, lweN,
                                                                  0, nullptr);
}

void launch_func_blind_rotate(const uint32_t* ct, uint32_t batch, int n, int logf, const uint32_t* tv,
                              const cplx* bskf, const FFTTables& T, uint32_t* lweN, hipStream_t st) {
    k_gate_blind_rotate<true><<<(batch + 3) / 4, 256, 0, st>>>(ct, nullptr, batch, n, 0, 0u, bskf, T, lweN, logf,
                                                                 tv);
}

// LWE arithmetic over power-of-two moduli, element-wise on [batch][n+1] (reference
// binfhe-base-scheme.cu:380-698 through the OpenFHE LWE operations): out = ((x & mx) - (y & my)
// + [k = n] c) & mo, with y optional.
__global__ __launch_bounds__(256) void k_lwe_axpb(const uint32_t* __restrict__ x, uint32_t mx,
                                                 const uint32_t* __restrict__ y, uint32_t my, uint32_t c,
                                                 uint32_t mo, int n, size_t count, uint32_t* __restrict__ out) {
    const uint32_t w = (uint32_t)n + 1u;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < count; k += (size_t)gridDim.x * blockDim.x) {
        uint32_t v = x[k] & mx;
        if (y) v -= y[k] & my;
        if (k % w == (uint32_t)n) v += c;
        out[k] = v & mo;
    }
}

// round(x 2^-shift) & mo on every element (the ModSwitch between power-of-two moduli)
__global__ __launch_bounds__(256) void k_lwe_modswitch(const uint32_t* __restrict__ x, int shift, uint32_t mo,
                                                      size_t count, uint32_t* __restrict__ out) {
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < count; k += (size_t)gridDim.x * blockDim.x) {
        const uint64_t v = x[k];
        out[k] = (uint32_t)(shift ? (v + (1ull << (shift - 1))) >> shift : v) & mo;
    }
}

static unsigned elementwise_blocks(size_t count) {
    const size_t b = (count + 255) / 256;
    return (unsigned)(b < 4096 ? (b ? b : 1) : 4096);
}

void launch_lwe_axpb(const uint32_t* x, uint32_t mx, const uint32_t* y, uint32_t my, uint32_t c, uint32_t mo, int n,
                     uint32_t batch, uint32_t* out, hipStream_t st) {
    const size_t count = (size_t)batch * (n + 1);
    k_lwe_axpb<<<elementwise_blocks(count), 256, 0, st>>>(x, mx, y, my, c, mo, n, count, out);
}

void launch_lwe_modswitch(const uint32_t* x, int shift, uint32_t mo, int n, uint32_t batch, uint32_t* out,
                          hipStream_t st) {
    const size_t count = (size_t)batch * (n + 1);
    k_lwe_modswitch<<<elementwise_blocks(count), 256, 0, st>>>(x, shift, mo, count, out);
}

void launch_plain_init(uint32_t* acc, uint32_t q1, uint32_t bt, hipStream_t st) {
    k_plain_init<<<1, 256, 0, st>>>(acc, q1, bt);
}

void launch_plain_step(uint32_t* acc, cplx* digf, uint32_t* t, const cplx* key, uint32_t m,
                       const FFTTables& T, hipStream_t st) {
    k_plain_decompose_fft<<<GATE_ROWS, 64, 0, st>>>(acc, T, digf);
    k_plain_mac_ifft<<<2, 64, 0, st>>>(digf, key, T, t);
    k_plain_monic<<<2, 256, 0, st>>>(acc, t, m);
}

void launch_plain_extract(const uint32_t* acc, uint32_t* lweN, hipStream_t st) {
    k_plain_extract<<<1, 256, 0, st>>>(acc, lweN);
}

void launch_polymul1024(const int32_t* x, const uint32_t* y, uint32_t batch, const FFTTables& T, uint32_t* out,
                        hipStream_t st) {
    k_polymul1024<<<(batch + 3) / 4, 256, 0, st>>>(x, y, batch, T, out);
}

void launch_external_product1024(uint32_t* acc, cplx* digf, uint32_t* t, const cplx* key, const FFTTables& T,
                                 hipStream_t st) {
    k_plain_decompose_fft<<<GATE_ROWS, 64, 0, st>>>(acc, T, digf);
    k_plain_mac_ifft<<<2, 64, 0, st>>>(digf, key, T, t);
}

uint32_t ks_bpad(uint32_t batch) { return (batch + KS_GROUP - 1) / KS_GROUP * KS_GROUP; }

unsigned device_cus() {
    static std::atomic<unsigned> cache[64];  // zero-initialised (static storage); 0 = not yet read
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    unsigned v = cache[dev].load(std::memory_order_relaxed);
    if (!v) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
        v = (unsigned)cus;
        cache[dev].store(v, std::memory_order_relaxed);
    }
    return v;
}

int ks_model_slices(uint32_t batch, int N, int n) {
    // i-slices: as many workgroups as run at once (two per CU: 100 VGPRs and 8 waves
    // each), so the launch is one full round; each slice >= 32 coefficients
    const unsigned cols = (unsigned)(n + 1 + 127) / 128, groups = ks_bpad(batch) / KS_GROUP;
    const unsigned slots = 2 * device_cus();
    unsigned slices = slots / (cols * groups);
    if (slices > (unsigned)N / 32) slices = (unsigned)N / 32;
    if (slices < 1) slices = 1;
    return (int)slices;
}

int launch_keyswitch32(const uint32_t* lweN, uint32_t batch, int N, int n, int logB, int l, const uint32_t* ksk,
                       uint32_t* at, uint32_t* out, hipStream_t st, int slices_in) {
    if (!(logB == 2 && l == 8) || N % 64 || n + 1 > 64 * 1024) return -1;
    const uint32_t Bpad = ks_bpad(batch);
    const unsigned cols = (unsigned)(n + 1 + 127) / 128, groups = Bpad / KS_GROUP;
    unsigned slices = slices_in > 0 ? (unsigned)slices_in : (unsigned)ks_model_slices(batch, N, n);
    if (slices > (unsigned)N) slices = (unsigned)N;
    const size_t count = (size_t)batch * (n + 1);
    k_fill_zero<<<(unsigned)((count + 255) / 256 < 4096 ? (count + 255) / 256 : 4096), 256, 0, st>>>(out, count);
    k_ks_digits<16><<<dim3((unsigned)N / 64, Bpad / 64), 256, 0, st>>>(lweN, batch, N, Bpad, at);
    k_keyswitch32<2, 8><<<dim3(cols * slices, groups), 64 * KS_WAVES, 0, st>>>(at, batch, Bpad, N, n, lweN, ksk,
                                                                               (int)cols, (int)slices, out);
    return 0;
}

}  // namespace tfhe
