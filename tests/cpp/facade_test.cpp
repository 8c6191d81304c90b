// Drives libtfhe_amd.so through its C++ interface (include/tfhe_amd.hpp) the way the
// reference's AES example drives its GPU contexts (bitwise/examples/aes_transcipher/main.cu):
// key generation, single and batched gates against their truth tables, circuit
// bootstraps from device LWEs in the reference's argument layout, and the AES-128 CTR
// keystream against FIPS-197 C.1 (the example's self-check, main.cu:1627-1630), and
// transciphering of public AES-CTR ciphertext back to LWE plaintext bits.
// Exit status 0 and a last line "facade ok" when every check passes.
#include <cstdio>
#include <cstring>
#include <random>

#include "tfhe_amd.hpp"

using namespace tfhe_amd;

#define EXPECT(cond, ...)                          \
    do {                                           \
        if (!(cond)) {                             \
            std::fprintf(stderr, "FAIL: ");        \
            std::fprintf(stderr, __VA_ARGS__);     \
            std::fprintf(stderr, "\n");            \
            return 1;                              \
        }                                          \
    } while (0)

template <class F>
static int expect_throw(F&& f, int code, const char* what) {
    try {
        f();
    } catch (const Error& e) {
        if (e.code != code) {
            std::fprintf(stderr, "FAIL: %s threw code %d (%s), expected %d\n", what, e.code, e.what(), code);
            return 1;
        }
        return 0;
    }
    std::fprintf(stderr, "FAIL: %s did not throw\n", what);
    return 1;
}

static int gates(const std::vector<int32_t>& sk) {
    GPUBinFHEContext cc(GateParams(), 64);
    cc.GPUBTKeyGen(sk, 7);
    struct G {
        int gate;
        int f[4];
        const char* name;
    } tabs[] = {{TFHE_AND, {0, 0, 0, 1}, "AND"},   {TFHE_OR, {0, 1, 1, 1}, "OR"},
                {TFHE_NAND, {1, 1, 1, 0}, "NAND"}, {TFHE_NOR, {1, 0, 0, 0}, "NOR"},
                {TFHE_XOR, {0, 1, 1, 0}, "XOR"},   {TFHE_XNOR, {1, 0, 0, 1}, "XNOR"}};
    uint32_t id = 0;
    for (const G& g : tabs)
        for (int x = 0; x < 4; ++x) {
            const LWECiphertext c1 = LWEEncrypt(sk, 9, id++, (uint32_t)(x >> 1));
            const LWECiphertext c2 = LWEEncrypt(sk, 9, id++, (uint32_t)(x & 1));
            const LWECiphertext out = cc.GPUEvalBinGate(g.gate, c1, c2);
            EXPECT(LWEDecrypt(sk, out) == (uint32_t)g.f[x], "GPUEvalBinGate %s(%d, %d)", g.name, x >> 1, x & 1);
        }
    std::printf("GPUEvalBinGate: 6 gates x 4 inputs ok\n");

    std::mt19937 rng(5);
    std::vector<LWECiphertext> v1, v2;
    std::vector<int> a, b;
    for (int k = 0; k < 64; ++k) {
        a.push_back((int)(rng() & 1));
        b.push_back((int)(rng() & 1));
        v1.push_back(LWEEncrypt(sk, 11, 2 * k, (uint32_t)a.back()));
        v2.push_back(LWEEncrypt(sk, 11, 2 * k + 1, (uint32_t)b.back()));
    }
    for (int gate : {TFHE_NAND, TFHE_XOR}) {
        const std::vector<LWECiphertext> out = cc.BatchGPUEvalBinGate(gate, v1, v2);
        EXPECT(out.size() == 64, "BatchGPUEvalBinGate size");
        for (int k = 0; k < 64; ++k) {
            const int want = gate == TFHE_NAND ? 1 - (a[k] & b[k]) : a[k] ^ b[k];
            EXPECT(LWEDecrypt(sk, out[k]) == (uint32_t)want, "BatchGPUEvalBinGate gate %d index %d", gate, k);
        }
    }
    std::printf("BatchGPUEvalBinGate: 2 x 64 ok\n");

    if (expect_throw([&] { (void)cc.GPUEvalBinGate(TFHE_AND, v1[0], v1[0]); }, TFHE_E_PARAM, "same ciphertext twice"))
        return 1;
    std::vector<LWECiphertext> big(65, v1[0]);
    if (expect_throw([&] { (void)cc.BatchGPUEvalBinGate(TFHE_AND, big, big); }, TFHE_E_CAPACITY, "65 > 64"))
        return 1;
    if (expect_throw([&] { (void)cc.GPUEvalBinGate(99, v1[0], v2[0]); }, TFHE_E_PARAM, "unknown gate")) return 1;
    std::printf("gate errors ok\n");
    return 0;
}

static int circuit_bootstrap_and_aes(const std::vector<int32_t>& sk) {
    GPUCirBTSContext cbs(CirBTSParams(), 256);  // two AES blocks
    cbs.KeyGen(sk, 7);
    const int n = cbs.params().n, N = cbs.params().N;

    // the reference's device-LWE layout: u64 masks a_stride apart, u64 bodies
    const uint32_t batch = 3, stride = (uint32_t)n + 3;
    std::vector<LWECiphertext> cts;
    for (uint32_t k = 0; k < batch; ++k) cts.push_back(LWEEncrypt(sk, 13, k, k & 1, 2));
    std::vector<uint64_t> ha((size_t)batch * stride, 0), hb(batch);
    for (uint32_t k = 0; k < batch; ++k) {
        for (int i = 0; i < n; ++i) ha[(size_t)k * stride + i] = cts[k].a[i];
        hb[k] = cts[k].b;
    }
    DeviceBuffer<uint64_t> da(ha.size()), db(hb.size());
    da.upload(ha.data(), ha.size(), hipStreamPerThread);
    db.upload(hb.data(), hb.size(), hipStreamPerThread);
    const auto view =
        cbs.gpu_CircuitBootstrappingBatchToDeviceFromDeviceLWE(da.get(), db.get(), batch, stride, hipStreamPerThread);
    EXPECT(view.batch == batch && view.digitsCC == 2 && view.N == (uint32_t)N &&
               view.per_ciphertext_elems == (size_t)2 * 4 * N,
           "view fields");
    std::vector<uint32_t> rows(batch * view.per_ciphertext_elems);
    hip_check(hipMemcpyAsync(rows.data(), view.d_rgsw, rows.size() * sizeof(uint32_t), hipMemcpyDeviceToHost,
                             view.stream),
              "view download");
    hip_check(hipStreamSynchronize(view.stream), "view sync");
    const auto host = cbs.gpu_CircuitBootstrappingBatch(cts);
    for (uint32_t k = 0; k < batch; ++k)
        EXPECT(std::memcmp(host[k].data(), rows.data() + k * view.per_ciphertext_elems,
                           view.per_ciphertext_elems * sizeof(uint32_t)) == 0,
               "FromDeviceLWE rows differ from gpu_CircuitBootstrappingBatch, ciphertext %u", k);
    EXPECT(cbs.gpu_CircuitBootstrapping(cts[1]) == host[1], "gpu_CircuitBootstrapping vs batch");
    const auto empty = cbs.gpu_CircuitBootstrappingBatchToDeviceFromDeviceLWE(nullptr, db.get(), batch, stride);
    EXPECT(empty.batch == 0 && empty.d_rgsw == nullptr, "empty view for NULL masks");
    if (expect_throw([&] { (void)cbs.gpu_CircuitBootstrappingBatchToDeviceFromDeviceLWE(da.get(), db.get(), batch, 5); },
                     TFHE_E_PARAM, "a_stride < n"))
        return 1;
    std::vector<LWECiphertext> many(257, cts[0]);
    if (expect_throw([&] { (void)cbs.gpu_CircuitBootstrappingBatch(many); }, TFHE_E_CAPACITY, "257 > 256")) return 1;
    std::printf("circuit bootstrap: FromDeviceLWE view == host batch == single, errors ok\n");

    // gpu_CircuitBootstrappingBatchToDevice: host ciphertexts in, a view of the context's rows
    const auto hview = cbs.gpu_CircuitBootstrappingBatchToDevice(cts, hipStreamPerThread);
    EXPECT(hview.batch == batch && hview.per_ciphertext_elems == view.per_ciphertext_elems, "ToDevice view fields");
    hip_check(hipMemcpyAsync(rows.data(), hview.d_rgsw, rows.size() * sizeof(uint32_t), hipMemcpyDeviceToHost,
                             hview.stream),
              "view download");
    hip_check(hipStreamSynchronize(hview.stream), "view sync");
    for (uint32_t k = 0; k < batch; ++k)
        EXPECT(std::memcmp(host[k].data(), rows.data() + k * hview.per_ciphertext_elems,
                           hview.per_ciphertext_elems * sizeof(uint32_t)) == 0,
               "ToDevice rows differ from gpu_CircuitBootstrappingBatch, ciphertext %u", k);
    EXPECT(cbs.gpu_CircuitBootstrappingBatchToDevice({}).batch == 0, "empty ToDevice view");

    // gpu_BootstrapManyLUT on a staircase polynomial LUT[c] = c 2^50: coefficient 0 of the
    // output decrypts to LUT[phase~] (negated past N) + Gpow_0 / 2 = 2^47
    {
        std::vector<uint64_t> lut(N);
        for (int c = 0; c < N; ++c) lut[c] = (uint64_t)c << 50;
        const uint32_t bw = 1;
        const LWECiphertext ct = LWEEncrypt(sk, 21, 5, 3, 8);
        const DeviceBuffer<uint64_t> d_rl = cbs.gpu_BootstrapManyLUT(ct, lut, bw);
        std::vector<uint64_t> rl((size_t)2 * N);
        d_rl.download(rl.data(), rl.size(), hipStreamPerThread);
        std::vector<int32_t> K2(N);
        check(tfhe_cbs_get_rlwe_sk(cbs.handle(), K2.data()));
        uint64_t ph = rl[N];  // b[0] - (a K2)[0], (a K2)[0] = a_0 K2_0 - sum_{i>0} a_i K2_{N-i}
        ph -= rl[0] * (uint64_t)(int64_t)K2[0];
        for (int i = 1; i < N; ++i) ph += rl[i] * (uint64_t)(int64_t)K2[N - i];
        auto ms = [&](uint32_t v) { return (uint32_t)((((uint64_t)v + (1ull << (19 + bw))) >> (20 + bw)) << bw) & 4095u; };
        uint32_t phi = ms(ct.b);
        for (int i = 0; i < n; ++i) phi = (phi - ms(ct.a[i]) * (uint32_t)sk[i]) & 4095u;
        const uint64_t want = (phi < (uint32_t)N ? lut[phi] : 0 - lut[phi - N]) + (1ull << 47);
        const int64_t err = (int64_t)(ph - want);
        EXPECT(err < (1ll << 40) && err > -(1ll << 40), "gpu_BootstrapManyLUT phase off by %lld", (long long)err);
        std::vector<uint64_t> short_lut(N - 1);
        if (expect_throw([&] { (void)cbs.gpu_BootstrapManyLUT(ct, short_lut, bw); }, TFHE_E_SHAPE, "LUT size"))
            return 1;
    }
    std::printf("gpu_CircuitBootstrappingBatchToDevice == host batch; gpu_BootstrapManyLUT decrypts to LUT[phase]\n");

    // AES-128 CTR keystream of the FIPS-197 C.1 block
    std::array<uint8_t, 16> key{}, pt{};
    for (int i = 0; i < 16; ++i) key[i] = (uint8_t)i, pt[i] = (uint8_t)(0x11 * i);
    cbs.SetRoundKeys(sk, 77, key);
    const uint8_t want[16] = {0x69, 0xc4, 0xe0, 0xd8, 0x6a, 0x7b, 0x04, 0x30,
                              0xd8, 0xcd, 0xb7, 0x80, 0x70, 0xb4, 0xc5, 0x5a};
    for (bool graph : {false, true}) {
        const auto ks = cbs.AesCtrKeystream({pt}, graph);
        uint8_t got[16] = {0};
        for (int bit = 0; bit < 128; ++bit) got[bit / 8] |= (uint8_t)(LWEDecrypt(sk, ks[0][bit], 2) << (bit % 8));
        EXPECT(std::memcmp(got, want, 16) == 0, "AES-128 FIPS-197 C.1 through the homomorphic keystream (graph=%d)",
               (int)graph);
    }
    std::printf("AES-128 CTR keystream: FIPS-197 C.1 ok (eager and graph)\n");

    // transciphering: public AES-CTR ciphertext of a random plaintext back to LWE plaintext bits
    std::mt19937 rng(17);
    std::vector<std::array<uint8_t, 16>> ctr(2), plain(2), cipher(2);
    for (int k = 0; k < 2; ++k) {
        for (int j = 0; j < 16; ++j) {
            ctr[k][j] = (uint8_t)(j < 8 ? 0xA5 ^ j : (j == 15 ? k : 0));
            plain[k][j] = (uint8_t)(rng() & 0xFF);
        }
        uint8_t ks[16];
        tfhe_aes128_encrypt_plain(key.data(), ctr[k].data(), ks);
        for (int j = 0; j < 16; ++j) cipher[k][j] = (uint8_t)(plain[k][j] ^ ks[j]);
    }
    const auto bits = cbs.Transcipher(ctr, cipher);
    for (int k = 0; k < 2; ++k) {
        uint8_t got[16] = {0};
        for (int bit = 0; bit < 128; ++bit) got[bit / 8] |= (uint8_t)(LWEDecrypt(sk, bits[k][bit], 2) << (bit % 8));
        EXPECT(std::memcmp(got, plain[k].data(), 16) == 0, "Transcipher block %d", k);
    }
    std::printf("AES-128 CTR transcipher: 2 random blocks ok\n");

    // the bench's fastest form: one-product PBS with one or two key bits per step chosen by
    // launch size, linear-key PFKS; 5 blocks (an uneven batch: a group of 4 and one of 1),
    // the FIPS-197 block first, the others against the library's plain AES
    GPUCirBTSContext fast(CirBTSParams(1, 3, 1), 5 * 128);
    fast.KeyGen(sk, 7);
    fast.SetRoundKeys(sk, 77, key);
    std::vector<std::array<uint8_t, 16>> ctr5(5);
    ctr5[0] = pt;
    for (int k = 1; k < 5; ++k)
        for (int j = 0; j < 16; ++j) ctr5[k][j] = (uint8_t)(rng() & 0xFF);
    const auto ks5 = fast.AesCtrKeystream(ctr5);
    for (int k = 0; k < 5; ++k) {
        uint8_t got[16] = {0}, plain_ks[16];
        for (int bit = 0; bit < 128; ++bit) got[bit / 8] |= (uint8_t)(LWEDecrypt(sk, ks5[k][bit], 2) << (bit % 8));
        tfhe_aes128_encrypt_plain(key.data(), ctr5[k].data(), plain_ks);
        EXPECT(std::memcmp(got, plain_ks, 16) == 0, "AES-128 keystream block %d, one-product PBS (auto key bits) + "
               "linear PFKS", k);
        if (k == 0) EXPECT(std::memcmp(got, want, 16) == 0, "AES-128 FIPS-197 C.1, one-product PBS + linear PFKS");
    }
    std::printf("AES-128 CTR keystream, one-product PBS (auto key bits) + linear PFKS: 5 blocks ok (FIPS-197 C.1)\n");
    return 0;
}

// EvalFunc, EvalFloor, EvalSign and EvalDecomp through the facade.  Inputs over Z_2^12 keep
// (v + beta) mod q_small near q/4 or 3q/4, far from the floor's branch points 0 and q/2 against
// the bootstrap noise (about 15 units at 2^12); the floor is checked to 6 sigma, the digits to
// 6 sigma of the switch to Z_512 (its rounding adds about 5 units there, 40 at 2^12).
static int functions(const std::vector<int32_t>& sk) {
    GPUBinFHEContext cc(GateParams(), 64);
    cc.GPUBTKeyGen(sk, 7);
    const uint32_t q = TFHE_FUNC_Q, beta = TFHE_FUNC_BETA, p = q / (2 * beta);
    const int F[8] = {1, 4, 7, 2, 5, 0, 3, 6};  // 3m + 1 mod 8: neither negacyclic nor periodic
    std::vector<uint32_t> lut(q);
    for (uint32_t x = 0; x < q; ++x) lut[x] = (uint32_t)F[x * p / q] * (q / p);
    EXPECT(tfhe_func_property(lut.data(), q) == TFHE_FUNC_ARBITRARY, "table class");
    std::vector<LWECiphertext> in;
    for (uint32_t m = 0; m < p; ++m) in.push_back(LWEEncryptMod(sk, 21, m, m * (q / p), q));
    const std::vector<LWECiphertext> fo = cc.BatchGPUEvalFunc(in, lut);
    for (uint32_t m = 0; m < p; ++m) {
        EXPECT(fo[m].modulus == q, "EvalFunc output modulus");
        const uint64_t ph = LWEPhase(sk, fo[m]);
        EXPECT(((ph + q / (2 * p)) / (q / p)) % p == (uint64_t)F[m], "BatchGPUEvalFunc m=%u", m);
    }
    {
        const LWECiphertext one = cc.GPUEvalFunc(in[3], lut);
        EXPECT(one.a == fo[3].a && one.b == fo[3].b, "GPUEvalFunc equals the batch form");
    }
    std::printf("BatchGPUEvalFunc: 8 messages of Z_8 ok\n");

    const uint64_t mod = 1u << 12;
    std::vector<LWECiphertext> v;
    std::vector<uint64_t> val;
    for (uint32_t k = 0; k < 8; ++k) {
        const uint64_t y = (k & 1) ? 3 * q / 4 : q / 4;  // (v + beta) mod q
        val.push_back(((k >> 1) * q + y - beta + 17 * k) % mod);
        v.push_back(LWEEncryptMod(sk, 23, k, val.back(), mod));
    }
    const auto fl = cc.BatchGPUEvalFloor(v);
    const auto sg = cc.BatchGPUEvalSign(v);
    const auto dc = cc.BatchGPUEvalDecomp(v);
    for (uint32_t k = 0; k < 8; ++k) {
        const uint64_t want = ((val[k] + beta) / q * q) % mod;
        const int64_t e = (int64_t)((LWEPhase(sk, fl[k]) - want + mod / 2) % mod) - (int64_t)(mod / 2);
        EXPECT(e > -128 && e < 128, "BatchGPUEvalFloor k=%u: error %lld", k, (long long)e);
        EXPECT(sg[k].modulus == q, "EvalSign output modulus");
        const uint64_t sph = LWEPhase(sk, sg[k]);
        EXPECT(((sph + q / 4) / (q / 2)) % 2 == val[k] / (mod / 2), "BatchGPUEvalSign k=%u", k);
        EXPECT(dc[k].size() == 2 && dc[k][0].modulus == q && dc[k][1].modulus == 512, "EvalDecomp digits");
        const uint64_t x = (8 * LWEPhase(sk, dc[k][1]) + (LWEPhase(sk, dc[k][0]) + beta) % q - beta) % mod;
        const int64_t d = (int64_t)((x - val[k] + mod / 2) % mod) - (int64_t)(mod / 2);
        EXPECT(d > -256 && d < 256, "BatchGPUEvalDecomp k=%u: error %lld", k, (long long)d);
    }
    std::printf("BatchGPUEvalFloor / Sign / Decomp over Z_2^12: 8 ciphertexts ok\n");
    if (expect_throw([&] { (void)cc.GPUEvalSign(LWEEncryptMod(sk, 1, 1, 1, q, 15)); }, TFHE_E_PARAM,
                     "EvalSign of a small-precision ciphertext"))
        return 1;
    return 0;
}

int main() {
    try {
        const std::vector<int32_t> sk = LWEKeyGen(636, 3);
        if (gates(sk)) return 1;
        if (functions(sk)) return 1;
        if (circuit_bootstrap_and_aes(sk)) return 1;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "FAIL: unexpected exception: %s\n", e.what());
        return 1;
    }
    std::printf("facade ok\n");
    return 0;
}
