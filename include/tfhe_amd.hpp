/* tfhe_amd.hpp -- C++ host interface over the C ABI of libtfhe_amd.so.
 *
 * Keeps the reference's class and method names so that its callers change types,
 * not call sites:
 *   tfhe_amd::GPUBinFHEContext  for phantom::bitwise::GPUBinFHEContext
 *                               (bitwise/include/binfhecontext.cuh:8-119)
 *   tfhe_amd::GPUCirBTSContext  for GPUCirBTSContext
 *                               (bitwise/include/cirbts/cirbts.cuh:31-498)
 * Errors throw tfhe_amd::Error, a std::runtime_error that carries the C ABI status,
 * where the reference throws OPENFHE_THROW(config_error) or PHANTOM_CHECK_CUDA's
 * runtime_error; nothing calls exit().  Host ciphertexts are LWECiphertext (n mask
 * words and the body on the 32-bit torus) in place of lbcrypto::LWECiphertext, and
 * the secret is a vector of n binary coefficients in place of LWEPrivateKey.
 * Methods that take host ciphertexts stage them through device buffers of their own;
 * the *Device / *FromDeviceLWE methods work on caller device memory.  One context
 * per host thread, as in the reference.
 *
 * Header-only: include it, link libtfhe_amd.so and the HIP runtime (amdhip64). */
#ifndef TFHE_AMD_HPP
#define TFHE_AMD_HPP

#include <hip/hip_runtime_api.h>

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "tfhe_amd.h"

namespace tfhe_amd {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& msg) : std::runtime_error(msg), code(c) {}
};

inline void check(int rc) {
    if (rc != TFHE_OK) throw Error(rc, tfhe_last_error());
}

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Error(TFHE_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

/* Device allocation owned by the facade (freed on destruction). */
template <class T>
class DeviceBuffer {
    T* p_ = nullptr;
    size_t n_ = 0;

public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t n) : n_(n) {
        if (n) hip_check(hipMalloc(reinterpret_cast<void**>(&p_), n * sizeof(T)), "hipMalloc");
    }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    DeviceBuffer(DeviceBuffer&& o) noexcept : p_(std::exchange(o.p_, nullptr)), n_(std::exchange(o.n_, 0)) {}
    ~DeviceBuffer() {
        if (p_) (void)hipFree(p_);
    }
    T* get() const { return p_; }
    size_t size() const { return n_; }
    void upload(const T* h, size_t n, hipStream_t s) {
        hip_check(hipMemcpyAsync(p_, h, n * sizeof(T), hipMemcpyHostToDevice, s), "upload");
    }
    void download(T* h, size_t n, hipStream_t s) const {
        hip_check(hipMemcpyAsync(h, p_, n * sizeof(T), hipMemcpyDeviceToHost, s), "download");
        hip_check(hipStreamSynchronize(s), "download sync");
    }
};

/* An LWE ciphertext: a (n words) and b over Z_modulus (a power of two; 2^32, the torus,
 * for gate and circuit-bootstrapping inputs), as lbcrypto::LWECiphertext carries GetModulus(). */
struct LWECiphertext {
    std::vector<uint32_t> a;
    uint32_t b = 0;
    uint64_t modulus = 1ull << 32;
};

/* The build's parameter sets: gate bootstrapping n=636, N=1024, BSK 2^7 x 3,
 * KS 2^2 x 8; circuit bootstrapping N=2048, PBS 2^9 x 4, CB 2^8 x 2, PFKS 2^4 x 8. */
inline tfhe_gate_params GateParams() { return tfhe_gate_params{636, 1024, 7, 3, 2, 8, 15, 25}; }
inline tfhe_cbs_params CirBTSParams(int pbs_approx = 0, int pbs_multibit = 1, int pfks_linear = 0,
                                    int lwe_to_rgsw = 0, int ks_linear = 0) {
    if (lwe_to_rgsw)  // HomTrace + SchemeSwitch: exact 2^10 x 4 PBS, keys at sigma 2^-54
        return tfhe_cbs_params{636, 2048, 10, 4, 8, 2, 4, 8, 2, 8, 15, 22, 30, 0, pbs_multibit > 1 ? 1 : pbs_multibit,
                               0, 1, ks_linear};
    return tfhe_cbs_params{636, 2048, 9, 4, 8, 2, 4, 8, 2, 8, 15, 19, 30, pbs_approx, pbs_multibit, pfks_linear, 0,
                           ks_linear};
}

/* Client-side helpers standing in for the OpenFHE fork's KeyGen / Encrypt / Decrypt. */
inline std::vector<int32_t> LWEKeyGen(int n, uint64_t seed) {
    std::vector<int32_t> sk(n);
    tfhe_lwe_keygen(n, seed, sk.data());
    return sk;
}
/* m in Z_p encoded as m * 2^32 / p (p = 4 for gate inputs, 2 for AES state bits) */
inline LWECiphertext LWEEncrypt(const std::vector<int32_t>& sk, uint64_t seed, uint32_t id, uint32_t m, uint32_t p = 4,
                                int nshift = 15) {
    std::vector<uint32_t> ct(sk.size() + 1);
    tfhe_lwe_encrypt((int)sk.size(), sk.data(), seed, id, m, p, nshift, ct.data());
    LWECiphertext c;
    c.b = ct.back();
    ct.pop_back();
    c.a = std::move(ct);
    return c;
}
inline uint32_t LWEDecrypt(const std::vector<int32_t>& sk, const LWECiphertext& c, uint32_t p = 4) {
    std::vector<uint32_t> ct(c.a);
    ct.push_back(c.b);
    return tfhe_lwe_decrypt((int)sk.size(), sk.data(), ct.data(), p);
}

/* m over Z_mod (a power of two <= 2^32): the torus encryption of m 2^32 / mod, every element
 * rounded to Z_mod (the inputs of the programmable bootstrap and the large-precision functions). */
inline LWECiphertext LWEEncryptMod(const std::vector<int32_t>& sk, uint64_t seed, uint32_t id, uint64_t m,
                                   uint64_t mod, int nshift = 15) {
    LWECiphertext c = LWEEncrypt(sk, seed, id, 0, 2, nshift);
    const int sh = [&] { int l = 0; while ((1ull << l) < mod) ++l; return 32 - l; }();
    c.b += (uint32_t)((m % mod) << sh);
    if (sh) {
        const uint64_t half = 1ull << (sh - 1), mask = mod - 1;
        for (auto& x : c.a) x = (uint32_t)((((uint64_t)x + half) >> sh) & mask);
        c.b = (uint32_t)((((uint64_t)c.b + half) >> sh) & mask);
    }
    c.modulus = mod;
    return c;
}
/* b - <a, s> over Z_modulus */
inline uint64_t LWEPhase(const std::vector<int32_t>& sk, const LWECiphertext& c) {
    uint64_t ph = c.b;
    for (size_t i = 0; i < c.a.size(); ++i) ph -= (uint64_t)c.a[i] * (uint64_t)(uint32_t)sk[i];
    return ph & (c.modulus - 1);
}

namespace detail {
inline std::vector<uint32_t> pack(const std::vector<LWECiphertext>& v, size_t n) {
    std::vector<uint32_t> out(v.size() * (n + 1));
    for (size_t k = 0; k < v.size(); ++k) {
        if (v[k].a.size() != n) throw Error(TFHE_E_SHAPE, "LWE ciphertext of the wrong dimension");
        for (size_t i = 0; i < n; ++i) out[k * (n + 1) + i] = v[k].a[i];
        out[k * (n + 1) + n] = v[k].b;
    }
    return out;
}
inline std::vector<LWECiphertext> unpack(const std::vector<uint32_t>& w, size_t count, size_t n) {
    std::vector<LWECiphertext> out(count);
    for (size_t k = 0; k < count; ++k) {
        out[k].a.assign(w.begin() + k * (n + 1), w.begin() + k * (n + 1) + n);
        out[k].b = w[k * (n + 1) + n];
    }
    return out;
}
}  // namespace detail

/* Gate bootstrapping (phantom::bitwise::GPUBinFHEContext). */
class GPUBinFHEContext {
    tfhe_gate_ctx* h_ = nullptr;
    tfhe_gate_params p_;
    uint32_t max_batch_;

public:
    /* GPUBinFHEContext(cc): the parameter set replaces the BinFHEContext argument;
     * max_batch bounds BatchGPUEvalBinGate. */
    explicit GPUBinFHEContext(const tfhe_gate_params& p = GateParams(), uint32_t max_batch = 4096)
        : p_(p), max_batch_(max_batch) {
        check(tfhe_gate_ctx_create(&p_, max_batch, &h_));
    }
    GPUBinFHEContext(const GPUBinFHEContext&) = delete;
    GPUBinFHEContext& operator=(const GPUBinFHEContext&) = delete;
    ~GPUBinFHEContext() { tfhe_gate_ctx_destroy(h_); }

    tfhe_gate_ctx* handle() const { return h_; }
    const tfhe_gate_params& params() const { return p_; }

    /* GPUBTKeyGen(sk, SYM_ENCRYPT) (binfhecontext.cuh:31-37); the seed replaces the
     * fork's PRNG state.  Public-key mode does not exist here, as it throws there. */
    void GPUBTKeyGen(const std::vector<int32_t>& lwe_sk, uint64_t seed) {
        if ((int)lwe_sk.size() != p_.n) throw Error(TFHE_E_SHAPE, "GPUBTKeyGen: secret of the wrong dimension");
        check(tfhe_gate_bt_keygen(h_, lwe_sk.data(), seed));
    }

    /* GPUEvalBinGate(gate, ct1, ct2, stream) (binfhecontext.cuh:48-51): the reference's
     * per-step launch structure for one gate; the same ciphertext twice throws, as in
     * binfhe-base-scheme.cu:158-159. */
    LWECiphertext GPUEvalBinGate(int gate, const LWECiphertext& ct1, const LWECiphertext& ct2,
                                 hipStream_t s = hipStreamPerThread) const {
        if (&ct1 == &ct2) throw Error(TFHE_E_PARAM, "GPUEvalBinGate: input ciphertexts should be independent");
        const size_t w = (size_t)p_.n + 1;
        std::vector<uint32_t> h = detail::pack({ct1, ct2}, p_.n);
        DeviceBuffer<uint32_t> d(3 * w);
        d.upload(h.data(), 2 * w, s);
        check(tfhe_gate_eval_plain(h_, gate, d.get(), d.get() + w, d.get() + 2 * w, s));
        std::vector<uint32_t> out(w);
        hip_check(hipMemcpyAsync(out.data(), d.get() + 2 * w, w * sizeof(uint32_t), hipMemcpyDeviceToHost, s), "download");
        hip_check(hipStreamSynchronize(s), "download sync");
        return detail::unpack(out, 1, p_.n)[0];
    }

    /* BatchGPUEvalBinGate(gate, v1, v2, stream) (binfhecontext.cuh:53-57): one
     * persistent blind-rotation launch and one batched key switch. */
    std::vector<LWECiphertext> BatchGPUEvalBinGate(int gate, const std::vector<LWECiphertext>& v1,
                                                   const std::vector<LWECiphertext>& v2,
                                                   hipStream_t s = hipStreamPerThread) const {
        if (v1.size() != v2.size()) throw Error(TFHE_E_SHAPE, "BatchGPUEvalBinGate: input vectors differ in size");
        if (v1.empty()) return {};
        if (v1.size() > max_batch_) throw Error(TFHE_E_CAPACITY, "BatchGPUEvalBinGate: batch exceeds the context");
        const size_t w = (size_t)p_.n + 1, B = v1.size();
        std::vector<uint32_t> h1 = detail::pack(v1, p_.n), h2 = detail::pack(v2, p_.n);
        DeviceBuffer<uint32_t> d1(B * w), d2(B * w), dout(B * w);
        d1.upload(h1.data(), B * w, s);
        d2.upload(h2.data(), B * w, s);
        BatchGPUEvalBinGateDevice(gate, d1.get(), d2.get(), (uint32_t)B, dout.get(), s);
        std::vector<uint32_t> out(B * w);
        dout.download(out.data(), B * w, s);
        return detail::unpack(out, B, p_.n);
    }

    /* Device form: [batch][n+1] uint32 ciphertexts in and out. */
    void BatchGPUEvalBinGateDevice(int gate, const uint32_t* d_ct1, const uint32_t* d_ct2, uint32_t batch,
                                   uint32_t* d_out, hipStream_t s = hipStreamPerThread) const {
        check(tfhe_gate_eval_batch(h_, gate, d_ct1, d_ct2, batch, d_out, s));
    }

    /* GPUEvalFunc(ct, LUT, stream) / BatchGPUEvalFunc(v_ct, LUT, stream) (binfhecontext.cuh:66-72):
     * the table has one entry per element of Z_q, q the ciphertexts' modulus. */
    LWECiphertext GPUEvalFunc(const LWECiphertext& ct, const std::vector<uint32_t>& LUT,
                              hipStream_t s = hipStreamPerThread) const {
        return BatchGPUEvalFunc({ct}, LUT, s)[0];
    }
    std::vector<LWECiphertext> BatchGPUEvalFunc(const std::vector<LWECiphertext>& v_ct,
                                                const std::vector<uint32_t>& LUT,
                                                hipStream_t s = hipStreamPerThread) const {
        const uint64_t q = common_modulus(v_ct, "BatchGPUEvalFunc");
        if (LUT.size() != q) throw Error(TFHE_E_SHAPE, "BatchGPUEvalFunc: the table needs one entry per element of Z_q");
        return run(v_ct, 1, q, s, [&](const uint32_t* in, uint32_t B, uint32_t* out) {
            check(tfhe_gate_eval_func_batch(h_, in, B, q, LUT.data(), out, s));
        })[0];
    }

    /* GPUEvalFloor(ct, roundbits, stream) / BatchGPUEvalFloor (binfhecontext.cuh:82-88). */
    LWECiphertext GPUEvalFloor(const LWECiphertext& ct, uint32_t roundbits = 0,
                               hipStream_t s = hipStreamPerThread) const {
        return BatchGPUEvalFloor({ct}, roundbits, s)[0];
    }
    std::vector<LWECiphertext> BatchGPUEvalFloor(const std::vector<LWECiphertext>& v_ct, uint32_t roundbits = 0,
                                                 hipStream_t s = hipStreamPerThread) const {
        const uint64_t mod = common_modulus(v_ct, "BatchGPUEvalFloor");
        return run(v_ct, 1, mod, s, [&](const uint32_t* in, uint32_t B, uint32_t* out) {
            check(tfhe_gate_eval_floor_batch(h_, in, B, mod, roundbits, out, s));
        })[0];
    }

    /* GPUEvalSign(ct, schemeSwitch, stream) / BatchGPUEvalSign (binfhecontext.cuh:98-104): the
     * outputs are over Z_TFHE_FUNC_Q. */
    LWECiphertext GPUEvalSign(const LWECiphertext& ct, bool schemeSwitch = false,
                              hipStream_t s = hipStreamPerThread) const {
        return BatchGPUEvalSign({ct}, schemeSwitch, s)[0];
    }
    std::vector<LWECiphertext> BatchGPUEvalSign(const std::vector<LWECiphertext>& v_ct, bool schemeSwitch = false,
                                                hipStream_t s = hipStreamPerThread) const {
        const uint64_t mod = common_modulus(v_ct, "BatchGPUEvalSign");
        return run(v_ct, 1, TFHE_FUNC_Q, s, [&](const uint32_t* in, uint32_t B, uint32_t* out) {
            check(tfhe_gate_eval_sign_batch(h_, in, B, mod, schemeSwitch ? 1 : 0, out, s));
        })[0];
    }

    /* GPUEvalDecomp(ct, stream) / BatchGPUEvalDecomp(v_ct, stream) (binfhecontext.cuh:113-118): the
     * digits from the least significant up; the batch form returns them per ciphertext, as the
     * reference's vector of vectors (binfhe-base-scheme-batch.cu:665-672). */
    std::vector<LWECiphertext> GPUEvalDecomp(const LWECiphertext& ct, hipStream_t s = hipStreamPerThread) const {
        return BatchGPUEvalDecomp({ct}, s)[0];
    }
    std::vector<std::vector<LWECiphertext>> BatchGPUEvalDecomp(const std::vector<LWECiphertext>& v_ct,
                                                               hipStream_t s = hipStreamPerThread) const {
        const uint64_t mod = common_modulus(v_ct, "BatchGPUEvalDecomp");
        constexpr uint32_t kMax = 16;  // 2^32 takes 9
        uint32_t nd = 0;
        uint64_t mods[kMax] = {};
        auto digits = run(v_ct, kMax, 0, s, [&](const uint32_t* in, uint32_t B, uint32_t* out) {
            check(tfhe_gate_eval_decomp_batch(h_, in, B, mod, kMax, out, &nd, mods, s));
        });
        std::vector<std::vector<LWECiphertext>> ret(v_ct.size(), std::vector<LWECiphertext>(nd));
        for (size_t k = 0; k < v_ct.size(); ++k)
            for (uint32_t d = 0; d < nd; ++d) {
                ret[k][d] = std::move(digits[d][k]);
                ret[k][d].modulus = mods[d];
            }
        return ret;
    }

private:
    static uint64_t common_modulus(const std::vector<LWECiphertext>& v, const char* what) {
        if (v.empty()) throw Error(TFHE_E_SHAPE, std::string(what) + ": no ciphertexts");
        for (const auto& c : v)
            if (c.modulus != v[0].modulus) throw Error(TFHE_E_SHAPE, std::string(what) + ": ciphertexts of different moduli");
        return v[0].modulus;
    }
    /* stage v through device memory, run f(d_in, batch, d_out[digits][batch][n+1]) and return
     * the outputs per digit, each over out_mod (0: set by the caller) */
    template <class F>
    std::vector<std::vector<LWECiphertext>> run(const std::vector<LWECiphertext>& v, uint32_t digits,
                                                uint64_t out_mod, hipStream_t s, F f) const {
        if (v.size() > max_batch_) throw Error(TFHE_E_CAPACITY, "function call: batch exceeds the context");
        const size_t w = (size_t)p_.n + 1, B = v.size();
        std::vector<uint32_t> h = detail::pack(v, p_.n);
        DeviceBuffer<uint32_t> din(B * w), dout((size_t)digits * B * w);
        din.upload(h.data(), B * w, s);
        f(din.get(), (uint32_t)B, dout.get());
        std::vector<uint32_t> o((size_t)digits * B * w);
        dout.download(o.data(), o.size(), s);
        std::vector<std::vector<LWECiphertext>> ret(digits);
        for (uint32_t d = 0; d < digits; ++d) {
            std::vector<uint32_t> part(o.begin() + d * B * w, o.begin() + (d + 1) * B * w);
            ret[d] = detail::unpack(part, B, p_.n);
            for (auto& c : ret[d]) c.modulus = out_mod;
        }
        return ret;
    }
};

/* Circuit bootstrapping and AES-128 transciphering (GPUCirBTSContext and the
 * example's round driver). */
class GPUCirBTSContext {
    tfhe_cbs_ctx* h_ = nullptr;
    tfhe_cbs_params p_;
    uint32_t max_batch_;

public:
    /* The reference's non-owning view (cirbts.cuh:406-418), elements torus-32 here. */
    struct DeviceRGSWBatchView {
        const uint32_t* d_rgsw{};
        uint32_t batch{};
        uint32_t digitsCC{};
        uint32_t N{};
        size_t per_ciphertext_elems{};
        hipStream_t stream{};
    };
    /* One RGSW on the host: [2 digitsCC rows][2 polys][N] uint32. */
    using RGSWCiphertext = std::vector<uint32_t>;

    /* GPUCirBTSContext(cc, maxBatchSize) (cirbts.cuh:420-426): maxBatchSize circuit
     * bootstraps per call (rounded up to whole AES blocks of 128). */
    explicit GPUCirBTSContext(const tfhe_cbs_params& p = CirBTSParams(), uint32_t maxBatchSize = 128) : p_(p) {
        const uint32_t blocks = maxBatchSize ? (maxBatchSize + 127) / 128 : 1;
        max_batch_ = 128 * blocks;
        check(tfhe_cbs_ctx_create(&p_, blocks, &h_));
    }
    GPUCirBTSContext(const GPUCirBTSContext&) = delete;
    GPUCirBTSContext& operator=(const GPUCirBTSContext&) = delete;
    ~GPUCirBTSContext() { tfhe_cbs_ctx_destroy(h_); }

    tfhe_cbs_ctx* handle() const { return h_; }
    const tfhe_cbs_params& params() const { return p_; }
    uint32_t max_batch_size() const { return max_batch_; }

    /* Key generation on the GPU (the fork's CirBTKeyGen, main.cu:232-235). */
    void KeyGen(const std::vector<int32_t>& lwe_sk, uint64_t seed) {
        if ((int)lwe_sk.size() != p_.n) throw Error(TFHE_E_SHAPE, "KeyGen: secret of the wrong dimension");
        check(tfhe_cbs_keygen(h_, lwe_sk.data(), seed));
    }

    /* gpu_CircuitBootstrappingBatchToDeviceFromDeviceLWE(params, ek, d_a, d_b, batch,
     * a_stride) (cirbts.cuh:466-474): u64 masks a_stride apart and u64 bodies, torus-32
     * values; the view stays valid until the next call that reuses the context's RGSW
     * buffer.  NULL inputs or batch 0 give an empty view, as there. */
    DeviceRGSWBatchView gpu_CircuitBootstrappingBatchToDeviceFromDeviceLWE(const uint64_t* d_a, const uint64_t* d_b,
                                                                           uint32_t batch, uint32_t a_stride,
                                                                           hipStream_t s = hipStreamPerThread) const {
        tfhe_rgsw_batch_view v{};
        check(tfhe_cbs_batch_from_device_lwe(h_, d_a, d_b, batch, a_stride, &v, s));
        return DeviceRGSWBatchView{v.d_rgsw, v.batch, v.digits, v.N, v.per_ciphertext_elems,
                                   static_cast<hipStream_t>(v.stream)};
    }

    /* gpu_CircuitBootstrappingBatchToDevice(params, ek, cts) (cirbts.cuh:464-466): host
     * ciphertexts in, a borrowed view of the outputs held in the context (valid until the
     * next call that reuses the context's RGSW buffer).  An empty vector gives an empty view. */
    DeviceRGSWBatchView gpu_CircuitBootstrappingBatchToDevice(const std::vector<LWECiphertext>& cts,
                                                              hipStream_t s = hipStreamPerThread) const {
        if (cts.size() > max_batch_)
            throw Error(TFHE_E_CAPACITY, "gpu_CircuitBootstrappingBatchToDevice: batch size exceeds max_batch_size_");
        std::vector<uint32_t> h = detail::pack(cts, p_.n);
        tfhe_rgsw_batch_view v{};
        check(tfhe_cbs_batch_to_device(h_, cts.empty() ? nullptr : h.data(), (uint32_t)cts.size(), &v, s));
        return DeviceRGSWBatchView{v.d_rgsw, v.batch, v.digits, v.N, v.per_ciphertext_elems,
                                   static_cast<hipStream_t>(v.stream)};
    }

    /* gpu_BootstrapManyLUT(params, ek, ct, LUT, bitwidth) (cirbts.cuh:478-482,
     * cirbts-bootstrap.cu:753-839): an owned device buffer of 2N words, the RLWE_K2
     * (mask, then body) on the 64-bit torus, coefficient domain; LUT: N torus-64 words. */
    DeviceBuffer<uint64_t> gpu_BootstrapManyLUT(const LWECiphertext& ct, const std::vector<uint64_t>& LUT,
                                               uint64_t bitwidth, hipStream_t s = hipStreamPerThread) const {
        if (LUT.size() != (size_t)p_.N) throw Error(TFHE_E_SHAPE, "gpu_BootstrapManyLUT: LUT of the wrong size");
        std::vector<uint32_t> h = detail::pack({ct}, p_.n);
        DeviceBuffer<uint32_t> d(h.size());
        DeviceBuffer<uint64_t> dl(LUT.size()), out((size_t)2 * p_.N);
        d.upload(h.data(), h.size(), s);
        dl.upload(LUT.data(), LUT.size(), s);
        check(tfhe_cbs_bootstrap_many_lut(h_, d.get(), 1, dl.get(), (uint32_t)bitwidth, out.get(), s));
        hip_check(hipStreamSynchronize(s), "gpu_BootstrapManyLUT");  // d and dl are freed on return
        return out;
    }

    /* gpu_CircuitBootstrappingBatch(params, ek, cts) (cirbts.cuh:460-462): host in, host out. */
    std::vector<RGSWCiphertext> gpu_CircuitBootstrappingBatch(const std::vector<LWECiphertext>& cts,
                                                              hipStream_t s = hipStreamPerThread) const {
        if (cts.empty()) return {};
        if (cts.size() > max_batch_)
            throw Error(TFHE_E_CAPACITY, "gpu_CircuitBootstrappingBatch: batch size exceeds max_batch_size_");
        const size_t w = (size_t)p_.n + 1, B = cts.size(), rw = (size_t)2 * p_.cb_l * 2 * p_.N;
        std::vector<uint32_t> h = detail::pack(cts, p_.n);
        DeviceBuffer<uint32_t> d(B * w), dr(B * rw);
        d.upload(h.data(), B * w, s);
        check(tfhe_cbs_batch(h_, d.get(), (uint32_t)B, dr.get(), s));
        std::vector<uint32_t> all(B * rw);
        dr.download(all.data(), B * rw, s);
        std::vector<RGSWCiphertext> out(B);
        for (size_t k = 0; k < B; ++k) out[k].assign(all.begin() + k * rw, all.begin() + (k + 1) * rw);
        return out;
    }

    /* gpu_CircuitBootstrapping(params, ek, ct) (cirbts.cuh:456-458): one ciphertext. */
    RGSWCiphertext gpu_CircuitBootstrapping(const LWECiphertext& ct, hipStream_t s = hipStreamPerThread) const {
        return gpu_CircuitBootstrappingBatch({ct}, s)[0];
    }

    /* AES-128 transciphering (the example's BuildEncryptedKeyMaterial and
     * RunTranscipherBatch, key_material.cuh:12-47, main.cu:1316-1619). */
    void SetRoundKeys(const std::vector<int32_t>& lwe_sk, uint64_t seed, const std::array<uint8_t, 16>& key) {
        std::vector<uint32_t> rk((size_t)11 * 128 * (p_.n + 1));
        tfhe_aes_encrypt_key_bits(p_.n, lwe_sk.data(), seed, key.data(), p_.lwe_nshift, rk.data());
        check(tfhe_aes_set_round_keys(h_, rk.data()));
    }
    /* CTR keystream LWE bits [blocks][128][n+1] (bit 8 byte + j, encoded bit * 2^31)
     * of AES_k(counter block) into device memory; graph = replay a captured hipGraph. */
    void AesCtrKeystreamDevice(const std::vector<std::array<uint8_t, 16>>& ctr, uint32_t* d_out, bool graph = false,
                               hipStream_t s = hipStreamPerThread) const {
        std::vector<uint8_t> h(ctr.size() * 16);
        for (size_t k = 0; k < ctr.size(); ++k)
            for (int j = 0; j < 16; ++j) h[16 * k + j] = ctr[k][j];
        check((graph ? tfhe_aes_ctr_keystream_graph : tfhe_aes_ctr_keystream_batch)(h_, h.data(), (uint32_t)ctr.size(),
                                                                                   d_out, s));
    }
    /* Transciphering (the CTR path of RunTranscipherBatch, main.cu:1555-1580): LWE bits
     * [blocks][128][n+1] of the plaintext of public AES-CTR ciphertext blocks. */
    void TranscipherDevice(const std::vector<std::array<uint8_t, 16>>& ctr,
                           const std::vector<std::array<uint8_t, 16>>& cipher, uint32_t* d_out, bool graph = false,
                           hipStream_t s = hipStreamPerThread) const {
        if (ctr.size() != cipher.size()) throw Error(TFHE_E_SHAPE, "Transcipher: one ciphertext block per counter");
        std::vector<uint8_t> hc(ctr.size() * 16), hp(cipher.size() * 16);
        for (size_t k = 0; k < ctr.size(); ++k)
            for (int j = 0; j < 16; ++j) hc[16 * k + j] = ctr[k][j], hp[16 * k + j] = cipher[k][j];
        check(tfhe_aes_ctr_transcipher_batch(h_, graph ? 1 : 0, hc.data(), hp.data(), (uint32_t)ctr.size(), d_out, s));
    }
    std::vector<std::vector<LWECiphertext>> Transcipher(const std::vector<std::array<uint8_t, 16>>& ctr,
                                                        const std::vector<std::array<uint8_t, 16>>& cipher,
                                                        bool graph = false, hipStream_t s = hipStreamPerThread) const {
        const size_t w = (size_t)p_.n + 1, B = ctr.size();
        DeviceBuffer<uint32_t> d(B * 128 * w);
        TranscipherDevice(ctr, cipher, d.get(), graph, s);
        return download_bits(d, B, s);
    }

    /* The same keystream copied to the host as 128 LWE bits per block. */
    std::vector<std::vector<LWECiphertext>> AesCtrKeystream(const std::vector<std::array<uint8_t, 16>>& ctr,
                                                            bool graph = false,
                                                            hipStream_t s = hipStreamPerThread) const {
        const size_t w = (size_t)p_.n + 1, B = ctr.size();
        DeviceBuffer<uint32_t> d(B * 128 * w);
        AesCtrKeystreamDevice(ctr, d.get(), graph, s);
        return download_bits(d, B, s);
    }

private:
    std::vector<std::vector<LWECiphertext>> download_bits(const DeviceBuffer<uint32_t>& d, size_t B,
                                                          hipStream_t s) const {
        const size_t w = (size_t)p_.n + 1;
        std::vector<uint32_t> all(B * 128 * w);
        d.download(all.data(), all.size(), s);
        std::vector<std::vector<LWECiphertext>> out(B);
        for (size_t k = 0; k < B; ++k)
            out[k] = detail::unpack(std::vector<uint32_t>(all.begin() + k * 128 * w, all.begin() + (k + 1) * 128 * w),
                                    128, p_.n);
        return out;
    }
};

}  // namespace tfhe_amd

#endif  // TFHE_AMD_HPP
