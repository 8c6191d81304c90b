// keyblob.cpp -- versioned key files of the circuit-bootstrapping context: the
// coefficient-form keys that tfhe_cbs_set_keys takes, behind a header with the
// parameter set, the key seed and a content hash per section.  The reference has
// no file format: its keys come from the OpenFHE fork's CPU context and are
// uploaded by gpu_init_cirbsk (cirbts-init.cu:200-865); a file lets keys made
// elsewhere, or once, be uploaded without regenerating them.
//
// Layout (little-endian): a 256-byte header, then the four sections back to back:
//   rlwe_sk [N] int32 | bsk2 rows as tfhe_gen_bsk64 writes them (uint64) |
//   conversion key (lookup PFKS [2][N+1][8][15][2N] uint32, the linear PFKS key's bytes as
//   tfhe_gen_pfkl40 writes them, or the HomTrace + SchemeSwitch keys as
//   tfhe_gen_htss_keys writes them) | ksk2 [N][8][4][n+1] uint32 ([N][8][n+1] for the linear
//   key switch).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

#include "../../include/tfhe_amd.h"
#include "capi_common.hpp"
#include "cbs_layout.hpp"
#include "kernels.hpp"

using namespace tfhe;

namespace {

constexpr char kMagic[8] = {'T', 'F', 'H', 'E', 'C', 'B', 'S', 'K'};
constexpr uint32_t kVersion = 2;  // 2: the parameter set holds ks_linear (18 fields)

struct BlobHeader {
    char magic[8];
    uint32_t version;
    uint32_t header_bytes;
    tfhe_cbs_params params;
    uint64_t seed;        // the seed the keys were generated from (0: made elsewhere)
    uint64_t bytes[4];    // section sizes: rlwe_sk, bsk2, pfks, ksk2
    uint64_t hash[4];     // content hash of each section
    uint64_t header_hash; // hash of all fields above
    uint8_t pad[88];
};
static_assert(sizeof(BlobHeader) == 256, "key file header is 256 bytes");
static_assert(sizeof(tfhe_cbs_params) == 72, "the header holds the 18-field parameter set");

// Streaming 64-bit content hash over little-endian 64-bit words (an integrity check,
// not a cryptographic one): h = mix(h ^ w) per word, with the multiply-xorshift
// finalizer of MurmurHash3; the section length is folded in at the end.
struct Hash64 {
    uint64_t h = 0x243F6A8885A308D3ull;
    uint64_t len = 0;
    static uint64_t mix(uint64_t x) {
        x ^= x >> 33;
        x *= 0xff51afd7ed558ccdull;
        x ^= x >> 33;
        x *= 0xc4ceb9fe1a85ec53ull;
        x ^= x >> 33;
        return x;
    }
    void update(const uint8_t* p, size_t n) {  // n a multiple of 8 except for the last call
        size_t k = 0;
        for (; k + 8 <= n; k += 8) {
            uint64_t w;
            std::memcpy(&w, p + k, 8);
            h = mix(h ^ w);
        }
        if (k < n) {
            uint64_t w = 0;
            std::memcpy(&w, p + k, n - k);
            h = mix(h ^ w);
        }
        len += n;
    }
    uint64_t final_value() const { return mix(h ^ len); }
};

uint64_t header_hash(const BlobHeader& b) {
    Hash64 hh;
    hh.update(reinterpret_cast<const uint8_t*>(&b), offsetof(BlobHeader, header_hash));
    return hh.final_value();
}

void section_sizes(const tfhe_cbs_params& p, uint64_t (&bytes)[4]) {
    bytes[0] = (uint64_t)kN * sizeof(int32_t);
    bytes[1] = (uint64_t)bsk2_words(p) * sizeof(uint64_t);
    bytes[2] = htss(p)          ? (uint64_t)htss_key_words() * sizeof(uint64_t)
               : p.pfks_linear ? (uint64_t)pfkl_key_bytes()
                               : (uint64_t)2 * pfks_words_per_u() * sizeof(uint32_t);
    bytes[3] = (uint64_t)ksk2_words(p) * sizeof(uint32_t);
}

struct FileCloser {
    void operator()(FILE* f) const {
        if (f) std::fclose(f);
    }
};
using File = std::unique_ptr<FILE, FileCloser>;

constexpr size_t kChunk = (size_t)64 << 20;  // staging chunk for device sections

}  // namespace

namespace tfhe {

// Writes a key file from host or device sections (hipMemcpyDefault follows the pointers).
int write_key_blob(const char* path, const tfhe_cbs_params& p, uint64_t seed, const void* const (&src)[4]) {
    BlobHeader hd{};
    std::memcpy(hd.magic, kMagic, 8);
    hd.version = kVersion;
    hd.header_bytes = sizeof(BlobHeader);
    hd.params = p;
    hd.seed = seed;
    section_sizes(p, hd.bytes);
    File f(std::fopen(path, "wb"));
    if (!f) return fail(TFHE_E_PARAM, std::string("key file: cannot open ") + path + " for writing");
    // header placeholder, then the sections (hashed on the way), then the final header
    if (std::fwrite(&hd, sizeof hd, 1, f.get()) != 1) return fail(TFHE_E_PARAM, "key file: write failed");
    std::vector<uint8_t> buf(kChunk);
    for (int s = 0; s < 4; ++s) {
        Hash64 hs;
        const uint8_t* base = static_cast<const uint8_t*>(src[s]);
        for (uint64_t off = 0; off < hd.bytes[s]; off += kChunk) {
            const size_t n = (size_t)std::min<uint64_t>(kChunk, hd.bytes[s] - off);
            int rc = hipcheck(hipMemcpy(buf.data(), base + off, n, hipMemcpyDefault), "key file: read section");
            if (rc) return rc;
            hs.update(buf.data(), n);
            if (std::fwrite(buf.data(), 1, n, f.get()) != n) return fail(TFHE_E_PARAM, "key file: write failed");
        }
        hd.hash[s] = hs.final_value();
    }
    hd.header_hash = header_hash(hd);
    if (std::fseek(f.get(), 0, SEEK_SET) != 0 || std::fwrite(&hd, sizeof hd, 1, f.get()) != 1)
        return fail(TFHE_E_PARAM, "key file: header write failed");
    if (std::fflush(f.get()) != 0) return fail(TFHE_E_PARAM, "key file: flush failed");
    return TFHE_OK;
}

}  // namespace tfhe

extern "C" int tfhe_cbs_write_keys(const char* path, const tfhe_cbs_params* p, uint64_t seed, const int32_t* rlwe_sk,
                                   const uint64_t* bsk, const void* pfks, const uint32_t* ksk2) {
    if (!path || !rlwe_sk || !bsk || !pfks || !ksk2) return fail(TFHE_E_PARAM, "tfhe_cbs_write_keys: bad argument");
    if (!cbs_params_supported(p)) return fail(TFHE_E_PARAM, "tfhe_cbs_write_keys: unsupported parameters");
    const void* const src[4] = {rlwe_sk, bsk, pfks, ksk2};
    return write_key_blob(path, *p, seed, src);
}

extern "C" int tfhe_cbs_load_keys(tfhe_cbs_ctx* ctx, const char* path) {
    if (!ctx || !path) return fail(TFHE_E_PARAM, "tfhe_cbs_load_keys: bad argument");
    const tfhe_cbs_params* cp = cbs_ctx_params(ctx);
    File f(std::fopen(path, "rb"));
    if (!f) return fail(TFHE_E_PARAM, std::string("tfhe_cbs_load_keys: cannot open ") + path);
    BlobHeader hd{};
    if (std::fread(&hd, sizeof hd, 1, f.get()) != 1) return fail(TFHE_E_SHAPE, "tfhe_cbs_load_keys: short header");
    if (std::memcmp(hd.magic, kMagic, 8) != 0) return fail(TFHE_E_SHAPE, "tfhe_cbs_load_keys: not a key file");
    if (hd.version != kVersion || hd.header_bytes != sizeof(BlobHeader))
        return fail(TFHE_E_SHAPE, "tfhe_cbs_load_keys: unsupported key file version");
    if (hd.header_hash != header_hash(hd)) return fail(TFHE_E_SHAPE, "tfhe_cbs_load_keys: header hash mismatch");
    if (std::memcmp(&hd.params, cp, sizeof(tfhe_cbs_params)) != 0)
        return fail(TFHE_E_PARAM, "tfhe_cbs_load_keys: the file's parameter set differs from the context's");
    uint64_t want[4];
    section_sizes(*cp, want);
    for (int s = 0; s < 4; ++s)
        if (hd.bytes[s] != want[s]) return fail(TFHE_E_SHAPE, "tfhe_cbs_load_keys: section size mismatch");
    std::vector<std::vector<uint8_t>> sec(4);
    for (int s = 0; s < 4; ++s) {
        sec[s].resize(hd.bytes[s]);
        Hash64 hs;
        for (uint64_t off = 0; off < hd.bytes[s]; off += kChunk) {
            const size_t n = (size_t)std::min<uint64_t>(kChunk, hd.bytes[s] - off);
            if (std::fread(sec[s].data() + off, 1, n, f.get()) != n)
                return fail(TFHE_E_SHAPE, "tfhe_cbs_load_keys: truncated key file");
            hs.update(sec[s].data() + off, n);
        }
        if (hs.final_value() != hd.hash[s]) return fail(TFHE_E_SHAPE, "tfhe_cbs_load_keys: content hash mismatch");
    }
    return tfhe_cbs_set_keys(ctx, reinterpret_cast<const int32_t*>(sec[0].data()),
                             reinterpret_cast<const uint64_t*>(sec[1].data()),
                             reinterpret_cast<const uint32_t*>(sec[2].data()),
                             reinterpret_cast<const uint32_t*>(sec[3].data()));
}

extern "C" int tfhe_cbs_key_file_info(const char* path, tfhe_cbs_params* params, uint64_t* seed) {
    if (!path || !params || !seed) return fail(TFHE_E_PARAM, "tfhe_cbs_key_file_info: bad argument");
    File f(std::fopen(path, "rb"));
    if (!f) return fail(TFHE_E_PARAM, std::string("tfhe_cbs_key_file_info: cannot open ") + path);
    BlobHeader hd{};
    if (std::fread(&hd, sizeof hd, 1, f.get()) != 1 || std::memcmp(hd.magic, kMagic, 8) != 0 ||
        hd.version != kVersion || hd.header_hash != header_hash(hd))
        return fail(TFHE_E_SHAPE, "tfhe_cbs_key_file_info: not a valid key file");
    *params = hd.params;
    *seed = hd.seed;
    return TFHE_OK;
}
