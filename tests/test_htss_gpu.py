"""HomTrace + SchemeSwitch circuit bootstrapping (lwe_to_rgsw = 1, the reference's own
LWE -> RGSW conversion) on the GPU against the oracle's restatement (orc_cbs_htss,
pinned on the CPU by the trace identity, key-row decryption and row decryption):
keys and RGSW rows bit for bit, AES-128 through it, key files."""
import numpy as np
import pytest

import oracle as O
from conftest import from_dev, to_dev

pytestmark = pytest.mark.gpu

SEED = 23


@pytest.fixture(scope="module")
def hkeys():
    return O.CbsKeys(SEED, p=O.cbs_params_htss())


@pytest.fixture(scope="module")
def hctx(gpu, hkeys):
    import tfhe_amd
    c = tfhe_amd.CbsContext(max_blocks=1, params=tfhe_amd.cbs_params(lwe_to_rgsw=1))
    c.keygen(hkeys.s, SEED)
    yield c
    c.close()


def test_htss_keys_and_bsk_equal_oracle(hctx, hkeys):
    """tfhe_gen_htss_keys writes the oracle's HomTrace and SchemeSwitch keys (Philox domains
    11 and 12) bit for bit, and tfhe_gen_bsk64 with the 2^10 x 4 gadget the oracle's BSK2."""
    import torch
    import tfhe_amd
    p = tfhe_amd.cbs_params(lwe_to_rgsw=1)
    hkeys._htss_keys()
    want = np.concatenate([hkeys.ht_key, hkeys.ss_key])
    L = tfhe_amd.lib()
    d_K2 = to_dev(hctx.rlwe_sk().view(np.uint32))
    d_keys = torch.empty(want.size, dtype=torch.int64, device="cuda")
    tfhe_amd.check(L.tfhe_gen_htss_keys(p, SEED, tfhe_amd._ptr(d_K2), tfhe_amd._ptr(d_keys), None))
    torch.cuda.synchronize()
    assert np.array_equal(from_dev(d_keys, np.uint64), want)
    d_s = to_dev(hkeys.s.view(np.uint32))
    d_bsk = torch.empty(hkeys.bsk.size, dtype=torch.int64, device="cuda")
    tfhe_amd.check(L.tfhe_gen_bsk64(p, SEED, tfhe_amd._ptr(d_s), tfhe_amd._ptr(d_K2), tfhe_amd._ptr(d_bsk), None))
    torch.cuda.synchronize()
    assert np.array_equal(from_dev(d_bsk, np.uint64), hkeys.bsk)


def test_cbs_htss_bit_exact(hctx, hkeys):
    """Circuit bootstraps through the exact 2^10 x 4 PBS and HomTrace + SchemeSwitch equal the
    oracle's rows word for word (every product exact on both sides), bits 0 and 1, rows in the
    reference's order (2j SS, 2j + 1 HT)."""
    import torch
    cts = np.stack([hkeys.encrypt_bit(b, 300 + k) for k, b in enumerate([0, 1, 1, 0])])
    d_rgsw = torch.empty((4, 4, 2, hkeys.p.N), dtype=torch.int32, device="cuda")
    hctx.cbs_batch(to_dev(cts), d_rgsw)
    torch.cuda.synchronize()
    got = from_dev(d_rgsw)
    for k in range(4):
        want, err = hkeys.cbs_htss(cts[k])
        assert err < 0.25
        assert np.array_equal(got[k].reshape(-1), want), f"input {k}"


def _decrypt_blocks(keys, bits_ct, blocks):
    out = []
    for b in range(blocks):
        v = bytearray(16)
        for bit in range(128):
            v[bit // 8] |= keys.decrypt_bit(bits_ct[b, bit]) << (bit % 8)
        out.append(bytes(v))
    return out


def test_aes128_through_htss(hctx, hkeys):
    """The AES-128 CTR keystream of FIPS-197 C.1 through HomTrace + SchemeSwitch circuit
    bootstraps decrypts to the published ciphertext; every phase inside the p = 2 margin."""
    import torch
    import tfhe_amd
    key = bytes(range(16))
    hctx.set_round_keys(tfhe_amd.aes_encrypt_key_bits(hkeys.s, 77, key))
    d_ks = torch.empty((1, 128, hkeys.p.n + 1), dtype=torch.int32, device="cuda")
    hctx.aes_ctr_keystream([bytes.fromhex("00112233445566778899aabbccddeeff")], d_ks)
    torch.cuda.synchronize()
    got = from_dev(d_ks)
    assert _decrypt_blocks(hkeys, got, 1)[0].hex() == "69c4e0d86a7b0430d8cdb78070b4c55a"
    s64 = hkeys.s.astype(np.uint64)
    ph = (got[:, :, -1].astype(np.uint64) - (got[:, :, :-1].astype(np.uint64) @ s64)) % 2 ** 32
    err = ((ph + 2 ** 30) % 2 ** 31).astype(np.int64) - 2 ** 30
    print(f"AES through HomTrace + SchemeSwitch: max |phase error| 2^{np.log2(np.abs(err).max()):.2f} of 2^32")
    assert np.abs(err).max() < 2 ** 29


def test_htss_key_file_round_trip(hctx, hkeys, tmp_path):
    """A HomTrace-form context's keys saved and loaded into a fresh context give the same rows."""
    import torch
    import tfhe_amd
    path = tmp_path / "htss.tfhe"
    hctx.save_keys(path)
    c2 = tfhe_amd.CbsContext(max_blocks=1, params=tfhe_amd.cbs_params(lwe_to_rgsw=1))
    c2.load_keys(path)
    cts = to_dev(np.stack([hkeys.encrypt_bit(1, 310)]))
    outs = []
    for c in (hctx, c2):
        d = torch.empty((1, 4, 2, hkeys.p.N), dtype=torch.int32, device="cuda")
        c.cbs_batch(cts, d)
        torch.cuda.synchronize()
        outs.append(from_dev(d))
    assert np.array_equal(outs[0], outs[1])
    c2.close()
