"""GPU parity of the circuit-bootstrapping level (N = 2048: PBS on the 64-bit
torus, PFKS / RGSW / LUT on the 32-bit torus) and the AES-128 transciphering
pipeline against the CPU oracle, through the C ABI."""
import numpy as np
import pytest

import oracle as O
from conftest import from_dev

pytestmark = pytest.mark.gpu

SEED = 21


def dev(a, dtype):
    import torch
    a = np.ascontiguousarray(a)
    t = torch.from_numpy(a.view(dtype).copy())
    return t.to("cuda")


def to_dev32(a):
    import torch
    return dev(np.asarray(a, np.uint32), np.int32)


def to_dev64(a):
    import torch
    return dev(np.asarray(a, np.uint64), np.int64)


@pytest.fixture(scope="module")
def okeys():
    return O.CbsKeys(SEED)


@pytest.fixture(scope="module")
def ctx(gpu, okeys):
    import tfhe_amd
    c = tfhe_amd.CbsContext(max_blocks=5)  # 640 inputs: the large-batch PBS form and a split batch
    c.keygen(okeys.s, SEED)
    yield c
    c.close()


def test_polymul2048_exact(gpu):
    import torch
    import tfhe_amd
    rng = np.random.default_rng(4)
    batch = 5
    x = rng.integers(-256, 256, (batch, 2048)).astype(np.int32)
    y = rng.integers(0, 2 ** 64, (batch, 2048), dtype=np.uint64)
    d_out = torch.empty((batch, 2048), dtype=torch.int64, device="cuda")
    tfhe_amd.polymul2048_batch(to_dev32(x.view(np.uint32)), to_dev64(y), d_out)
    torch.cuda.synchronize()
    got = from_dev(d_out, np.uint64)
    for k in range(batch):
        assert np.array_equal(got[k], O.nc_mul_exact64(x[k], y[k])), k


def test_polymul2048_one_product_error(gpu):
    """The PBS's one-product form of a digit-polynomial x torus-64 product differs
    from the exact integer product (the oracle's schoolbook restatement) by FP64
    rounding only: every coefficient within 2^-36 of the torus (a CPU emulation
    gives about 2^-39)."""
    import torch
    import tfhe_amd
    rng = np.random.default_rng(14)
    batch = 8
    x = rng.integers(-256, 256, (batch, 2048)).astype(np.int32)
    y = rng.integers(0, 2 ** 64, (batch, 2048), dtype=np.uint64)
    d_out = torch.empty((batch, 2048), dtype=torch.int64, device="cuda")
    tfhe_amd.polymul2048_batch(to_dev32(x.view(np.uint32)), to_dev64(y), d_out, approx=True)
    torch.cuda.synchronize()
    got = from_dev(d_out, np.uint64)
    worst = 0
    for k in range(batch):
        d = np.abs((got[k] - O.nc_mul_exact64(x[k], y[k])).view(np.int64))
        worst = max(worst, int(d.max()))
    print(f"one-product polymul: max |error| = 2^{np.log2(max(worst, 1)):.2f} of 2^64")
    assert 0 < worst < 2 ** 28


def test_keygen_parity(ctx, okeys):
    """GPU key generation reproduces the oracle's K2, BSK2 and PFKS entries bit for bit."""
    import torch
    import tfhe_amd
    assert np.array_equal(ctx.rlwe_sk(), okeys.K2)
    p = tfhe_amd.cbs_params()
    d_bsk = torch.empty(okeys.bsk.size, dtype=torch.int64, device="cuda")
    d_s = to_dev32(okeys.s.view(np.uint32))  # keep the device buffers alive across the call
    d_K2 = to_dev32(okeys.K2.view(np.uint32))
    tfhe_amd.check(tfhe_amd.lib().tfhe_gen_bsk64(p, SEED, tfhe_amd._ptr(d_s), tfhe_amd._ptr(d_K2),
                                                 tfhe_amd._ptr(d_bsk), None))
    torch.cuda.synchronize()
    assert np.array_equal(from_dev(d_bsk, np.uint64), okeys.bsk)
    for (u, i, t, v) in [(0, 0, 0, 1), (1, 0, 0, 1), (0, 2048, 7, 15), (1, 1000, 3, 9), (0, 17, 5, 4)]:
        assert np.array_equal(ctx.pfks_entry(u, i, t, v), okeys.pfks_entry(u, i, t, v)), (u, i, t, v)


def _bits_cts(okeys, bits, base_id):
    return np.stack([okeys.encrypt_bit(int(b), base_id + k) for k, b in enumerate(bits)])


def test_pbs_manylut_bit_exact(ctx, okeys):
    import torch
    bits = [0, 1, 1, 0]
    cts = _bits_cts(okeys, bits, 300)
    d_out = torch.empty((len(bits), 2, 2049), dtype=torch.int64, device="cuda")
    ctx.pbs_batch(to_dev32(cts), d_out)
    torch.cuda.synchronize()
    got = from_dev(d_out, np.uint64)
    for k, b in enumerate(bits):
        want, err = okeys.pbs(cts[k])
        assert err < 0.25
        assert np.array_equal(got[k], want), k


def test_pbs_large_batch_form_matches(ctx, okeys):
    """Batches above 256 run the two-waves-per-SIMD PBS form, small ones the
    branch-free form: 300 bootstraps equal the same inputs bootstrapped in small
    sub-batches and the oracle on sampled inputs, bit for bit."""
    import torch
    import tfhe_amd
    batch = 300
    rng = np.random.default_rng(17)
    bits = rng.integers(0, 2, batch)
    cts = np.stack([tfhe_amd.lwe_encrypt(okeys.s, 31, k, int(bits[k]), p=2) for k in range(batch)])
    d_cts = to_dev32(cts)
    d_big = torch.empty((batch, 2, 2049), dtype=torch.int64, device="cuda")
    ctx.pbs_batch(d_cts, d_big)
    torch.cuda.synchronize()
    big = from_dev(d_big, np.uint64)
    for lo in (0, 150, 296):
        hi = min(lo + 4, batch)
        d_small = torch.empty((hi - lo, 2, 2049), dtype=torch.int64, device="cuda")
        ctx.pbs_batch(d_cts[lo:hi].contiguous(), d_small)
        torch.cuda.synchronize()
        assert np.array_equal(from_dev(d_small, np.uint64), big[lo:hi]), lo
    for k in (0, 299):
        want, err = okeys.pbs(cts[k])
        assert err < 0.25
        assert np.array_equal(big[k], want), k


def test_pfks_bit_exact(ctx, okeys):
    """Both functions on the two PBS outputs, plus random torus-64 inputs whose
    rounding hits the 2^32 wrap (a_i >= 2^64 - 2^31) and every digit value; the
    output rows are strided (out_stride > 2N) and the gap stays untouched."""
    import torch
    cts = _bits_cts(okeys, [1], 400)
    d_pbs = torch.empty((1, 2, 2049), dtype=torch.int64, device="cuda")
    ctx.pbs_batch(to_dev32(cts), d_pbs)
    torch.cuda.synchronize()
    lwes = from_dev(d_pbs, np.uint64)[0]
    rng = np.random.default_rng(12)
    extra = rng.integers(0, 2 ** 64, (3, 2049), dtype=np.uint64)
    extra[0, :64] = np.uint64(2 ** 64 - 2 ** 31 + 5)
    extra[1, :16] = np.array([(v * 0x11111111) << 32 for v in range(16)], np.uint64)
    lwes = np.concatenate([lwes, extra])
    d_lwe64 = to_dev64(lwes)
    count, stride = len(lwes), 4096 + 64
    for u in (0, 1):
        d_out = torch.full((count, stride), 7, dtype=torch.int32, device="cuda")
        ctx.pfks_batch(u, d_lwe64, count, stride, d_out)
        torch.cuda.synchronize()
        got = from_dev(d_out)
        for j in range(count):
            assert np.array_equal(got[j, :4096], okeys.pfks(u, lwes[j])), (u, j)
        assert (got[:, 4096:] == 7).all()


def test_cbs_split_batch_equals_parts(ctx, okeys):
    """600 inputs are one full round of 512 PBS workgroups plus a tail of 88, which
    runs on the context's second stream beside the first part's PFKS: the rows equal
    two unsplit calls of 300 bit for bit."""
    import torch
    import tfhe_amd
    batch = 600
    rng = np.random.default_rng(41)
    bits = rng.integers(0, 2, batch)
    cts = np.stack([tfhe_amd.lwe_encrypt(okeys.s, 35, k, int(bits[k]), p=2) for k in range(batch)])
    d_cts = to_dev32(cts)
    d_all = torch.empty((batch, 4, 2, 2048), dtype=torch.int32, device="cuda")
    ctx.cbs_batch(d_cts, d_all)
    d_part = torch.empty((batch // 2, 4, 2, 2048), dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    whole = from_dev(d_all)
    for lo in (0, batch // 2):
        ctx.cbs_batch(d_cts[lo:lo + batch // 2].contiguous(), d_part)
        torch.cuda.synchronize()
        assert np.array_equal(from_dev(d_part), whole[lo:lo + batch // 2]), lo


def test_cbs_from_device_lwe_view(ctx, okeys):
    """The reference's argument layout (u64 masks a_stride apart, u64 bodies, a
    borrowed view of outputs held in the context) gives the same RGSW rows as
    tfhe_cbs_batch on the packed LWEs; only the low 32 bits of each word count.
    NULL inputs give an empty view, a_stride < n an error (cirbts-batch.cu:1423-1440)."""
    import ctypes
    import torch
    import tfhe_amd
    L = tfhe_amd.lib()
    cts = _bits_cts(okeys, [1, 0, 1, 1, 0], 950)
    n = cts.shape[1] - 1
    stride = n + 7
    a = np.full((5, stride), 0xDEADBEEF12345678, np.uint64)  # row padding the entry must skip
    a[:, :n] = cts[:, :n].astype(np.uint64) | (np.uint64(5) << np.uint64(40))
    b = cts[:, n].astype(np.uint64)
    d_a, d_b = to_dev64(a), to_dev64(b)
    v = ctx.cbs_batch_from_device_lwe(d_a, d_b, 5, stride)
    assert (v.batch, v.digits, v.N, v.per_ciphertext_elems) == (5, 2, 2048, 2 * 4 * 2048)
    d_f_view = torch.empty((5, 8, 1024, 2), dtype=torch.float64, device="cuda")
    tfhe_amd.check(L.tfhe_rgsw_to_fourier(ctx.h, ctypes.c_void_p(v.d_rgsw), 5, tfhe_amd._ptr(d_f_view), None))
    d_rgsw = torch.empty((5, 4, 2, 2048), dtype=torch.int32, device="cuda")
    ctx.cbs_batch(to_dev32(cts), d_rgsw)
    d_f = torch.empty_like(d_f_view)
    ctx.rgsw_to_fourier(d_rgsw, 5, d_f)
    torch.cuda.synchronize()
    assert torch.equal(d_f_view, d_f)  # the exact transforms of equal rows
    for args in ((None, d_b, 5), (d_a, None, 5), (d_a, d_b, 0)):
        e = ctx.cbs_batch_from_device_lwe(*args, stride)
        assert e.batch == 0 and not e.d_rgsw
    with pytest.raises(tfhe_amd.TfheError) as err:
        ctx.cbs_batch_from_device_lwe(d_a, d_b, 5, n - 1)
    assert err.value.code == tfhe_amd.TFHE_E_PARAM


def test_cbs_rgsw_rows_decrypt(ctx, okeys):
    """Circuit bootstrap of 6 bits: row 2j+u decrypts to f_u(bit * 2^(16+8j)) on
    the 32-bit torus; the first input is also checked bit for bit against
    the oracle's circuit bootstrap."""
    import torch
    bits = [1, 0, 1, 1, 0, 1]
    cts = _bits_cts(okeys, bits, 500)
    d_rgsw = torch.empty((len(bits), 4, 2, 2048), dtype=torch.int32, device="cuda")
    ctx.cbs_batch(to_dev32(cts), d_rgsw)
    torch.cuda.synchronize()
    rg = from_dev(d_rgsw)
    want0, err0 = okeys.cbs(cts[0])
    assert err0 < 0.25
    assert np.array_equal(rg[0].reshape(-1), want0)
    for k, b in enumerate(bits):
        for j in range(2):
            mu = np.zeros(2048, np.uint32)
            mu[0] = np.uint32(b << (16 + 8 * j))
            for u in range(2):
                ph = O.rlwe_phase32(okeys.K2, rg[k, 2 * j + u].reshape(-1))
                want = mu if u == 1 else (-O.nc_mul_exact32(okeys.K2, mu)).astype(np.uint32)
                err = ((ph - want).astype(np.uint32)).view(np.int32)
                assert np.abs(err).max() < 2 ** 16, (k, j, u)  # ~2^12 expected


def _sbox_poly():
    S = [O.aes_sbox(x) for x in range(256)]
    poly = np.zeros(2048, np.uint32)
    for x in range(256):
        for j in range(8):
            poly[8 * x + j] = np.uint32(((S[x] >> j) & 1) << 31)
    return poly, S


def test_lut_bit_exact(ctx, okeys):
    """Vertical packing on fresh RGSWs equals the oracle (extracted, switched to 32 bits)."""
    import torch
    poly, S = _sbox_poly()
    xs = [0x53, 0x00]
    rgsws = [okeys.rgsw_encrypt((x >> i) & 1, 2000 + 16 * k + 2 * i) for k, x in enumerate(xs) for i in range(8)]
    d_rgsw = to_dev32(np.concatenate(rgsws))
    d_f = torch.empty(len(rgsws) * 8 * 1024 * 2, dtype=torch.float64, device="cuda")
    ctx.rgsw_to_fourier(d_rgsw, len(rgsws), d_f)
    d_out = torch.empty((len(xs), 8, 2049), dtype=torch.int32, device="cuda")
    ctx.lut8(d_f, len(xs), to_dev32(poly), 0, 8, d_out)
    torch.cuda.synchronize()
    got = from_dev(d_out)
    for k, x in enumerate(xs):
        fs = [okeys.rgsw_fourier(r) for r in rgsws[8 * k:8 * k + 8]]
        acc, err = okeys.vertical_packing(fs, poly)
        assert err < 0.25
        for j in range(8):
            assert np.array_equal(got[k, j], okeys.extract32(acc, j)), (x, j)


def _xtime(b):
    return ((b << 1) ^ (0x1B if b & 0x80 else 0)) & 0xFF


def test_lut8x32_ttable(ctx, okeys):
    """8 -> 32-bit LUT on the AES T-table Te0 (bytes 2 S(x), S(x), S(x), 3 S(x)): four
    table polynomials share each byte's 8 RGSWs; every output equals the oracle's
    vertical packing of its part, and the 32 bits decrypt to Te0[x]."""
    import torch
    S = [O.aes_sbox(x) for x in range(256)]
    te0 = [_xtime(S[x]) | (S[x] << 8) | (S[x] << 16) | ((_xtime(S[x]) ^ S[x]) << 24) for x in range(256)]
    polys = np.zeros((4, 2048), np.uint32)
    for x in range(256):
        for b in range(32):
            polys[b // 8, 8 * x + b % 8] = np.uint32(((te0[x] >> b) & 1) << 31)
    xs = [0x53, 0x00, 0xA7]
    rgsws = [okeys.rgsw_encrypt((x >> i) & 1, 3000 + 16 * k + 2 * i) for k, x in enumerate(xs) for i in range(8)]
    d_rgsw = to_dev32(np.concatenate(rgsws))
    d_f = torch.empty(len(rgsws) * 8 * 1024 * 2, dtype=torch.float64, device="cuda")
    ctx.rgsw_to_fourier(d_rgsw, len(rgsws), d_f)
    d_out = torch.empty((len(xs), 32, 2049), dtype=torch.int32, device="cuda")
    ctx.lut8x32(d_f, len(xs), to_dev32(polys), d_out)
    torch.cuda.synchronize()
    got = from_dev(d_out)
    K = okeys.K2.astype(np.int64).astype(np.uint64)
    for k, x in enumerate(xs):
        fs = [okeys.rgsw_fourier(r) for r in rgsws[8 * k:8 * k + 8]]
        word = 0
        for p in range(4):
            acc, err = okeys.vertical_packing(fs, polys[p])
            assert err < 0.25
            for j in range(8):
                lwe = got[k, 8 * p + j]
                assert np.array_equal(lwe, okeys.extract32(acc, j)), (hex(x), p, j)
                ph = (int(lwe[-1]) - int(np.dot(lwe[:-1].astype(np.uint64), K))) % 2 ** 32
                word |= (((ph + 2 ** 30) >> 31) & 1) << (8 * p + j)
        assert word == te0[x], (hex(x), hex(word), hex(te0[x]))


def test_ks2_bit_exact(ctx, okeys):
    import torch
    rng = np.random.default_rng(9)
    count = 19
    lwe64 = rng.integers(0, 2 ** 64, (count, 2049), dtype=np.uint64)
    l32 = ((lwe64 + np.uint64(1 << 31)) >> np.uint64(32)).astype(np.uint32)
    d_out = torch.empty((count, 637), dtype=torch.int32, device="cuda")
    ctx.keyswitch_batch(to_dev32(l32), count, d_out)
    torch.cuda.synchronize()
    got = from_dev(d_out)
    for k in range(count):
        assert np.array_equal(got[k], okeys.ks2(lwe64[k])), k


def test_sbox_end_to_end(ctx, okeys):
    """Encrypted bytes -> circuit bootstrap of each bit -> S-box LUT -> key switch
    decrypts to S(x) for every probed byte."""
    import torch
    poly, S = _sbox_poly()
    xs = [0x00, 0x53, 0xFF, 0x1C, 0xA7, 0x80]
    bits = [(x >> i) & 1 for x in xs for i in range(8)]
    cts = _bits_cts(okeys, bits, 700)
    nb = len(bits)
    d_rgsw = torch.empty((nb, 4, 2, 2048), dtype=torch.int32, device="cuda")
    ctx.cbs_batch(to_dev32(cts), d_rgsw)
    d_f = torch.empty(nb * 8 * 1024 * 2, dtype=torch.float64, device="cuda")
    ctx.rgsw_to_fourier(d_rgsw, nb, d_f)
    d_l = torch.empty((len(xs), 8, 2049), dtype=torch.int32, device="cuda")
    ctx.lut8(d_f, len(xs), to_dev32(poly), 0, 8, d_l)
    d_out = torch.empty((len(xs) * 8, 637), dtype=torch.int32, device="cuda")
    ctx.keyswitch_batch(d_l, len(xs) * 8, d_out)
    torch.cuda.synchronize()
    lut = from_dev(d_l)
    got = from_dev(d_out)
    for k, x in enumerate(xs):
        val = sum(okeys.decrypt_bit(got[8 * k + j]) << j for j in range(8))
        assert val == S[x], (hex(x), hex(val), hex(S[x]))
        for j in range(8):  # LUT output noise (torus-32 phase under K2) well inside the 2^30 margin
            ph = (int(lut[k, j, -1]) - int(np.dot(lut[k, j, :-1].astype(np.uint64), okeys.K2.astype(np.uint64)))) % 2 ** 32
            e = ((ph - (((S[x] >> j) & 1) << 31) + 2 ** 31) % 2 ** 32) - 2 ** 31
            assert abs(e) < 2 ** 28, (hex(x), j, e)


def _decrypt_blocks(okeys, bits_ct, blocks):
    out = []
    for b in range(blocks):
        v = bytearray(16)
        for bit in range(128):
            v[bit // 8] |= okeys.decrypt_bit(bits_ct[b, bit]) << (bit % 8)
        out.append(bytes(v))
    return out


def test_aes128_fips197_and_random_blocks(ctx, okeys):
    """Homomorphic AES-128 (CTR keystream with the counter block as input):
    the FIPS-197 C.1 vector and a seeded random block, both bit-exact after
    decryption; every output phase inside the p = 2 decision margin."""
    import torch
    import tfhe_amd
    key = bytes(range(16))
    rk = tfhe_amd.aes_encrypt_key_bits(okeys.s, 77, key)
    ctx.set_round_keys(rk)
    blocks = [bytes.fromhex("00112233445566778899aabbccddeeff"), bytes(np.random.default_rng(3).integers(0, 256, 16, dtype=np.uint8))]
    d_ks = torch.empty((2, 128, 637), dtype=torch.int32, device="cuda")
    ctx.aes_ctr_keystream(blocks, d_ks)
    torch.cuda.synchronize()
    got = from_dev(d_ks)
    dec = _decrypt_blocks(okeys, got, 2)
    assert dec[0].hex() == "69c4e0d86a7b0430d8cdb78070b4c55a"
    assert dec[1] == O.aes128_encrypt(key, blocks[1]) == tfhe_amd.aes128_encrypt_plain(key, blocks[1])
    s64 = okeys.s.astype(np.uint64)
    ph = (got[:, :, -1].astype(np.uint64) - (got[:, :, :-1].astype(np.uint64) @ s64)) % 2 ** 32
    err = ((ph + 2 ** 30) % 2 ** 31).astype(np.int64) - 2 ** 30
    assert np.abs(err).max() < 2 ** 29  # within half of the 2^30 margin


def test_aes_keystream_graph_replays_match_eager(ctx, okeys):
    """The hipGraph-captured keystream equals the eager launch sequence on every
    replay, with eager work and new counter blocks in between replays.  5 blocks
    (640 inputs) split each round's PBS across the context's two streams, so the
    capture holds the fork and join."""
    import torch
    import tfhe_amd
    key = bytes(range(100, 116))
    ctx.set_round_keys(tfhe_amd.aes_encrypt_key_bits(okeys.s, 79, key))
    rng = np.random.default_rng(8)
    d_eager = torch.empty((5, 128, 637), dtype=torch.int32, device="cuda")
    d_graph = torch.empty((5, 128, 637), dtype=torch.int32, device="cuda")
    for rep in range(3):
        ctrs = [bytes(rng.integers(0, 256, 16, dtype=np.uint8)) for _ in range(5)]
        ctx.aes_ctr_keystream(ctrs, d_eager)
        torch.cuda.synchronize()
        ctx.aes_ctr_keystream(ctrs, d_graph, graph=True)
        torch.cuda.synchronize()
        assert np.array_equal(from_dev(d_graph), from_dev(d_eager)), rep
        assert _decrypt_blocks(okeys, from_dev(d_graph), 5) == [O.aes128_encrypt(key, c) for c in ctrs]


def test_aes_ctr_transcipher_recovers_plaintext(ctx, okeys):
    """Transciphering (main.cu:1555-1580): the public AES-CTR ciphertext of seeded
    random plaintext blocks, added as public bits to the keystream, decrypts to the
    plaintext, in the batched and graph modes; the batched mode equals the keystream
    with the ciphertext bits added on the host."""
    import torch
    import tfhe_amd
    key = bytes(range(200, 216))
    ctx.set_round_keys(tfhe_amd.aes_encrypt_key_bits(okeys.s, 85, key))
    rng = np.random.default_rng(14)
    ctrs = [bytes(8) + int(b).to_bytes(8, "big") for b in range(3)]
    plain = [bytes(rng.integers(0, 256, 16, dtype=np.uint8)) for _ in range(3)]
    cipher = [bytes(p ^ k for p, k in zip(pt, O.aes128_encrypt(key, c))) for pt, c in zip(plain, ctrs)]
    d_ks = torch.empty((3, 128, 637), dtype=torch.int32, device="cuda")
    ctx.aes_ctr_keystream(ctrs, d_ks)
    torch.cuda.synchronize()
    ks = from_dev(d_ks).copy()
    pub = np.array([[(cipher[b][bit // 8] >> (bit % 8)) & 1 for bit in range(128)] for b in range(3)], np.uint64)
    ks[:, :, -1] = ((ks[:, :, -1].astype(np.uint64) + (pub << np.uint64(31))) % 2 ** 32).astype(np.uint32)
    for mode in ("batch", "graph", "graph"):
        d = torch.empty_like(d_ks)
        ctx.aes_ctr_transcipher(ctrs, cipher, d, mode=mode)
        torch.cuda.synchronize()
        got = from_dev(d)
        assert _decrypt_blocks(okeys, got, 3) == plain, mode
        assert np.array_equal(got, ks), mode


def test_uneven_keystream_groups_equal_separate_calls(lctx, okeys):
    """5 blocks are one full round of resident PBS workgroups (4 blocks) plus one: in a
    linear-PFKS context the remainder runs as its own group on the second stream; the
    keystream equals the keystreams of the 4 and the 1 block computed in separate calls,
    bit for bit."""
    import torch
    import tfhe_amd
    ctx = lctx
    key = bytes(range(7, 23))
    ctx.set_round_keys(tfhe_amd.aes_encrypt_key_bits(okeys.s, 87, key))
    rng = np.random.default_rng(15)
    ctrs = [bytes(rng.integers(0, 256, 16, dtype=np.uint8)) for _ in range(5)]
    d_all = torch.empty((5, 128, 637), dtype=torch.int32, device="cuda")
    ctx.aes_ctr_keystream(ctrs, d_all)
    d4 = torch.empty((4, 128, 637), dtype=torch.int32, device="cuda")
    d1 = torch.empty((1, 128, 637), dtype=torch.int32, device="cuda")
    ctx.aes_ctr_keystream(ctrs[:4], d4)
    ctx.aes_ctr_keystream(ctrs[4:], d1)
    torch.cuda.synchronize()
    whole = from_dev(d_all)
    assert np.array_equal(whole[:4], from_dev(d4))
    assert np.array_equal(whole[4:], from_dev(d1))
    assert _decrypt_blocks(okeys, whole, 5) == [O.aes128_encrypt(key, c) for c in ctrs]


def test_set_keys_from_stateless_generators(ctx, okeys):
    """Keys made outside a context (tfhe_gen_bsk64 / tfhe_gen_pfks32 on the device,
    the oracle's KSK2 on the host) uploaded with tfhe_cbs_set_keys give the same
    PBS and circuit-bootstrap outputs as the keygen'd context, bit for bit."""
    import torch
    import tfhe_amd
    p = tfhe_amd.cbs_params()
    L = tfhe_amd.lib()
    d_s = to_dev32(okeys.s.view(np.uint32))
    d_K2 = to_dev32(okeys.K2.view(np.uint32))
    d_bsk = torch.empty(okeys.bsk.size, dtype=torch.int64, device="cuda")
    tfhe_amd.check(L.tfhe_gen_bsk64(p, SEED, tfhe_amd._ptr(d_s), tfhe_amd._ptr(d_K2), tfhe_amd._ptr(d_bsk), None))
    d_pfks = torch.empty(2 * 2049 * 8 * 15 * 2 * 2048, dtype=torch.int32, device="cuda")
    tfhe_amd.check(L.tfhe_gen_pfks32(p, SEED, tfhe_amd._ptr(d_K2), tfhe_amd._ptr(d_pfks), None))
    c2 = tfhe_amd.CbsContext(max_blocks=1)
    try:
        c2.set_keys(okeys.K2, d_bsk, d_pfks, okeys.ksk2)
        del d_pfks
        for (u, i, t, v) in [(0, 3, 1, 2), (1, 2048, 7, 15)]:
            assert np.array_equal(c2.pfks_entry(u, i, t, v), okeys.pfks_entry(u, i, t, v)), (u, i, t, v)
        cts = to_dev32(_bits_cts(okeys, [1, 0, 1], 900))
        outs = []
        for c in (ctx, c2):
            d_pbs = torch.empty((3, 2, 2049), dtype=torch.int64, device="cuda")
            d_rgsw = torch.empty((3, 4, 2, 2048), dtype=torch.int32, device="cuda")
            c.pbs_batch(cts, d_pbs)
            c.cbs_batch(cts, d_rgsw)
            torch.cuda.synchronize()
            outs.append((from_dev(d_pbs, np.uint64), from_dev(d_rgsw)))
        assert np.array_equal(outs[0][0], outs[1][0])
        assert np.array_equal(outs[0][1], outs[1][1])
    finally:
        c2.close()


def test_aes_plain_form_equals_batched(ctx, okeys):
    """The per-bootstrap keystream (every circuit bootstrap, LUT and key switch a
    launch of one) equals the batched keystream bit for bit and decrypts to AES."""
    import torch
    import tfhe_amd
    key = bytes(range(50, 66))
    ctx.set_round_keys(tfhe_amd.aes_encrypt_key_bits(okeys.s, 81, key))
    ctr = [bytes(np.random.default_rng(6).integers(0, 256, 16, dtype=np.uint8))]
    d_batch = torch.empty((1, 128, 637), dtype=torch.int32, device="cuda")
    d_plain = torch.empty((1, 128, 637), dtype=torch.int32, device="cuda")
    ctx.aes_ctr_keystream(ctr, d_batch)
    ctx.aes_ctr_keystream(ctr, d_plain, plain=True)
    torch.cuda.synchronize()
    assert np.array_equal(from_dev(d_plain), from_dev(d_batch))
    assert _decrypt_blocks(okeys, from_dev(d_plain), 1) == [O.aes128_encrypt(key, ctr[0])]


def _check_choices(ch, forms=(0, 1)):
    for k in range(2):
        assert ch["pbs_form"][k] == -1 or ch["pbs_form"][k] in forms  # the forms the context's PBS has
        assert ch["pfks_slices"][k] == -1 or 4 <= ch["pfks_slices"][k] <= 32
    assert ch["pfks_slices"][1] >= 4  # the whole batch or its tail always exists
    assert 1 <= ch["ks_slices"] <= 64


def test_autotune_keeps_keystream_words(ctx, okeys):
    """tfhe_cbs_autotune changes launch shapes only: after tuning 5 blocks (a batch
    split into a full round and a tail) and 1 block, the eager and the recaptured
    graph keystreams equal the untuned keystream bit for bit."""
    import torch
    import tfhe_amd
    key = bytes(range(30, 46))
    ctx.set_round_keys(tfhe_amd.aes_encrypt_key_bits(okeys.s, 83, key))
    rng = np.random.default_rng(12)
    ctrs = [bytes(rng.integers(0, 256, 16, dtype=np.uint8)) for _ in range(5)]
    d_ref = torch.empty((5, 128, 637), dtype=torch.int32, device="cuda")
    ctx.aes_ctr_keystream(ctrs, d_ref, graph=True)  # captured before tuning: dropped by it
    torch.cuda.synchronize()
    ref = from_dev(d_ref)
    assert _decrypt_blocks(okeys, ref, 5) == [O.aes128_encrypt(key, c) for c in ctrs]
    for blocks in (5, 1):
        _check_choices(ctx.autotune(blocks))
    torch.cuda.synchronize()
    for graph in (False, True, True):
        d = torch.empty_like(d_ref)
        ctx.aes_ctr_keystream(ctrs, d, graph=graph)
        torch.cuda.synchronize()
        assert np.array_equal(from_dev(d), ref), graph
    d1 = torch.empty((1, 128, 637), dtype=torch.int32, device="cuda")
    ctx.aes_ctr_keystream(ctrs[:1], d1)
    torch.cuda.synchronize()
    assert np.array_equal(from_dev(d1), ref[:1])


@pytest.fixture(scope="module", params=[1, 2], ids=["one_bit", "two_bits"])
def actx(gpu, okeys, request):
    """The same LWE key with the one-product FP64 PBS (pbs_approx = 1), blind-rotating
    one or two key bits per step (pbs_multibit)."""
    import tfhe_amd
    c = tfhe_amd.CbsContext(max_blocks=3, params=tfhe_amd.cbs_params(pbs_approx=1, pbs_multibit=request.param))
    c.keygen(okeys.s, SEED)
    yield c
    c.close()


def _phase_err(K2, out, bit):
    """Signed phase - bit * 2^(48 + 8j) of both levels of a PBS output (torus-64)."""
    K = K2.astype(np.int64).astype(np.uint64)
    errs = []
    for j in range(2):
        e = (int(out[j, -1]) - int(np.dot(out[j, :-1], K)) - (bit << (48 + 8 * j))) % 2 ** 64
        errs.append(e - 2 ** 64 if e >= 2 ** 63 else e)
    return errs


def _oracle_pbs_many(okeys, cts, mb):
    """Exact oracle PBS (one or two key bits per step) of many inputs on the host cores
    (the C calls release the GIL)."""
    from concurrent.futures import ThreadPoolExecutor
    import os
    fn = okeys.pbs_mb2 if mb == 2 else okeys.pbs
    if mb == 2:
        okeys.bsk_mb2_fourier()  # build the pattern key once, before the threads
    with ThreadPoolExecutor(max(1, min(16, os.cpu_count() or 1))) as pool:
        res = list(pool.map(fn, list(cts)))
    assert all(err < 0.25 for _, err in res)
    return [out for out, _ in res]


def test_pbs_approx_within_noise_bound(actx, okeys):
    """The one-product PBS (one or two key bits per step) against the oracle's exact
    PBS of the same key form on the same 300 inputs, and on 4 of them in the
    small-launch form (for two bits: the wide key batch at one wave per SIMD).
    Its FP64 rounding (about 2^-33 of the torus per bootstrap) can move a gadget
    digit by one, after which the words differ from the exact form's while the
    phase stays within the PBS noise.  The bound, per level and input: the GPU phase
    lies within 2^-24 of the torus of the message and of the oracle's exact phase;
    the rms of the differences stays below 2^-26.
    Noise model.  Both forms' noise is dominated by errors in the accumulator's mask
    times K2 (N/2 key coefficients add up): in the exact form the gadget rounding
    residual (up to 2^-37) of every step with a non-zero key pattern, 2^35.6-2^35.9 of
    2^64 rms; the one-product form adds its FP64 product error (about 2^-38.6 rms per
    step and RLWE component) on every product, whatever the key bit, so one bit per
    step adds 2 x 636 such terms and two bits 6 x 318 (three pattern products per
    pair, each times X^e - 1).  Expected GPU / oracle rms: about 1.3 (one bit) and
    1.5-1.9 (two bits); asserted: below 2.2 and above 1/1.5, with the GPU's rms
    below 2^37 (2^-27 of the torus).  The per-step agreement that pins the pattern
    factors and key rows is test_pbs_few_steps_phase_matches_oracle.  Two bits: the
    4-input launch runs split over 8 waves (the size model's form for launches of at most
    one workgroup per CU), inside the same bounds; the 4-input wide launch equals the
    first 4 outputs of the 300-input launch bit for bit."""
    import torch
    import tfhe_amd
    mb = actx.p.pbs_multibit
    batch = 300
    rng = np.random.default_rng(29)
    bits = rng.integers(0, 2, batch)
    cts = np.stack([tfhe_amd.lwe_encrypt(okeys.s, 33, k, int(bits[k]), p=2) for k in range(batch)])
    d_cts = to_dev32(cts)
    d_big = torch.empty((batch, 2, 2049), dtype=torch.int64, device="cuda")
    actx.pbs_batch(d_cts, d_big)
    d_small = torch.empty((4, 2, 2049), dtype=torch.int64, device="cuda")
    actx.pbs_batch(d_cts[:4].contiguous(), d_small)
    torch.cuda.synchronize()
    big, small = from_dev(d_big, np.uint64), from_dev(d_small, np.uint64)
    exact = _oracle_pbs_many(okeys, cts, mb)
    gpu_e = np.array([_phase_err(okeys.K2, big[k], int(bits[k])) for k in range(batch)], np.float64)
    gpu_s = np.array([_phase_err(okeys.K2, small[k], int(bits[k])) for k in range(4)], np.float64)
    ora_e = np.array([_phase_err(okeys.K2, exact[k], int(bits[k])) for k in range(batch)], np.float64)
    diff = np.concatenate([(gpu_e - ora_e).ravel(), (gpu_s - ora_e[:4]).ravel()])
    rms = lambda v: float(np.sqrt(np.mean(np.square(v))))
    print(f"PBS ({mb} key bit(s) per step): GPU noise max 2^{np.log2(np.abs(gpu_e).max()):.2f} "
          f"rms 2^{np.log2(rms(gpu_e)):.2f}; exact oracle max 2^{np.log2(np.abs(ora_e).max()):.2f} "
          f"rms 2^{np.log2(rms(ora_e)):.2f}; GPU - oracle max 2^{np.log2(np.abs(diff).max()):.2f} "
          f"rms 2^{np.log2(rms(diff)):.2f} (of 2^64)")
    assert np.abs(gpu_e).max() < 2 ** 40 and np.abs(gpu_s).max() < 2 ** 40 and np.abs(ora_e).max() < 2 ** 40
    assert np.abs(diff).max() < 2 ** 40
    assert rms(diff) < 2 ** 38
    assert rms(gpu_e) < 2 ** 37
    assert rms(gpu_e) < 2.2 * rms(ora_e) and rms(ora_e) < 1.5 * rms(gpu_e)
    if mb == 2:
        actx.set_pbs_form(4, 1)  # the wide key batch: the same sums as two waves per SIMD
        d_wide = torch.empty((4, 2, 2049), dtype=torch.int64, device="cuda")
        actx.pbs_batch(d_cts[:4].contiguous(), d_wide)
        torch.cuda.synchronize()
        actx.set_pbs_form(4, -1)
        assert np.array_equal(from_dev(d_wide, np.uint64), big[:4])


def test_pbs_few_steps_phase_matches_oracle(actx, okeys):
    """LWE inputs whose mask is zero except for 1-3 random key-bit pairs: the blind
    rotation then makes only that many non-trivial steps (a zero rotation index gives
    the factor X^0 - 1 = 0), so the one-product GPU output must agree with the
    oracle's exact PBS of the same key form to within a few steps of FP64 rounding
    plus rare gadget-digit moves (2^28 of 2^64 each).  Measured on MI355X: the first
    step (zero mask in) adds 2^25 rms, each later step about 2^31.2 rms (the product
    error of 2^-38.4 per mask coefficient through K2's N/2 ones), up to 6 key bits here:
    rms 2^32.9, max 2^34.6 over 192 phases.  Asserted: every phase within 2^36 of 2^64
    (2^-28 of the torus; beyond 8 sigma of 6 steps).  A wrong pattern factor, key row or
    key pair would miss by at least a gadget level, 2^44 and more.  Every launch form: 96
    inputs in each form of the context (one bit: branchy and branch-free FFT steps; two bits:
    two waves per SIMD, wide, split over 8 waves), and 4 in the size model's form."""
    import torch
    mb = actx.p.pbs_multibit
    n = okeys.p.n
    batch = 96
    rng = np.random.default_rng(31)
    cts = np.zeros((batch, n + 1), np.uint32)
    for k in range(batch):
        pairs = rng.choice(n // 2, size=1 + k % 3, replace=False)
        for q in pairs:
            for i in (2 * q, 2 * q + 1):
                if rng.integers(0, 4) or i == 2 * q:  # sometimes only one bit of the pair rotates
                    cts[k, i] = rng.integers(1 << 22, (1 << 32) - (1 << 22))
        cts[k, n] = rng.integers(0, 1 << 32)
    d_cts = to_dev32(cts)
    outs = []
    for b, form in [(batch, f) for f in ((0, 1, 2) if mb == 2 else (0, 1))] + [(4, -1)]:
        actx.set_pbs_form(b, form)
        d_o = torch.empty((b, 2, 2049), dtype=torch.int64, device="cuda")
        actx.pbs_batch(d_cts[:b].contiguous(), d_o)
        torch.cuda.synchronize()
        actx.set_pbs_form(b, -1)
        outs.append((from_dev(d_o, np.uint64), b))
    exact = _oracle_pbs_many(okeys, cts, mb)
    K = okeys.K2.astype(np.int64).astype(np.uint64)

    def phases(o):
        return [(int(o[j, -1]) - int(np.dot(o[j, :-1], K))) % 2 ** 64 for j in range(2)]

    worst = 0
    for got, cnt in outs:
        for k in range(cnt):
            for pg, pe in zip(phases(got[k]), phases(exact[k])):
                d = (pg - pe) % 2 ** 64
                worst = max(worst, min(d, 2 ** 64 - d))
    print(f"few-step PBS ({mb} key bit(s) per step): max |GPU - oracle phase| 2^{np.log2(max(worst, 1)):.2f} of 2^64")
    assert worst < 2 ** 36


def _many_lut_inputs(okeys, count, seed):
    """LWE encryptions of random phases (uniform torus-32 messages) plus one input with a
    zero mask; a random torus-64 test polynomial."""
    import tfhe_amd
    rng = np.random.default_rng(seed)
    n = okeys.p.n
    cts = np.stack([tfhe_amd.lwe_encrypt(okeys.s, seed, k, int(rng.integers(0, 2 ** 16)), p=2 ** 16)
                    for k in range(count - 1)] + [np.concatenate([np.zeros(n, np.uint32),
                                                                  [np.uint32(rng.integers(0, 2 ** 32))]])])
    lut = rng.integers(0, 2 ** 64, okeys.p.N, dtype=np.uint64)
    return cts.astype(np.uint32), lut


def _rlwe_phase64(K2, r):
    """phase of a torus-64 RLWE [2][N] under K2 (negacyclic product, exact mod 2^64)."""
    return O.rlwe_phase64(K2, r.reshape(-1))


@pytest.mark.parametrize("bw", [0, 1, 3])
def test_bootstrap_many_lut_bit_exact(ctx, okeys, bw):
    """tfhe_cbs_bootstrap_many_lut (gpu_BootstrapManyLUT) in the exact two-limb form
    equals the oracle's restatement word for word on a random test polynomial, for
    bitwidths 0, 1, 3, including a zero-mask input (no rotation steps)."""
    import torch
    cts, lut = _many_lut_inputs(okeys, 6, 50 + bw)
    d_out = torch.empty((len(cts), 2, okeys.p.N), dtype=torch.int64, device="cuda")
    ctx.bootstrap_many_lut(to_dev32(cts), to_dev64(lut), bw, d_out)
    torch.cuda.synchronize()
    got = from_dev(d_out, np.uint64)
    for k in range(len(cts)):
        want, err = okeys.bootstrap_many_lut(cts[k], lut, bw)
        assert err < 0.25
        assert np.array_equal(got[k], want), f"input {k}, bitwidth {bw}"


def test_bootstrap_many_lut_one_product_phase(actx, okeys):
    """The one-product forms (one and two key bits per step) of gpu_BootstrapManyLUT:
    every coefficient's phase within 2^40 of 2^64 (2^-24 of the torus) of the exact
    oracle's (both carry the PBS noise, about 2^36 rms), and each output's phase equals
    LUT X^-phase~ plus Gpow_j / 2 at j < 2, phase~ from the plaintext secret, within
    the same bound: coefficient 0 is LUT[phase~]."""
    import torch
    mb = actx.p.pbs_multibit
    N = okeys.p.N
    cts, lut = _many_lut_inputs(okeys, 5, 70)
    d_out = torch.empty((len(cts), 2, N), dtype=torch.int64, device="cuda")
    actx.bootstrap_many_lut(to_dev32(cts), to_dev64(lut), 1, d_out)
    torch.cuda.synchronize()
    got = from_dev(d_out, np.uint64)
    s = okeys.s.astype(np.int64)
    worst = 0
    for k in range(len(cts)):
        want, _ = okeys.bootstrap_many_lut(cts[k], lut, 1, mb=mb)
        pg, pw = _rlwe_phase64(okeys.K2, got[k]), _rlwe_phase64(okeys.K2, want)
        d = (pg.astype(np.uint64) - pw.astype(np.uint64)).view(np.int64)
        worst = max(worst, int(np.abs(d).max()))
        # closed form of the message: LUT X^-phase~ with the switched phase from s
        at = [O.special_ms(int(a), N, 1) for a in cts[k][:-1]]
        ph = (O.special_ms(int(cts[k][-1]), N, 1) - int(np.dot(np.array(at, np.int64), s))) % (2 * N)
        m = (np.arange(N) + ph) % (2 * N)
        msg = np.where(m < N, lut[m % N], (0 - lut[m % N].astype(np.uint64)))
        msg = msg.astype(np.uint64)
        msg[0] += np.uint64(1 << 47)
        msg[1] += np.uint64(1 << 55)
        e = (pg.astype(np.uint64) - msg).view(np.int64)
        assert np.abs(e).max() < 2 ** 40, f"input {k}"
    print(f"ManyLUT ({mb} key bit(s) per step): max |GPU - oracle phase| 2^{np.log2(max(worst, 1)):.2f} of 2^64")
    assert worst < 2 ** 40


def test_cbs_batch_to_device_view(ctx, okeys):
    """tfhe_cbs_batch_to_device (gpu_CircuitBootstrappingBatchToDevice): host ciphertexts
    in, a view of the context's outputs equal to tfhe_cbs_batch's words; an empty batch
    gives an empty view."""
    import ctypes
    import torch
    rng = np.random.default_rng(8)
    cts = np.stack([okeys.encrypt_bit(int(b), 900 + k) for k, b in enumerate(rng.integers(0, 2, 3))])
    d_ref = torch.empty((3, 4, 2, okeys.p.N), dtype=torch.int32, device="cuda")
    ctx.cbs_batch(to_dev32(cts), d_ref)
    v = ctx.cbs_batch_to_device(cts)
    torch.cuda.synchronize()
    assert v.batch == 3 and v.digits == 2 and v.N == okeys.p.N and v.per_ciphertext_elems == 4 * 2 * okeys.p.N
    view = np.zeros(3 * v.per_ciphertext_elems, np.uint32)
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    assert hip.hipMemcpy(view.ctypes.data, v.d_rgsw, view.nbytes, 2) == 0  # device to host
    assert np.array_equal(view, from_dev(d_ref).reshape(-1))
    e = ctx.cbs_batch_to_device(np.zeros((0, okeys.p.n + 1), np.uint32))
    assert e.batch == 0 and not e.d_rgsw


def test_pbs_two_bit_pattern_key_equals_oracle(gpu, okeys):
    """The GPU generator of the two-bit pattern key (tfhe_gen_bsk64, pbs_multibit = 2:
    RGSWs of the indicators of (1,0), (0,1), (1,1) per key-bit pair, Philox domain 9)
    writes the oracle's key bit for bit, all 318 x 3 RGSWs."""
    import torch
    import tfhe_amd
    p = tfhe_amd.cbs_params(pbs_approx=1, pbs_multibit=2)
    want = okeys.bsk_mb2()
    d_bsk = torch.empty(want.size, dtype=torch.int64, device="cuda")
    d_s = to_dev32(okeys.s.view(np.uint32))
    d_K2 = to_dev32(okeys.K2.view(np.uint32))
    tfhe_amd.check(tfhe_amd.lib().tfhe_gen_bsk64(p, SEED, tfhe_amd._ptr(d_s), tfhe_amd._ptr(d_K2),
                                                 tfhe_amd._ptr(d_bsk), None))
    torch.cuda.synchronize()
    assert np.array_equal(from_dev(d_bsk, np.uint64), want)


def test_aes128_approx_pbs_form(actx, okeys):
    """Homomorphic AES-128 with the one-product PBS: the FIPS-197 C.1 vector and a
    seeded random block decrypt bit-exactly, phases inside half the margin."""
    import torch
    import tfhe_amd
    key = bytes(range(16))
    actx.set_round_keys(tfhe_amd.aes_encrypt_key_bits(okeys.s, 77, key))
    blocks = [bytes.fromhex("00112233445566778899aabbccddeeff"), bytes(np.random.default_rng(5).integers(0, 256, 16, dtype=np.uint8))]
    d_ks = torch.empty((2, 128, 637), dtype=torch.int32, device="cuda")
    actx.aes_ctr_keystream(blocks, d_ks)
    torch.cuda.synchronize()
    got = from_dev(d_ks)
    dec = _decrypt_blocks(okeys, got, 2)
    assert dec[0].hex() == "69c4e0d86a7b0430d8cdb78070b4c55a"
    assert dec[1] == O.aes128_encrypt(key, blocks[1])
    s64 = okeys.s.astype(np.uint64)
    ph = (got[:, :, -1].astype(np.uint64) - (got[:, :, :-1].astype(np.uint64) @ s64)) % 2 ** 32
    err = ((ph + 2 ** 30) % 2 ** 31).astype(np.int64) - 2 ** 30
    assert np.abs(err).max() < 2 ** 29
    d_graph = torch.empty_like(d_ks)  # the captured keystream of this form equals its eager one
    for _ in range(2):
        actx.aes_ctr_keystream(blocks, d_graph, graph=True)
        torch.cuda.synchronize()
        assert np.array_equal(from_dev(d_graph), got)
    # tuned shapes (the one-product PBS may change FFT form, which moves rounding
    # but not decryption): the same blocks again decrypt to the same AES output
    _check_choices(actx.autotune(2), (0, 1, 2) if actx.p.pbs_multibit == 2 else (0, 1))
    d_tuned = torch.empty_like(d_ks)
    actx.aes_ctr_keystream(blocks, d_tuned)
    torch.cuda.synchronize()
    assert _decrypt_blocks(okeys, from_dev(d_tuned), 2) == dec


@pytest.fixture(scope="module")
def lctx(gpu, okeys):
    """The same keys with the linear-key PFKS (pfks_linear = 1, int8 GEMM on the matrix cores)."""
    import tfhe_amd
    c = tfhe_amd.CbsContext(max_blocks=33, params=tfhe_amd.cbs_params(pfks_linear=1))
    c.keygen(okeys.s, SEED)
    yield c
    c.close()


@pytest.fixture(scope="module")
def kctx(gpu, okeys):
    """The same keys with the linear-key N -> n key switch (ks_linear = 1, int8 GEMM on the
    matrix cores) and the linear PFKS; 65 blocks, so that one key switch of 8320 inputs
    runs as two GEMM chunks (8192 + 128)."""
    import tfhe_amd
    c = tfhe_amd.CbsContext(max_blocks=65, params=tfhe_amd.cbs_params(pfks_linear=1, ks_linear=1))
    c.keygen(okeys.s, SEED)
    yield c
    c.close()


def test_ks2_linear_bit_exact(kctx, okeys):
    """The linear-key key switch equals the oracle's restatement (orc_ks2_linear32) bit for
    bit: 8320 random torus-32 inputs in one call (checked on both sides of the 8192-input
    chunk boundary and at random), then calls of 1 and 129 inputs.  The oracle's key rows
    are regenerated from the seed, so this also pins the GPU key generator."""
    import torch
    rng = np.random.default_rng(43)
    nb = 8320
    lwe32 = rng.integers(0, 2 ** 32, (nb, okeys.p.N + 1), dtype=np.uint64).astype(np.uint32)
    d_in = to_dev32(lwe32)
    d_out = torch.empty((nb, okeys.p.n + 1), dtype=torch.int32, device="cuda")
    kctx.keyswitch_batch(d_in, nb, d_out)
    torch.cuda.synchronize()
    got = from_dev(d_out)
    idx = [0, 1, 127, 128, 4095, 8190, 8191, 8192, 8193, 8319] + [int(k) for k in rng.integers(0, nb, 6)]
    for k in idx:
        assert np.array_equal(got[k], okeys.ks2_linear32(lwe32[k])), k
    for cnt in (1, 129):
        d_small = torch.full((cnt, okeys.p.n + 1), 5, dtype=torch.int32, device="cuda")
        kctx.keyswitch_batch(d_in[nb - cnt:].contiguous(), cnt, d_small)
        torch.cuda.synchronize()
        assert np.array_equal(from_dev(d_small), got[nb - cnt:])


def test_ks2_linear_key_file_and_aes(kctx, okeys, tmp_path):
    """A key file of the linear-KS context (version 2: the parameter set holds ks_linear)
    carries the key-switching key as the oracle's rows [N][8][n+1]; loaded into a fresh
    context it gives the same key switch words; AES-128 (FIPS-197 C.1 and random blocks)
    through the linear key switch decrypts exactly."""
    import struct
    import torch
    import tfhe_amd
    path = tmp_path / "kskl.tfhe"
    kctx.save_keys(path)
    data = open(path, "rb").read()
    assert data[:8] == b"TFHECBSK" and struct.unpack_from("<I", data, 8)[0] == 2
    sizes = struct.unpack_from("<4Q", data, 96)
    assert sizes[3] == okeys.p.N * 8 * (okeys.p.n + 1) * 4
    off = 256 + sum(sizes[:3])
    assert np.array_equal(np.frombuffer(data, np.uint32, okeys.p.N * 8 * (okeys.p.n + 1), off), okeys.kskl)
    c2 = tfhe_amd.CbsContext(max_blocks=1, params=tfhe_amd.cbs_params(pfks_linear=1, ks_linear=1))
    try:
        c2.load_keys(path)
        rng = np.random.default_rng(47)
        lwe32 = rng.integers(0, 2 ** 32, (5, okeys.p.N + 1), dtype=np.uint64).astype(np.uint32)
        outs = []
        for c in (kctx, c2):
            d = torch.empty((5, okeys.p.n + 1), dtype=torch.int32, device="cuda")
            c.keyswitch_batch(to_dev32(lwe32), 5, d)
            torch.cuda.synchronize()
            outs.append(from_dev(d))
        assert np.array_equal(outs[0], outs[1])
    finally:
        c2.close()
    key = bytes(range(16))
    kctx.set_round_keys(tfhe_amd.aes_encrypt_key_bits(okeys.s, 77, key))
    rng = np.random.default_rng(11)
    blocks = [bytes.fromhex("00112233445566778899aabbccddeeff")] + \
        [bytes(rng.integers(0, 256, 16, dtype=np.uint8)) for _ in range(4)]
    d_ks = torch.empty((5, 128, 637), dtype=torch.int32, device="cuda")
    kctx.aes_ctr_keystream(blocks, d_ks)
    torch.cuda.synchronize()
    got = from_dev(d_ks)
    dec = _decrypt_blocks(okeys, got, 5)
    assert dec[0].hex() == "69c4e0d86a7b0430d8cdb78070b4c55a"
    assert dec[1:] == [O.aes128_encrypt(key, b) for b in blocks[1:]]
    s64 = okeys.s.astype(np.uint64)
    ph = (got[:, :, -1].astype(np.uint64) - (got[:, :, :-1].astype(np.uint64) @ s64)) % 2 ** 32
    err = ((ph + 2 ** 30) % 2 ** 31).astype(np.int64) - 2 ** 30
    print(f"AES with the linear key switch: max |phase error| 2^{np.log2(np.abs(err).max()):.2f} of 2^32", flush=True)
    assert np.abs(err).max() < 2 ** 29


def test_pfkl_key_entries_match_oracle(lctx, okeys):
    """GPU linear-key entries (5 signed bytes per 40-bit word, k = 8 i + t) equal the
    oracle's, including the b row (i = N) and both functions."""
    N = okeys.p.N
    for (u, i, t) in [(0, 0, 0), (1, 1, 7), (0, 777, 3), (1, N, 5), (0, N, 0)]:
        assert np.array_equal(lctx.pfkl_entry(u, i, t), okeys.pfkl_entry(u, i, t)), (u, i, t)


def test_pfkl_switch_bit_exact(lctx, okeys):
    """The GEMM-form PFKS of random torus-64 LWE_K2 inputs equals the oracle's
    linear-key switch bit for bit, for each function, at out_stride > 2N."""
    import torch
    N = okeys.p.N
    rng = np.random.default_rng(23)
    lwe = rng.integers(0, 2 ** 64, (3, N + 1), dtype=np.uint64)
    d_lwe = to_dev64(lwe)
    stride = 2 * N + 64
    for u in (0, 1):
        d_out = torch.zeros((3, stride), dtype=torch.int32, device="cuda")
        lctx.pfks_batch(u, d_lwe, 3, stride, d_out)
        torch.cuda.synchronize()
        got = from_dev(d_out)
        for k in range(3):
            assert np.array_equal(got[k, :2 * N], okeys.pfks_linear(u, lwe[k])), (u, k)
        assert not got[:, 2 * N:].any()  # nothing written past the row


@pytest.mark.parametrize("nb", [2000, 8192])
def test_pfkl_staged_gemm_equals_tiled_and_oracle(lctx, okeys, nb):
    """Launches of more than 1792 PFKS inputs take the LDS-staged GEMM (256-input tiles,
    8 waves); smaller ones the 128-input tiles.  nb random inputs in one call (2000: a
    zero-padded last tile; 8192: a whole launch chunk of 32 tiles) equal calls of 1000
    word for word (integer sums: the same words whatever the tiling), and sampled rows,
    including the last one, equal the oracle."""
    import torch
    N = okeys.p.N
    rng = np.random.default_rng(61 + nb)
    lwe = rng.integers(0, 2 ** 64, (nb, N + 1), dtype=np.uint64)
    d_lwe = to_dev64(lwe)
    for u in (0, 1):
        d_big = torch.empty((nb, 2 * N), dtype=torch.int32, device="cuda")
        lctx.pfks_batch(u, d_lwe, nb, 2 * N, d_big)
        d_small = torch.empty((nb, 2 * N), dtype=torch.int32, device="cuda")
        for c0 in range(0, nb, 1000):
            c = min(1000, nb - c0)
            lctx.pfks_batch(u, d_lwe[c0:c0 + c].contiguous(), c, 2 * N, d_small[c0:c0 + c])
        torch.cuda.synchronize()
        big = from_dev(d_big)
        assert np.array_equal(big, from_dev(d_small)), u
        for k in (0, 255, 256, nb - 1):
            assert np.array_equal(big[k], okeys.pfks_linear(u, lwe[k])), (u, k)


def test_cbs_linear_bit_exact(lctx, okeys):
    """A whole circuit bootstrap with the linear PFKS equals the oracle's (the PBS
    and all four RGSW rows)."""
    import torch
    cts = _bits_cts(okeys, [1], 910)
    d_rgsw = torch.empty((1, 4, 2, 2048), dtype=torch.int32, device="cuda")
    lctx.cbs_batch(to_dev32(cts), d_rgsw)
    torch.cuda.synchronize()
    want, err = okeys.cbs(cts[0], linear=True)
    assert err < 0.25
    assert np.array_equal(from_dev(d_rgsw).reshape(-1), want)


def test_cbs_linear_chunked_batch_equals_parts(lctx, okeys):
    """4100 inputs give 8200 PFKS inputs, over the 8192-input launch chunk, and a PBS
    split at 4096: the rows equal two calls of 2050 bit for bit."""
    import torch
    import tfhe_amd
    batch = 4100
    rng = np.random.default_rng(43)
    bits = rng.integers(0, 2, batch)
    cts = np.stack([tfhe_amd.lwe_encrypt(okeys.s, 37, k, int(bits[k]), p=2) for k in range(batch)])
    d_cts = to_dev32(cts)
    d_all = torch.empty((batch, 4, 2, 2048), dtype=torch.int32, device="cuda")
    lctx.cbs_batch(d_cts, d_all)
    torch.cuda.synchronize()
    whole = from_dev(d_all)
    d_part = torch.empty((batch // 2, 4, 2, 2048), dtype=torch.int32, device="cuda")
    for lo in (0, batch // 2):
        lctx.cbs_batch(d_cts[lo:lo + batch // 2].contiguous(), d_part)
        torch.cuda.synchronize()
        assert np.array_equal(from_dev(d_part), whole[lo:lo + batch // 2]), lo
    del d_all, d_part


def test_cbs_linear_staged_pfks_equals_tiled(lctx, okeys):
    """1000 inputs give 2000 PFKS inputs, which take the LDS-staged GEMM with both functions
    in one launch; calls of 250 (500 PFKS inputs) take the 128-input tiles.  The RGSW rows
    are equal bit for bit (the same PBS words, integer key-switch sums)."""
    import torch
    import tfhe_amd
    batch = 1000
    rng = np.random.default_rng(71)
    bits = rng.integers(0, 2, batch)
    cts = np.stack([tfhe_amd.lwe_encrypt(okeys.s, 39, k, int(bits[k]), p=2) for k in range(batch)])
    d_cts = to_dev32(cts)
    d_all = torch.empty((batch, 4, 2, 2048), dtype=torch.int32, device="cuda")
    lctx.cbs_batch(d_cts, d_all)
    torch.cuda.synchronize()
    whole = from_dev(d_all)
    d_part = torch.empty((batch // 4, 4, 2, 2048), dtype=torch.int32, device="cuda")
    for lo in range(0, batch, batch // 4):
        lctx.cbs_batch(d_cts[lo:lo + batch // 4].contiguous(), d_part)
        torch.cuda.synchronize()
        assert np.array_equal(from_dev(d_part), whole[lo:lo + batch // 4]), lo


def test_aes128_linear_pfks(lctx, okeys):
    """Homomorphic AES-128 with the linear PFKS: FIPS-197 C.1 and random blocks
    decrypt exactly with phases inside half the margin; the graph replay equals the
    eager keystream; transciphered plaintexts come back."""
    import torch
    import tfhe_amd
    key = bytes(range(16))
    lctx.set_round_keys(tfhe_amd.aes_encrypt_key_bits(okeys.s, 77, key))
    rng = np.random.default_rng(9)
    blocks = [bytes.fromhex("00112233445566778899aabbccddeeff")] + \
        [bytes(rng.integers(0, 256, 16, dtype=np.uint8)) for _ in range(4)]
    d_ks = torch.empty((5, 128, 637), dtype=torch.int32, device="cuda")
    lctx.aes_ctr_keystream(blocks, d_ks)
    torch.cuda.synchronize()
    got = from_dev(d_ks)
    dec = _decrypt_blocks(okeys, got, 5)
    assert dec[0].hex() == "69c4e0d86a7b0430d8cdb78070b4c55a"
    assert dec[1:] == [O.aes128_encrypt(key, b) for b in blocks[1:]]
    s64 = okeys.s.astype(np.uint64)
    ph = (got[:, :, -1].astype(np.uint64) - (got[:, :, :-1].astype(np.uint64) @ s64)) % 2 ** 32
    err = ((ph + 2 ** 30) % 2 ** 31).astype(np.int64) - 2 ** 30
    assert np.abs(err).max() < 2 ** 29
    d_graph = torch.empty_like(d_ks)
    for _ in range(2):
        lctx.aes_ctr_keystream(blocks, d_graph, graph=True)
        torch.cuda.synchronize()
        assert np.array_equal(from_dev(d_graph), got)
    plain = [bytes(rng.integers(0, 256, 16, dtype=np.uint8)) for _ in range(5)]
    cipher = [bytes(p ^ k for p, k in zip(pt, O.aes128_encrypt(key, c))) for pt, c in zip(plain, blocks)]
    lctx.aes_ctr_transcipher(blocks, cipher, d_ks)
    torch.cuda.synchronize()
    assert _decrypt_blocks(okeys, from_dev(d_ks), 5) == plain


def test_pbs_auto_multibit_matches_fixed_forms(gpu, okeys):
    """Two key bits per step: the wide form (a 128-input launch pinned to it, one wave per
    SIMD and all pattern keys in one batch) and the two-waves-per-SIMD form (8192 inputs)
    give the same words on the same inputs.  pbs_multibit = 3 holds both key forms and
    picks two bits untuned (its launches equal the two-bit context's bit for bit); tuned
    for 1 block it equals the fixed context of the form it picked (the split form: the same
    phases within 2^40 of 2^64); AES through it decrypts."""
    import torch
    import tfhe_amd
    outs = {}
    rng = np.random.default_rng(29)
    bits = rng.integers(0, 2, 64)
    base = _bits_cts(okeys, bits, 1200)
    cts = np.tile(base, (128, 1))  # 8192 inputs from 64 encryptions
    d_cts = to_dev32(cts)
    for mb in (1, 2, 3):
        c = tfhe_amd.CbsContext(max_blocks=64, params=tfhe_amd.cbs_params(pbs_approx=1, pbs_multibit=mb,
                                                                           pfks_linear=1))
        try:
            c.keygen(okeys.s, SEED)
            if mb >= 2:
                c.set_pbs_form(128, {2: 1, 3: 3}[mb])  # the wide key batch
            for nb in (128, 8192):
                d_out = torch.empty((nb, 2, 2049), dtype=torch.int64, device="cuda")
                c.pbs_batch(d_cts[:nb].contiguous(), d_out)
                torch.cuda.synchronize()
                outs[(mb, nb)] = from_dev(d_out, np.uint64)
            c.set_pbs_form(128, -1)
            if mb == 3:
                tuned = c.autotune(1)["pbs_form"][1]  # 1 block: no full round, all 128 in the rest
                assert tuned in (0, 1, 2, 3, 4)
                d_out = torch.empty((128, 2, 2049), dtype=torch.int64, device="cuda")
                c.pbs_batch(d_cts[:128].contiguous(), d_out)
                torch.cuda.synchronize()
                outs[("tuned", tuned)] = from_dev(d_out, np.uint64)
                key = bytes(range(16))
                c.set_round_keys(tfhe_amd.aes_encrypt_key_bits(okeys.s, 77, key))
                ctr = [bytes.fromhex("00112233445566778899aabbccddeeff")]
                d_ks = torch.empty((1, 128, 637), dtype=torch.int32, device="cuda")
                c.aes_ctr_keystream(ctr, d_ks)
                torch.cuda.synchronize()
                assert _decrypt_blocks(okeys, from_dev(d_ks), 1)[0].hex() == "69c4e0d86a7b0430d8cdb78070b4c55a"
        finally:
            c.close()
    assert np.array_equal(outs[(2, 128)], outs[(2, 8192)][:128])  # wide form = two waves per SIMD
    assert np.array_equal(outs[(3, 128)], outs[(2, 128)])
    assert np.array_equal(outs[(3, 8192)], outs[(2, 8192)])
    assert np.array_equal(outs[(1, 128)], outs[(1, 8192)][:128])
    assert not np.array_equal(outs[(1, 128)], outs[(2, 128)])  # one and two bits per step differ in their words
    (tuned, got), = [(k[1], v) for k, v in outs.items() if k[0] == "tuned"]
    if tuned == 4:  # split over 8 waves: other FP64 rounding, the same phases
        for k in range(0, 128, 9):
            for e1, e2 in zip(_phase_err(okeys.K2, got[k], int(bits[k % 64])),
                              _phase_err(okeys.K2, outs[(2, 128)][k], int(bits[k % 64]))):
                assert abs(e1 - e2) < 2 ** 40, k
    else:
        assert np.array_equal(got, outs[(2, 128)] if tuned >= 2 else outs[(1, 128)])


def test_pinned_pbs_forms_give_the_same_words(actx, okeys):
    """tfhe_cbs_set_pbs_form pins the launch form per size.  The two forms give the same
    words on the same launch sizes (301, 300 and 3 inputs) whichever the size model would
    pick: with one key bit per step the branchy and branch-free cross-wave FFT steps
    (rounding at the same points), with two the two-waves-per-SIMD and wide key batches;
    with two bits the split form (8 waves) gives the same phases within 2^40 of 2^64.
    Inputs mix encryptions, zero masks (every step a zero rotation) and masks with one
    rotating key-bit pair.  A form the context's PBS does not have is refused."""
    import torch
    import tfhe_amd
    n = okeys.p.n
    rng = np.random.default_rng(37)
    batch = 301
    cts = np.zeros((batch, n + 1), np.uint32)
    for k in range(batch):
        if k % 5 == 1:  # zero mask, random body
            cts[k, n] = rng.integers(0, 1 << 32)
        elif k % 7 == 3:  # one rotating key-bit pair
            q = int(rng.integers(0, n // 2))
            cts[k, 2 * q] = rng.integers(1 << 22, (1 << 32) - (1 << 22))
            cts[k, n] = rng.integers(0, 1 << 32)
        else:
            cts[k] = tfhe_amd.lwe_encrypt(okeys.s, 41, k, int(rng.integers(0, 2)), p=2)
    d_cts = to_dev32(cts)
    mb = actx.p.pbs_multibit
    with pytest.raises(tfhe_amd.TfheError):
        actx.set_pbs_form(batch, 3 if mb == 2 else 2)  # forms 0, 1 (and 2: split, two bits)
    for nb in (301, 300, 3):
        outs = {}
        for form in (0, 1):
            actx.set_pbs_form(nb, form)
            d_out = torch.full((nb, 2, 2049), 7, dtype=torch.int64, device="cuda")
            actx.pbs_batch(d_cts[:nb].contiguous(), d_out)
            torch.cuda.synchronize()
            outs[form] = from_dev(d_out, np.uint64)
        actx.set_pbs_form(nb, -1)
        assert np.array_equal(outs[0], outs[1])
        if mb == 2:  # the split form: the same phases (other FP64 rounding of the sums)
            actx.set_pbs_form(nb, 2)
            d_out = torch.full((nb, 2, 2049), 7, dtype=torch.int64, device="cuda")
            actx.pbs_batch(d_cts[:nb].contiguous(), d_out)
            torch.cuda.synchronize()
            actx.set_pbs_form(nb, -1)
            split = from_dev(d_out, np.uint64)
            for k in range(0, nb, 7):
                for e1, e2 in zip(_phase_err(okeys.K2, split[k], 0), _phase_err(okeys.K2, outs[0][k], 0)):
                    d = (e1 - e2) % 2 ** 64
                    assert min(d, 2 ** 64 - d) < 2 ** 40, (nb, k)


def test_empty_batches_leave_outputs_untouched(ctx, okeys):
    """Every batched entry point takes a count of 0 as a no-op that returns OK and
    launches nothing: outputs filled with a sentinel keep it."""
    import torch
    import tfhe_amd
    L, P, C = tfhe_amd.lib(), tfhe_amd._ptr, tfhe_amd.check
    ctx.set_round_keys(tfhe_amd.aes_encrypt_key_bits(okeys.s, 77, bytes(range(16))))
    SENT = 0x5A5A5A5A
    lwe = torch.full((2, 637), SENT, dtype=torch.int32, device="cuda")
    pbs = torch.full((2, 2, 2049), SENT, dtype=torch.int64, device="cuda")
    rgsw = torch.full((2, 4, 2, 2048), SENT, dtype=torch.int32, device="cuda")
    rgsw_f = torch.full((2, 8, 1024, 2), 0.5, dtype=torch.float64, device="cuda")
    lut = torch.full((2, 32, 2049), SENT, dtype=torch.int32, device="cuda")
    poly64 = torch.full((2, 2048), SENT, dtype=torch.int64, device="cuda")
    poly32 = torch.full((4, 2048), SENT, dtype=torch.int32, device="cuda")
    state = torch.full((1, 128, 637), SENT, dtype=torch.int32, device="cuda")
    outs = [lwe, pbs, rgsw, rgsw_f, lut, poly64, poly32, state]
    before = [o.clone() for o in outs]
    h = ctx.h
    C(L.tfhe_cbs_batch(h, P(lwe), 0, P(rgsw), None))
    C(L.tfhe_cbs_pbs_batch(h, P(lwe), 0, P(pbs), None))
    C(L.tfhe_cbs_pfks_batch(h, 0, P(pbs), 0, 4096, P(rgsw), None))
    C(L.tfhe_rgsw_to_fourier(h, P(rgsw), 0, P(rgsw_f), None))
    C(L.tfhe_lut8_eval_batch(h, P(rgsw_f), 0, P(poly32), 0, 8, P(lut), None))
    C(L.tfhe_lut8x32_eval_batch(h, P(rgsw_f), 0, P(poly32), P(lut), None))
    C(L.tfhe_cbs_keyswitch_batch(h, P(lut), 0, P(lwe), None))
    C(L.tfhe_polymul2048_batch(P(poly32), P(poly64), 0, P(poly64), None))
    C(L.tfhe_polymul2048_approx_batch(P(poly32), P(poly64), 0, P(poly64), None))
    C(L.tfhe_aes_round_batch(h, 1, P(state), 0, P(state), None))
    C(L.tfhe_lwe_xor_public(P(state), P(state), 0, 636, None))
    empty = np.zeros(1, np.uint8)
    for fn in (L.tfhe_aes_ctr_keystream_batch, L.tfhe_aes_ctr_keystream_graph, L.tfhe_aes_ctr_keystream_plain):
        C(fn(h, empty, 0, P(state), None))
    for mode in (0, 1, 2):
        C(L.tfhe_aes_ctr_transcipher_batch(h, mode, empty, empty, 0, P(state), None))
    torch.cuda.synchronize()
    for k, (o, b) in enumerate(zip(outs, before)):
        assert torch.equal(o, b), k


def test_cbs_error_paths(gpu):
    """Capacity, argument and state errors come back as codes (the reference throws,
    cirbts-batch.cu:97-100); nothing is launched."""
    import torch
    import tfhe_amd
    L = tfhe_amd.lib()
    with pytest.raises(tfhe_amd.TfheError) as e:  # no such PBS precision form
        tfhe_amd.CbsContext(max_blocks=1, params=tfhe_amd.cbs_params(pbs_approx=2))
    assert e.value.code == tfhe_amd.TFHE_E_PARAM
    with pytest.raises(tfhe_amd.TfheError) as e:  # two key bits per step need the one-product form
        tfhe_amd.CbsContext(max_blocks=1, params=tfhe_amd.cbs_params(pbs_approx=0, pbs_multibit=2))
    assert e.value.code == tfhe_amd.TFHE_E_PARAM
    with pytest.raises(tfhe_amd.TfheError) as e:  # above the 4096-block capacity bound
        tfhe_amd.CbsContext(max_blocks=4097)
    assert e.value.code == tfhe_amd.TFHE_E_PARAM
    c = tfhe_amd.CbsContext(max_blocks=1)
    try:
        d = torch.zeros((4, 637), dtype=torch.int32, device="cuda")
        out = torch.zeros((4, 4, 2, 2048), dtype=torch.int32, device="cuda")
        with pytest.raises(tfhe_amd.TfheError) as e:  # no keys yet
            c.cbs_batch(d, out)
        assert e.value.code == tfhe_amd.TFHE_E_STATE
        with pytest.raises(tfhe_amd.TfheError) as e:  # no round keys
            c.aes_ctr_keystream([bytes(16)], torch.zeros((1, 128, 637), dtype=torch.int32, device="cuda"))
        assert e.value.code == tfhe_amd.TFHE_E_STATE
        with pytest.raises(tfhe_amd.TfheError) as e:  # no round keys: nothing to transcipher
            c.aes_ctr_transcipher([bytes(16)], [bytes(16)], torch.zeros((1, 128, 637), dtype=torch.int32,
                                                                         device="cuda"))
        assert e.value.code == tfhe_amd.TFHE_E_STATE
        with pytest.raises(tfhe_amd.TfheError) as e:  # nothing to time without keys
            c.autotune(1)
        assert e.value.code == tfhe_amd.TFHE_E_STATE
        c.keygen(tfhe_amd.lwe_keygen(636, 4), 4)
        with pytest.raises(tfhe_amd.TfheError) as e:
            c.autotune(2)
        assert e.value.code == tfhe_amd.TFHE_E_CAPACITY
        with pytest.raises(tfhe_amd.TfheError) as e:  # 129 inputs in a 1-block (128) context
            tfhe_amd.check(L.tfhe_cbs_batch(c.h, tfhe_amd._ptr(d), 129, tfhe_amd._ptr(out), None))
        assert e.value.code == tfhe_amd.TFHE_E_CAPACITY
        with pytest.raises(tfhe_amd.TfheError) as e:
            tfhe_amd.check(L.tfhe_cbs_keyswitch_batch(c.h, tfhe_amd._ptr(d), 129, tfhe_amd._ptr(d), None))
        assert e.value.code == tfhe_amd.TFHE_E_CAPACITY
        with pytest.raises(tfhe_amd.TfheError) as e:  # function index out of range
            tfhe_amd.check(L.tfhe_cbs_pfks_batch(c.h, 2, tfhe_amd._ptr(d), 1, 4096, tfhe_amd._ptr(out), None))
        assert e.value.code == tfhe_amd.TFHE_E_PARAM
        with pytest.raises(tfhe_amd.TfheError) as e:  # rows closer than 2N apart
            tfhe_amd.check(L.tfhe_cbs_pfks_batch(c.h, 0, tfhe_amd._ptr(d), 1, 4095, tfhe_amd._ptr(out), None))
        assert e.value.code == tfhe_amd.TFHE_E_PARAM
        with pytest.raises(tfhe_amd.TfheError) as e:
            tfhe_amd.check(L.tfhe_aes_round_batch(c.h, 11, tfhe_amd._ptr(d), 1, tfhe_amd._ptr(d), None))
        assert e.value.code == tfhe_amd.TFHE_E_PARAM
        c.cbs_batch(d, out, batch=0)  # empty batch is a no-op
    finally:
        c.close()


def test_key_file_round_trip(gpu, okeys, tmp_path):
    """Key files: a context's keys saved (tfhe_cbs_save_keys) and loaded into a fresh
    context give the same circuit bootstraps bit for bit; a file written from keys made
    elsewhere (tfhe_cbs_write_keys: device BSK2 rows and linear PFKS key from the
    stateless generators, the oracle's KSK2 on the host) does too; the header records
    the parameter set and seed; a flipped byte (content hash), another parameter set and
    a truncated file are refused."""
    import os
    import torch
    import tfhe_amd
    p = tfhe_amd.cbs_params(pfks_linear=1)
    c1 = tfhe_amd.CbsContext(max_blocks=1, params=p)
    c1.keygen(okeys.s, SEED)
    path = tmp_path / "keys.tfhe"
    c1.save_keys(path)
    prm, seed = tfhe_amd.key_file_info(path)
    assert seed == SEED and bytes(prm) == bytes(p)
    cts = np.stack([okeys.encrypt_bit(b, 950 + k) for k, b in enumerate([0, 1, 1])])
    d_in = to_dev32(cts)

    def cbs(c):
        d = torch.empty((3, 4, 2, okeys.p.N), dtype=torch.int32, device="cuda")
        c.cbs_batch(d_in, d)
        torch.cuda.synchronize()
        return from_dev(d)

    want = cbs(c1)
    c2 = tfhe_amd.CbsContext(max_blocks=1, params=p)
    c2.load_keys(path)
    assert np.array_equal(cbs(c2), want)
    # keys made elsewhere: device sections from the stateless generators, KSK2 from the oracle
    L = tfhe_amd.lib()
    d_s, d_K2 = to_dev32(okeys.s.view(np.uint32)), to_dev32(c1.rlwe_sk().view(np.uint32))
    d_bsk = torch.empty(okeys.p.n * 8 * 2 * okeys.p.N, dtype=torch.int64, device="cuda")
    tfhe_amd.check(L.tfhe_gen_bsk64(p, SEED, tfhe_amd._ptr(d_s), tfhe_amd._ptr(d_K2), tfhe_amd._ptr(d_bsk), None))
    d_pfkl = torch.empty(2 * 5 * 2 * okeys.p.N * 16416, dtype=torch.int8, device="cuda")
    tfhe_amd.check(L.tfhe_gen_pfkl40(p, SEED, tfhe_amd._ptr(d_K2), tfhe_amd._ptr(d_pfkl), None))
    path2 = tmp_path / "external.tfhe"
    tfhe_amd.write_keys(path2, p, 0, c1.rlwe_sk(), d_bsk, d_pfkl, okeys.ksk2)
    del d_bsk, d_pfkl
    c3 = tfhe_amd.CbsContext(max_blocks=1, params=p)
    c3.load_keys(path2)
    assert np.array_equal(cbs(c3), want)
    assert tfhe_amd.key_file_info(path2)[1] == 0
    # refusals: a flipped byte in the PFKS section, another parameter set, a truncated file
    size = os.path.getsize(path)
    with open(path, "r+b") as f:
        f.seek(size // 2)
        b = f.read(1)
        f.seek(size // 2)
        f.write(bytes([b[0] ^ 0x10]))
    with pytest.raises(tfhe_amd.TfheError) as e:
        c2.load_keys(path)
    assert e.value.code == tfhe_amd.TFHE_E_SHAPE
    c4 = tfhe_amd.CbsContext(max_blocks=1, params=tfhe_amd.cbs_params(pbs_approx=1, pfks_linear=1))
    with pytest.raises(tfhe_amd.TfheError) as e:
        c4.load_keys(path2)
    assert e.value.code == tfhe_amd.TFHE_E_PARAM
    with open(path2, "r+b") as f:
        f.truncate(size // 3)
    with pytest.raises(tfhe_amd.TfheError) as e:
        c2.load_keys(path2)
    assert e.value.code == tfhe_amd.TFHE_E_SHAPE
    for c in (c1, c2, c3, c4):
        c.close()


def _phase_margin_log2(okeys, got, blocks):
    """log2 of the largest |phase - bit 2^31| over the output bits (torus-32)."""
    s64 = okeys.s.astype(np.uint64)
    ph = (got[:blocks, :, -1].astype(np.uint64) - (got[:blocks, :, :-1].astype(np.uint64) @ s64)) % 2 ** 32
    err = ((ph + 2 ** 30) % 2 ** 31).astype(np.int64) - 2 ** 30
    return float(np.log2(max(int(np.abs(err).max()), 1)))


@pytest.mark.parametrize("form", [0, 1], ids=["sbox_rounds", "ttable_rounds"])
def test_aes128_decrypt_fips197_and_random(lctx, okeys, form):
    """Homomorphic AES-128 decryption (the reference's Td direction, tfhe_aes_decrypt_batch)
    in both round forms: FIPS-197 C.1 and Appendix-B ciphertexts and three seeded random
    blocks (5 blocks: a full-round group plus a tail group in this linear-PFKS context)
    decrypt to the plaintexts of the oracle's inverse cipher, every phase inside the p = 2
    margin (2^30), with the decryption round keys of the equivalent inverse cipher."""
    import torch
    import tfhe_amd
    key = bytes.fromhex("2b7e151628aed2a6abf7158809cf4f3c")
    lctx.set_dec_round_keys(tfhe_amd.aes_encrypt_dec_key_bits(okeys.s, 91, key))
    lctx.set_round_form(form)
    try:
        rng = np.random.default_rng(12)
        cts = [O.aes128_encrypt(key, bytes.fromhex("00112233445566778899aabbccddeeff")),
               bytes.fromhex("3925841d02dc09fbdc118597196a0b32")] + \
              [bytes(rng.integers(0, 256, 16, dtype=np.uint8)) for _ in range(3)]
        d_out = torch.empty((5, 128, okeys.p.n + 1), dtype=torch.int32, device="cuda")
        lctx.aes_decrypt(cts, d_out)
        torch.cuda.synchronize()
        got = from_dev(d_out)
        dec = _decrypt_blocks(okeys, got, 5)
        assert dec[1].hex() == "3243f6a8885a308d313198a2e0370734"
        assert dec == [O.aes128_decrypt(key, c) for c in cts]
        m = _phase_margin_log2(okeys, got, 5)
        print(f"AES decryption, round form {form}: max |phase error| 2^{m:.2f} of 2^32")
        assert m < 30
    finally:
        lctx.set_round_form(0)


def test_aes128_ttable_keystream_and_rounds(lctx, okeys):
    """T-table rounds (tfhe_aes_set_round_form(1), the reference's data flow) in the
    encryption direction: the CTR keystream of FIPS-197 C.1 and a random block decrypts
    to plain AES, eager and graph-replayed (equal words); one decryption round and one
    encryption round in each form decrypt to the same state bits."""
    import torch
    import tfhe_amd
    key = bytes(range(16))
    lctx.set_round_keys(tfhe_amd.aes_encrypt_key_bits(okeys.s, 77, key))
    lctx.set_dec_round_keys(tfhe_amd.aes_encrypt_dec_key_bits(okeys.s, 78, key))
    blocks = [bytes.fromhex("00112233445566778899aabbccddeeff"),
              bytes(np.random.default_rng(6).integers(0, 256, 16, dtype=np.uint8))]
    lctx.set_round_form(1)
    try:
        d_ks = torch.empty((2, 128, okeys.p.n + 1), dtype=torch.int32, device="cuda")
        lctx.aes_ctr_keystream(blocks, d_ks)
        torch.cuda.synchronize()
        got = from_dev(d_ks)
        dec = _decrypt_blocks(okeys, got, 2)
        assert dec[0].hex() == "69c4e0d86a7b0430d8cdb78070b4c55a"
        assert dec[1] == O.aes128_encrypt(key, blocks[1])
        assert _phase_margin_log2(okeys, got, 2) < 30
        d_g = torch.empty_like(d_ks)
        lctx.aes_ctr_keystream(blocks, d_g, graph=True)
        torch.cuda.synchronize()
        assert np.array_equal(from_dev(d_g), got)
        # single rounds from a fresh state of random bits, both forms
        rng = np.random.default_rng(7)
        state = np.stack([okeys.encrypt_bit(int(b), 1200 + k) for k, b in enumerate(rng.integers(0, 2, 128))])
        d_in = to_dev32(state[None])
        outs = {}
        for form in (1, 0):
            lctx.set_round_form(form)
            for name, fn in (("enc", lctx.aes_round), ("dec", lctx.aes_dec_round)):
                d_o = torch.empty((1, 128, okeys.p.n + 1), dtype=torch.int32, device="cuda")
                fn(3, d_in, 1, d_o)
                torch.cuda.synchronize()
                outs[(name, form)] = _decrypt_blocks(okeys, from_dev(d_o), 1)[0]
        assert outs[("enc", 0)] == outs[("enc", 1)]
        assert outs[("dec", 0)] == outs[("dec", 1)]
    finally:
        lctx.set_round_form(0)
