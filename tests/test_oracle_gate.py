"""Pins of the CPU oracle's gate-bootstrapping path against published vectors,
closed forms and brute force (no GPU needed)."""
import ctypes
import random

import numpy as np
import pytest

import oracle as O


@pytest.fixture(scope="module")
def keys():
    return O.GateKeys(1)


# ---------------------------------------------------------------- Philox4x32-10
def test_philox_known_answers():
    # Random123 kat_vectors, philox4x32 with 10 rounds
    assert list(O.philox([0, 0], [0, 0, 0, 0])) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert list(O.philox([0xFFFFFFFF] * 2, [0xFFFFFFFF] * 4)) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert list(O.philox([0xA4093822, 0x299F31D0], [0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344])) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_noise_moments():
    L = O.lib()
    v = np.array([L.orc_noise(7, 99, i // 64, i % 64, 15) for i in range(20000)], np.float64)
    # Irwin-Hall(12) scaled to sigma = 2^17 units; bounded by 6 sigma
    assert abs(v.mean()) < 0.05 * 2 ** 17
    assert abs(v.std() / 2 ** 17 - 1.0) < 0.03
    assert np.abs(v).max() <= 6 * 2 ** 17 + 1


# ---------------------------------------------------------------- AES (FIPS-197)
def test_aes_fips197_vectors():
    # Appendix C.1 (the reference's kNistAes128, aes128.cuh:42-47)
    ct = O.aes128_encrypt(bytes(range(16)), bytes.fromhex("00112233445566778899aabbccddeeff"))
    assert ct.hex() == "69c4e0d86a7b0430d8cdb78070b4c55a"
    # Appendix B cipher example
    key = bytes.fromhex("2b7e151628aed2a6abf7158809cf4f3c")
    ct = O.aes128_encrypt(key, bytes.fromhex("3243f6a8885a308d313198a2e0370734"))
    assert ct.hex() == "3925841d02dc09fbdc118597196a0b32"
    # Appendix A.1 key expansion, last round key w[40..43]
    rk = O.aes128_expand(key)
    assert bytes(rk[160:176]).hex() == "d014f9a8c9ee2589e13f0cc8b6630ca6"


def test_aes_sbox_properties():
    S = [O.aes_sbox(x) for x in range(256)]
    assert S[0x00] == 0x63 and S[0x53] == 0xED and S[0xFF] == 0x16
    assert sorted(S) == list(range(256))
    assert all(S[x] != x and S[x] != x ^ 0xFF for x in range(256))  # no fixed / opposite fixed points


def test_aes_inverse_sbox_and_decryption_vectors():
    """S^-1 inverts S everywhere (FIPS-197 Fig. 14 values at 0x00, 0x63, 0xff); the
    inverse cipher (FIPS-197 5.3) and the Td-table equivalent inverse cipher (5.3.5, the
    reference's decryption data flow) both decrypt Appendix C.1 and B, and random blocks,
    to the plaintext; the Te-table cipher equals the plain cipher."""
    assert all(O.aes_inv_sbox(O.aes_sbox(x)) == x for x in range(256))
    assert O.aes_inv_sbox(0x00) == 0x52 and O.aes_inv_sbox(0x63) == 0x00 and O.aes_inv_sbox(0xFF) == 0x7D
    k1, c1, p1 = bytes(range(16)), "69c4e0d86a7b0430d8cdb78070b4c55a", "00112233445566778899aabbccddeeff"
    k2 = bytes.fromhex("2b7e151628aed2a6abf7158809cf4f3c")
    c2, p2 = "3925841d02dc09fbdc118597196a0b32", "3243f6a8885a308d313198a2e0370734"
    for k, c, p in ((k1, c1, p1), (k2, c2, p2)):
        assert O.aes128_decrypt(k, bytes.fromhex(c)).hex() == p
        assert O.aes128_decrypt_ttable(k, bytes.fromhex(c)).hex() == p
        assert O.aes128_encrypt_ttable(k, bytes.fromhex(p)).hex() == c
    rng = random.Random(17)
    for _ in range(20):
        k = bytes(rng.randrange(256) for _ in range(16))
        m = bytes(rng.randrange(256) for _ in range(16))
        ct = O.aes128_encrypt(k, m)
        assert O.aes128_encrypt_ttable(k, m) == ct
        assert O.aes128_decrypt(k, ct) == m and O.aes128_decrypt_ttable(k, ct) == m


def test_aes_ttables():
    """Te0[x] = (2 S, S, S, 3 S) and Td0[x] = (14 S^-1, 9 S^-1, 13 S^-1, 11 S^-1) big-endian
    (the identities of the reference's aes128.cuh:227-276), Te_r / Td_r = Te0 / Td0 rotated
    right by 8 r bits; the published first entries Te0[0] = c66363a5, Td0[0] = 51f4a750."""
    assert O.aes_ttable(0, 0, 0) == 0xC66363A5 and O.aes_ttable(1, 0, 0) == 0x51F4A750

    def gm(a, b):
        r = 0
        while b:
            if b & 1:
                r ^= a
            a = ((a << 1) ^ (0x1B if a & 0x80 else 0)) & 0xFF
            b >>= 1
        return r
    for x in range(256):
        s, si = O.aes_sbox(x), O.aes_inv_sbox(x)
        te0 = (gm(s, 2) << 24) | (s << 16) | (s << 8) | gm(s, 3)
        td0 = (gm(si, 14) << 24) | (gm(si, 9) << 16) | (gm(si, 13) << 8) | gm(si, 11)
        for r in range(4):
            rot = lambda w: ((w >> (8 * r)) | (w << (32 - 8 * r))) & 0xFFFFFFFF if r else w
            assert O.aes_ttable(0, r, x) == rot(te0) and O.aes_ttable(1, r, x) == rot(td0)


# ---------------------------------------------------------------- decomposition
@pytest.mark.parametrize("logB,l", [(7, 3), (8, 4), (8, 2), (2, 8)])
def test_decompose32_reconstruction(logB, l):
    rng = random.Random(logB * 100 + l)
    ign = 32 - l * logB
    vals = [0, 1, 2 ** 31, 2 ** 31 - 1, 2 ** 32 - 1, 2 ** 31 + 2 ** 30] + [rng.getrandbits(32) for _ in range(3000)]
    if ign > 0:
        vals += [(k << ign) + (1 << (ign - 1)) for k in range(5)]  # rounding ties
    for t in vals:
        d = O.decompose32(t, logB, l)
        assert all(-(1 << (logB - 1)) <= int(x) < (1 << (logB - 1)) for x in d)
        rec = sum(int(d[j]) << (ign + j * logB) for j in range(l)) % 2 ** 32
        err = ((t - rec + 2 ** 31) % 2 ** 32) - 2 ** 31
        assert abs(err) <= (1 << (ign - 1) if ign > 0 else 0), (t, d, err)


@pytest.mark.parametrize("logB,l", [(9, 4), (8, 2), (4, 8)])
def test_decompose64_reconstruction(logB, l):
    rng = random.Random(7 + logB)
    ign = 64 - l * logB
    for t in [0, 2 ** 63, 2 ** 64 - 1] + [rng.getrandbits(64) for _ in range(3000)]:
        d = O.decompose64(t, logB, l)
        assert all(-(1 << (logB - 1)) <= int(x) < (1 << (logB - 1)) for x in d)
        rec = sum(int(d[j]) << (ign + j * logB) for j in range(l)) % 2 ** 64
        err = ((t - rec + 2 ** 63) % 2 ** 64) - 2 ** 63
        assert abs(err) <= (1 << (ign - 1))


# ---------------------------------------------------------------- negacyclic products
def test_negacyclic_monomial_closed_form():
    N = 16
    for a in range(N):
        for b in range(N):
            x = np.zeros(N, np.int32); x[a] = 1
            y = np.zeros(N, np.uint32); y[b] = 1
            out = O.nc_mul_exact32(x, y)
            want = np.zeros(N, np.uint32)
            if a + b < N:
                want[a + b] = 1
            else:
                want[a + b - N] = 0xFFFFFFFF
            assert np.array_equal(out, want)


@pytest.mark.parametrize("N,logB", [(1024, 7), (2048, 9), (2048, 8)])
def test_fft_product_matches_schoolbook(N, logB):
    rng = np.random.default_rng(N + logB)
    half = 1 << (logB - 1)
    for _ in range(3):
        x = rng.integers(-half, half, N).astype(np.int64)
        y = rng.integers(-2 ** 31, 2 ** 31, N).astype(np.int64)
        prod, err = O.nc_mul_fft(x, y)
        exact = O.nc_mul_exact32(x.astype(np.int32), y.astype(np.uint32).astype(np.uint32))
        assert err < 0.25
        assert np.array_equal(prod.astype(np.uint64).astype(np.uint32), exact)


# ---------------------------------------------------------------- key material
def test_bsk_rows_decrypt(keys):
    p, K, s = keys.p, keys.K, keys.s
    rows, N = 2 * p.bk_l, p.N
    bsk = keys.bsk.reshape(p.n, rows, 2, N)
    for i in (0, 1, 317, p.n - 1):
        for r in range(rows):
            j, u = r >> 1, r & 1
            g = 1 << (32 - p.bk_l * p.bk_logB + j * p.bk_logB)
            a, b = bsk[i, r, 0], bsk[i, r, 1]
            ph = (b - O.nc_mul_exact32(K, a)).astype(np.uint32)  # b - a*K
            want = np.zeros(N, np.uint32)
            # row u=0 holds (a + s g, a K + e): phase = e - s g K ; row u=1: phase = e + s g
            if u == 1:
                want[0] = (int(s[i]) * g) & 0xFFFFFFFF
            else:
                want = (-(O.nc_mul_exact32(K, np.array([int(s[i]) * g] + [0] * (N - 1), np.uint32)))).astype(np.uint32)
            err = ((ph - want).astype(np.uint32).view(np.int32)).astype(np.int64)
            assert np.abs(err).max() < 7 * 2 ** 7  # 6 sigma of 2^-25 plus slack


def test_ksk_entries_decrypt(keys):
    p = keys.p
    B = 1 << p.ks_logB
    ksk = keys.ksk.reshape(p.N, p.ks_l, B, p.n + 1)
    for (i, j, v) in [(0, 0, 1), (5, 7, 3), (1023, 3, 2), (400, 0, 0)]:
        mu = (v * int(keys.K[i]) << (32 - p.ks_l * p.ks_logB + j * p.ks_logB)) & 0xFFFFFFFF
        ph = keys.phase(ksk[i, j, v])
        err = ((ph - mu + 2 ** 31) % 2 ** 32) - 2 ** 31
        assert abs(err) <= 6 * 2 ** 17 + 1


# ---------------------------------------------------------------- bootstrap pieces
def test_modswitch_exact_rounding():
    L = O.lib()
    N = 1024
    rng = random.Random(3)
    vals = [0, 2 ** 32 - 1, 1 << 20, (1 << 20) - 1, (1 << 21) * 5 + (1 << 20)] + [rng.getrandbits(32) for _ in range(2000)]
    for v in vals:
        want = ((2 * v * 2 * N + 2 ** 32) // (2 * 2 ** 32)) % (2 * N)  # floor(v*2N/2^32 + 1/2)
        assert L.orc_modswitch_2N(v, N) == want


def test_monic_update_closed_form():
    L = O.lib()
    N = 64
    rng = np.random.default_rng(5)
    for m in [0, 1, 5, 63, 64, 65, 100, 127]:
        t = rng.integers(0, 2 ** 32, 2 * N, dtype=np.uint64).astype(np.uint32)
        acc = rng.integers(0, 2 ** 32, 2 * N, dtype=np.uint64).astype(np.uint32)
        want = acc.copy()
        mono = np.zeros(N, np.int32)
        if m < N:
            mono[m] = 1
        else:
            mono[m - N] = -1
        for o in range(2):
            rot = O.nc_mul_exact32(mono, t[o * N:(o + 1) * N])
            want[o * N:(o + 1) * N] = (acc[o * N:(o + 1) * N] + rot - t[o * N:(o + 1) * N]).astype(np.uint32)
        got = acc.copy()
        L.orc_monic_update32(N, m, t, got)
        assert np.array_equal(got, want)


def test_sample_extract_phase(keys):
    """The extracted LWE has, under the RLWE key as an LWE key, the phase of coefficient idx."""
    L = O.lib()
    p, K = keys.p, keys.K
    N = p.N
    rng = np.random.default_rng(11)
    acc = rng.integers(0, 2 ** 32, 2 * N, dtype=np.uint64).astype(np.uint32)
    ph_poly = (acc[N:] - O.nc_mul_exact32(K, acc[:N])).astype(np.uint32)
    for idx in (0, 1, 511, N - 1):
        lwe = np.zeros(N + 1, np.uint32)
        L.orc_sample_extract32(N, acc, idx, lwe)
        ph = (int(lwe[N]) - int(np.dot(lwe[:N].astype(np.uint64), K.astype(np.uint64)))) % 2 ** 32
        assert ph == int(ph_poly[idx])


def test_keyswitch_preserves_phase(keys):
    L = O.lib()
    p = keys.p
    N = p.N
    rng = np.random.default_rng(12)
    for mu in (0, 1 << 30, 3 << 30):
        a = rng.integers(0, 2 ** 32, N, dtype=np.uint64).astype(np.uint32)
        b = (int(np.dot(a.astype(np.uint64), keys.K.astype(np.uint64))) + mu + 12345) % 2 ** 32
        lweN = np.concatenate([a, np.array([b], np.uint32)])
        out = np.zeros(p.n + 1, np.uint32)
        L.orc_keyswitch32(ctypes.byref(p), keys.ksk, lweN, out)
        err = ((keys.phase(out) - mu + 2 ** 31) % 2 ** 32) - 2 ** 31
        # key noise N*l*sigma^2 -> sigma ~ 2^-8.5 of the torus; rounding of a~ adds < 2^-13
        assert abs(err) < 2 ** 32 / 32


def test_blind_rotation_rotates_test_vector(keys):
    """acc after blind rotation decrypts to X^-<a~,s> * TV (the test vector already folds in b~)."""
    L = O.lib()
    p = keys.p
    N = p.N
    rng = random.Random(21)
    for trial in range(2):
        at = np.array([rng.randrange(2 * N) for _ in range(p.n)], np.uint32)
        bt = rng.randrange(2 * N)
        acc = np.zeros(2 * N, np.uint32)
        L.orc_gate_test_vector(N, O.NAND, bt, acc[N:])
        tv = acc[N:].copy()
        err = L.orc_blind_rotate32(ctypes.byref(p), keys.bsk_f, at, acc)
        assert err < 0.25
        phi = int(np.dot(at.astype(np.int64), keys.s.astype(np.int64))) % (2 * N)
        mono = np.zeros(N, np.int32)
        sh = (2 * N - phi) % (2 * N)
        if sh < N:
            mono[sh] = 1
        else:
            mono[sh - N] = -1
        want = O.nc_mul_exact32(mono, tv)
        ph_poly = (acc[N:] - O.nc_mul_exact32(keys.K, acc[:N])).astype(np.uint32)
        e = ((ph_poly - want).astype(np.uint32).view(np.int32)).astype(np.int64)
        assert np.abs(e).max() < 2 ** 32 / 64


def test_external_product_is_cmux_selector(keys):
    """ExtProd(RGSW(s_i), acc) decrypts to s_i * phase(acc)."""
    L = O.lib()
    p = keys.p
    N, M, rows = p.N, p.N // 2, 2 * p.bk_l
    rng = np.random.default_rng(4)
    msg = rng.integers(0, 8, N).astype(np.uint32) << 29
    a = rng.integers(0, 2 ** 32, N, dtype=np.uint64).astype(np.uint32)
    b = (O.nc_mul_exact32(keys.K, a) + msg).astype(np.uint32)
    acc = np.concatenate([a, b])
    for i in (0, 1, 2, 3):
        key = np.ascontiguousarray(keys.bsk_f.reshape(p.n, rows * 2 * 2 * M)[i])
        out = np.zeros(2 * N, np.uint32)
        err = L.orc_external_product32(N, p.bk_logB, p.bk_l, key, acc, out)
        assert err < 0.25
        ph = (out[N:] - O.nc_mul_exact32(keys.K, out[:N])).astype(np.uint32)
        want = msg * np.uint32(keys.s[i])
        e = ((ph - want).astype(np.uint32).view(np.int32)).astype(np.int64)
        assert np.abs(e).max() < 2 ** 32 / 256


# ---------------------------------------------------------------- gate truth tables
@pytest.mark.parametrize("gate", [O.AND, O.OR, O.NAND, O.NOR, O.XOR, O.XNOR, O.XOR_FAST, O.XNOR_FAST])
def test_gate_truth_table(keys, gate):
    for a in (0, 1):
        for b in (0, 1):
            c1 = keys.encrypt(a, 1000 + 4 * gate + 2 * a + b)
            c2 = keys.encrypt(b, 5000 + 4 * gate + 2 * a + b)
            out, err = keys.bootstrap(gate, c1, c2)
            assert err < 0.25
            assert keys.decrypt(out) == O.GATE_FUNCS[gate](a, b), (gate, a, b)


def test_gate_rejects_unsupported():
    L = O.lib()
    z = np.zeros(637, np.uint32)
    assert L.orc_gate_prep(6, 636, z, z, z) == -1  # MAJORITY is not a two-input bootstrap gate
