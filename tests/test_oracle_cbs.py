"""Pins of the CPU oracle's circuit-bootstrapping level (N = 2048): exact
torus-64 limb products and torus-32 products against schoolbook, the
multi-value PBS (torus-64), private functional key switching to torus-32,
CMux / vertical packing (torus-32) and the 64 -> 32 key switch, each checked
by decryption under the secret keys (no GPU needed)."""
import random

import numpy as np
import pytest

import oracle as O


@pytest.fixture(scope="module")
def keys():
    return O.CbsKeys(2)


def centered64(v):
    return ((int(v) + 2 ** 63) % 2 ** 64) - 2 ** 63


def test_external_product64_exact_against_schoolbook():
    """sum_r dec_r(acc) * C_r computed with two FP64 limb products equals the
    exact integer result mod 2^64 (random torus-64 RGSW rows, N=2048, 2^8 x 2)."""
    L = O._cbs_lib()
    N, logB, l = 2048, 8, 2
    rng = np.random.default_rng(31)
    rows = 2 * l
    rgsw = rng.integers(0, 2 ** 64, rows * 2 * N, dtype=np.uint64)
    acc = rng.integers(0, 2 ** 64, 2 * N, dtype=np.uint64)
    rgsw_f = np.concatenate([O.to_fourier64(rgsw[q * N:(q + 1) * N]) for q in range(rows * 2)])
    out = np.zeros(2 * N, np.uint64)
    err = L.orc_external_product64(N, logB, l, rgsw_f, acc, out)
    assert err < 0.25
    want = np.zeros(2 * N, np.uint64)
    for u in range(2):
        digs = np.array([O.decompose64(int(t), logB, l) for t in acc[u * N:(u + 1) * N]])  # [N][l]
        for j in range(l):
            r = 2 * j + u
            for o in range(2):
                want[o * N:(o + 1) * N] += O.nc_mul_exact64(digs[:, j], rgsw[(r * 2 + o) * N:(r * 2 + o + 1) * N])
    assert np.array_equal(out, want)


def test_external_product32_n2048_exact_against_schoolbook():
    """The torus-32 CMux product at N = 2048 (one FP64 FFT product per digit
    polynomial, 2^8 x 2 digits times full 32-bit RGSW words) equals the exact
    integer result mod 2^32; also pins orc_rgsw_to_fourier32's layout."""
    L = O._cbs_lib()
    N, logB, l = 2048, 8, 2
    rng = np.random.default_rng(32)
    rows = 2 * l
    rgsw = rng.integers(0, 2 ** 32, rows * 2 * N, dtype=np.uint64).astype(np.uint32)
    acc = rng.integers(0, 2 ** 32, 2 * N, dtype=np.uint64).astype(np.uint32)
    rgsw_f = np.zeros(rows * 2 * N, np.float64)
    L.orc_rgsw_to_fourier32(N, rows, rgsw, rgsw_f)
    out = np.zeros(2 * N, np.uint32)
    err = L.orc_external_product32(N, logB, l, rgsw_f, acc, out)
    assert err < 0.25
    want = np.zeros(2 * N, np.uint32)
    for u in range(2):
        digs = np.array([O.decompose32(int(t), logB, l) for t in acc[u * N:(u + 1) * N]])  # [N][l]
        for j in range(l):
            r = 2 * j + u
            for o in range(2):
                want[o * N:(o + 1) * N] += O.nc_mul_exact32(digs[:, j], rgsw[(r * 2 + o) * N:(r * 2 + o + 1) * N])
    assert np.array_equal(out, want)


def test_special_ms_closed_form():
    L = O._cbs_lib()
    rng = random.Random(8)
    for v in [0, 2 ** 32 - 1, 1 << 20, (1 << 20) - 1, 3 << 20] + [rng.getrandbits(32) for _ in range(3000)]:
        for N, bw in ((2048, 1), (2048, 0), (1024, 1)):
            lg = (2 * N).bit_length() - 1
            want = (((v * 2 * N * 2 + (2 ** 32 << bw)) // (2 * 2 ** 32 << bw)) << bw) % (2 * N)
            assert L.orc_special_ms(v, N, bw) == want, (v, N, bw)


def test_rlwe64_encryption_and_bsk_rows(keys):
    p = keys.p
    rows = 2 * p.bk_l
    bsk = keys.bsk.reshape(p.n, rows, 2, p.N)
    for i in (0, p.n - 1):
        for r in range(rows):
            j, u = r >> 1, r & 1
            g = 1 << (64 - p.bk_l * p.bk_logB + j * p.bk_logB)
            ph = O.rlwe_phase64(keys.K2, bsk[i, r].reshape(-1))
            if u == 1:
                want = np.zeros(p.N, np.uint64)
                want[0] = (int(keys.s[i]) * g) % 2 ** 64
            else:
                m = np.zeros(p.N, np.uint64)
                m[0] = (int(keys.s[i]) * g) % 2 ** 64
                want = (-O.nc_mul_exact64(keys.K2, m)).astype(np.uint64)
            err = O.signed64((ph - want).astype(np.uint64))
            assert np.abs(err).max() < 7 * 2 ** 13


@pytest.mark.parametrize("bit", [0, 1])
def test_pbs_manylut_levels(keys, bit):
    """One blind rotation yields both circuit-bootstrapping levels:
    LWE_K2(bit * 2^48) and LWE_K2(bit * 2^56)."""
    for rep in range(2):
        ct = keys.encrypt_bit(bit, 100 + 2 * rep + bit)
        out, err = keys.pbs(ct)
        assert err < 0.25
        for j in range(2):
            ph = keys.lwe64_phase(out[j])
            e = centered64(ph - bit * 2 ** (48 + 8 * j))
            assert abs(e) < 2 ** 40  # noise ~2^-28 of the torus, bound 2^-24


def test_pfks_entries_and_functions(keys):
    """Key entries (torus-32) decrypt to f_u(v kappa_i 2^(4t)) with sigma 2^-30
    noise; the switched RLWE decrypts to f_u(phase) of the torus-64 input."""
    p = keys.p
    N = p.N
    for (u, i, t, v) in [(1, 0, 0, 1), (1, N, 7, 15), (0, 5, 3, 9), (0, N, 0, 2)]:
        ct = keys.pfks_entry(u, i, t, v)
        kappa = int(keys.K2[i]) if i < N else -1
        c = (v * kappa * 2 ** (4 * t)) % 2 ** 32
        m = np.zeros(N, np.uint32)
        m[0] = c
        want = m if u == 1 else (-O.nc_mul_exact32(keys.K2, m)).astype(np.uint32)
        err = ((O.rlwe_phase32(keys.K2, ct) - want).astype(np.uint32)).view(np.int32)
        assert np.abs(err).max() <= 7 * 4  # sigma = 4 units of 2^-32; Irwin-Hall tail is +-6 sigma
        assert np.abs(err).max() > 0
    # a random LWE_K2 of mu = 2^56
    rng = np.random.default_rng(3)
    a = rng.integers(0, 2 ** 64, N, dtype=np.uint64)
    mu = 2 ** 56
    b = (int(np.dot(a, keys.K2.astype(np.uint64))) + mu + 999) % 2 ** 64
    lwe = np.concatenate([a, np.array([b], np.uint64)])
    for u in (0, 1):
        r = keys.pfks(u, lwe)
        ph = O.rlwe_phase32(keys.K2, r)
        m = np.zeros(N, np.uint32)
        m[0] = mu >> 32
        want = m if u == 1 else (-O.nc_mul_exact32(keys.K2, m)).astype(np.uint32)
        err = ((ph - want).astype(np.uint32)).view(np.int32)
        # sqrt(sum over 16392 digits of v^2) * 4 ~ 2^12 units (2^-20 of the torus)
        assert np.abs(err).max() < 2 ** 16


def test_pfks_sum_equals_entry_by_entry(keys):
    """orc_pfks sums the selected entries by linearity (one mask product); on a
    sparse input it equals subtracting each orc_gen_pfks_entry ciphertext, for
    both functions, including the b row (i = N) and a digit-15 word."""
    l2 = np.zeros(keys.p.N + 1, np.uint64)
    l2[3] = np.uint64(0x5 << 32)
    l2[keys.p.N] = np.uint64(0xA3 << 32)
    l2[700] = np.uint64((0xF0000001 << 32) - (1 << 31))  # rounds up to 0xF0000001
    for u in (0, 1):
        want = np.zeros(2 * keys.p.N, np.uint32)
        for i in (3, keys.p.N, 700):
            at = ((int(l2[i]) + (1 << 31)) >> 32) & 0xFFFFFFFF
            for t in range(8):
                v = (at >> (4 * t)) & 15
                if v:
                    want = (want - keys.pfks_entry(u, i, t, v)).astype(np.uint32)
        assert np.array_equal(keys.pfks(u, l2), want), u


M40 = (1 << 40) - 1


def _nc_mul_binary_u64(K, a):
    """Exact negacyclic a * K mod 2^64 for binary K (wrapping uint64 adds)."""
    N = len(K)
    out = np.zeros(N, np.uint64)
    for j in np.nonzero(K)[0]:
        r = np.roll(a, j)
        r[:j] = (np.uint64(0) - r[:j])
        out += r
    return out


def test_pfks_linear_entries_and_functions(keys):
    """Linear-key entries (Z/2^40) decrypt to f_u(kappa_i 2^(4t)) 2^8 with sigma 2^7
    units (2^-33 of the torus); the switched RLWE decrypts to f_u(phase) of the
    torus-64 input on the 32-bit torus."""
    p = keys.p
    N = p.N
    for (u, i, t) in [(1, 0, 0), (1, N, 7), (0, 5, 3), (0, N, 0)]:
        ct = keys.pfkl_entry(u, i, t)
        a, b = ct[:N], ct[N:]
        assert int(a.max()) <= M40 and int(b.max()) <= M40
        kappa = int(keys.K2[i]) if i < N else -1
        c = (((kappa * 2 ** (4 * t)) % 2 ** 32) << 8) & M40
        m = np.zeros(N, np.uint64)
        m[0] = c
        want = m if u == 1 else (np.uint64(0) - _nc_mul_binary_u64(keys.K2, m)) & np.uint64(M40)
        ph = (b - _nc_mul_binary_u64(keys.K2, a)) & np.uint64(M40)
        err = ((ph - want) & np.uint64(M40)).astype(np.int64)
        err = np.where(err >= 2 ** 39, err - 2 ** 40, err)
        assert np.abs(err).max() <= 7 * 2 ** 7  # Irwin-Hall tail is +-6 sigma
        assert np.abs(err).max() > 2 ** 5
    rng = np.random.default_rng(4)
    a = rng.integers(0, 2 ** 64, N, dtype=np.uint64)
    mu = 2 ** 56
    b = (int(np.dot(a, keys.K2.astype(np.uint64))) + mu + 999) % 2 ** 64
    lwe = np.concatenate([a, np.array([b], np.uint64)])
    for u in (0, 1):
        r = keys.pfks_linear(u, lwe)
        ph = O.rlwe_phase32(keys.K2, r)
        m = np.zeros(N, np.uint32)
        m[0] = mu >> 32
        want = m if u == 1 else (-O.nc_mul_exact32(keys.K2, m)).astype(np.uint32)
        err = ((ph - want).astype(np.uint32)).view(np.int32)
        # key noise sqrt(16392 * 21.5) * 2^-33 ~ 2^8.2 units plus the input rounding
        # (half a unit per coefficient, ~2^9 over 2049 coefficients)
        assert np.abs(err).max() < 2 ** 13


def test_pfks_linear_sum_equals_entry_by_entry(keys):
    """orc_pfks_linear sums d * entry by linearity (one mask product mod 2^40); on a
    sparse input (zero coefficients give all-zero signed digits) it equals
    -sum_t d_t KL[i][t] taken entry by entry, rounded to 32 bits, for both functions,
    including the b row (i = N) and words whose top digit carries out."""
    N = keys.p.N
    l2 = np.zeros(N + 1, np.uint64)
    l2[3] = np.uint64(0x5 << 32)
    l2[N] = np.uint64(0xA3 << 32)
    l2[700] = np.uint64((0xF8000009 << 32) - (1 << 31))  # rounds up to 0xF8000009
    for u in (0, 1):
        acc = np.zeros(2 * N, np.uint64)
        for i in (3, N, 700):
            at = ((int(l2[i]) + (1 << 31)) >> 32) & 0xFFFFFFFF
            w = (at + 0x88888888) & 0xFFFFFFFF
            digits = [((w >> (4 * t)) & 15) - 8 for t in range(8)]
            assert sum(d * 16 ** t for t, d in enumerate(digits)) % 2 ** 32 == at
            for t, d in enumerate(digits):
                if d:
                    acc = acc - np.uint64(d % 2 ** 64) * keys.pfkl_entry(u, i, t)
        out40 = acc & np.uint64(M40)
        want = (((out40 + np.uint64(128)) >> np.uint64(8)) & np.uint64(0xFFFFFFFF)).astype(np.uint32)
        assert np.array_equal(keys.pfks_linear(u, l2), want), u


@pytest.mark.parametrize("m", [0, 1])
def test_cmux_selects(keys, m):
    p = keys.p
    N = p.N
    rng = np.random.default_rng(10 + m)
    C = keys.rgsw_fourier(keys.rgsw_encrypt(m, 500))
    msgs = [(rng.integers(0, 2, N).astype(np.uint32) << np.uint32(31)).astype(np.uint32) for _ in range(2)]
    ds = []
    for k, msg in enumerate(msgs):
        a = rng.integers(0, 2 ** 32, N, dtype=np.uint64).astype(np.uint32)
        b = (O.nc_mul_exact32(keys.K2, a) + msg).astype(np.uint32)
        ds.append(np.concatenate([a, b]))
    out = np.zeros(2 * N, np.uint32)
    err = O._cbs_lib().orc_cmux32(N, p.cb_logB, p.cb_l, C, ds[0], ds[1], out)
    assert err < 0.25
    ph = O.rlwe_phase32(keys.K2, out)
    e = ((ph - msgs[m]).astype(np.uint32)).view(np.int32)
    # decomposition keeps 16 of 32 bits: error ~ sqrt(N/6) 2^15 ~ 2^19 typical
    assert np.abs(e).max() < 2 ** 24


def test_vertical_packing_table_lookup(keys):
    """Eight CMux rotations by X^{-8 * 2^i} select table[x] (8 output bits) from a
    table packed as coefficient 8x + j = bit j of table[x]."""
    p = keys.p
    rng = np.random.default_rng(77)
    table = rng.integers(0, 256, 256)
    poly = np.zeros(p.N, np.uint32)
    for x in range(256):
        for j in range(8):
            poly[8 * x + j] = np.uint32(((int(table[x]) >> j) & 1) << 31)
    for x in (0, 0xA7, 255):
        fs = [keys.rgsw_fourier(keys.rgsw_encrypt((x >> i) & 1, 1000 + 8 * i + 64 * x)) for i in range(8)]
        acc, err = keys.vertical_packing(fs, poly)
        assert err < 0.25
        ph = O.rlwe_phase32(keys.K2, acc)
        got = [int(((int(ph[j]) + 2 ** 30) >> 31) & 1) for j in range(8)]
        assert got == [(int(table[x]) >> j) & 1 for j in range(8)], x


def test_ks2_preserves_phase(keys):
    p = keys.p
    N = p.N
    rng = np.random.default_rng(5)
    for bit in (0, 1):
        a = rng.integers(0, 2 ** 64, N, dtype=np.uint64)
        b = (int(np.dot(a, keys.K2.astype(np.uint64))) + (bit << 63) + 12345) % 2 ** 64
        lwe = np.concatenate([a, np.array([b], np.uint64)])
        out = keys.ks2(lwe)
        assert keys.decrypt_bit(out) == bit
        e = ((keys.phase32(out) - (bit << 31) + 2 ** 31) % 2 ** 32) - 2 ** 31
        assert abs(e) < 2 ** 32 / 16


def test_kskl_rows_decrypt(keys):
    """Linear key-switching rows KL[i][j] decrypt under the level-0 key to
    K2_i 2^(16 + 2 j) (mod 2^32) within 6 sigma (sigma = 2^-15 of the torus), and the
    full key holds the same rows."""
    p = keys.p
    rng = np.random.default_rng(8)
    kl = keys.kskl.reshape(p.N, p.ks_l, p.n + 1)
    for i, j in [(0, 0), (p.N - 1, p.ks_l - 1)] + [(int(rng.integers(0, p.N)), int(rng.integers(0, p.ks_l)))
                                                  for _ in range(20)]:
        row = keys.kskl_entry(i, j)
        assert np.array_equal(row, kl[i, j])
        mu = (int(keys.K2[i]) * (1 << (16 + 2 * j))) % 2 ** 32
        e = ((keys.phase32(row) - mu + 2 ** 31) % 2 ** 32) - 2 ** 31
        assert abs(e) < 6 * 2 ** 17


def _balanced_digits(at, B, l):
    """Digits d_j in [-B/2, B/2) with sum_j d_j B^j = at (mod B^l), by carrying."""
    out = []
    for _ in range(l):
        d = ((at + B // 2) % B) - B // 2
        out.append(d)
        at = (at - d) // B
    return out


def test_ks2_linear_equals_signed_digit_sum(keys):
    """The oracle's linear key switch equals an explicit sum over the key rows with digits
    found by carrying (not the offset method): out = (0, b) - sum_ij d_ij KL[i][j] mod 2^32,
    on random torus-32 inputs."""
    p = keys.p
    rng = np.random.default_rng(9)
    kl = keys.kskl.reshape(p.N * p.ks_l, p.n + 1).astype(np.int64)
    B, kbits = 1 << p.ks_logB, p.ks_l * p.ks_logB
    for _ in range(3):
        lwe = rng.integers(0, 2 ** 32, p.N + 1, dtype=np.uint64).astype(np.uint32)
        d = np.zeros(p.N * p.ks_l, np.int64)
        for i in range(p.N):
            at = ((int(lwe[i]) + (1 << (31 - kbits))) >> (32 - kbits)) % (1 << kbits)
            d[i * p.ks_l:(i + 1) * p.ks_l] = _balanced_digits(at, B, p.ks_l)
        assert d.min() >= -B // 2 and d.max() < B // 2
        want = (-(d @ kl)) % 2 ** 32
        want[p.n] = (want[p.n] + int(lwe[p.N])) % 2 ** 32
        assert np.array_equal(keys.ks2_linear32(lwe), want.astype(np.uint32))


def test_ks2_linear_preserves_phase(keys):
    """The linear key switch keeps the phase within the key-switching noise, as the lookup
    form does (the same digit sum, other key rows and signed digits)."""
    p = keys.p
    N = p.N
    rng = np.random.default_rng(5)
    for bit in (0, 1):
        a = rng.integers(0, 2 ** 64, N, dtype=np.uint64)
        b = (int(np.dot(a, keys.K2.astype(np.uint64))) + (bit << 63) + 12345) % 2 ** 64
        lwe = np.concatenate([a, np.array([b], np.uint64)])
        for out in (keys.ks2_linear(lwe), keys.ks2(lwe)):
            assert keys.decrypt_bit(out) == bit
            e = ((keys.phase32(out) - (bit << 31) + 2 ** 31) % 2 ** 32) - 2 ** 31
            assert abs(e) < 2 ** 32 / 16


# ---------------------------------------------------------------- two key bits per step
def test_mb2_patterns_reproduce_the_pair_rotation():
    """For every key-bit pair (b0, b1) and rotations (m1, m2): at most one pattern
    indicator is 1, and 1 + sum_p ind_p (X^(e_p) - 1) with e_p = m1, m2, m1 + m2 is the
    monomial X^(m1 b0 + m2 b1) -- the identity the two-bit step rests on.  Checked on
    exponents mod 2N (X^(2N) = 1 in the negacyclic ring)."""
    L = O._cbs_lib()
    Q2 = 4096
    rng = random.Random(3)
    for b0 in (0, 1):
        for b1 in (0, 1):
            s = np.array([b0, b1], np.int32)
            ind = [int(L.orc_mb2_indicator(s, 0, p)) for p in range(3)]
            assert sum(ind) == (1 if (b0 or b1) else 0)
            for _ in range(50):
                m1, m2 = rng.randrange(Q2), rng.randrange(Q2)
                e = [m1, m2, (m1 + m2) % Q2]
                active = [e[p] for p in range(3) if ind[p]]
                want = (m1 * b0 + m2 * b1) % Q2
                assert (active[0] if active else 0) == want


def test_mb2_pattern_rows_decrypt(keys):
    """Pattern-key rows q = 3k + p, r = 2j + u decrypt to ind_p(s_2k, s_2k+1) * 2^(28 + 9j)
    on component u (u = 1: the body; u = 0: the mask, i.e. -K2 times it)."""
    p = keys.p
    s = keys.s
    for k in (0, 1, 200, p.n // 2 - 1):
        for pat in range(3):
            q = 3 * k + pat
            b0, b1 = int(s[2 * k]), int(s[2 * k + 1])
            ind = [b0 & (1 - b1), (1 - b0) & b1, b0 & b1][pat]
            for r in (0, 1, 6, 7):
                j, u = r >> 1, r & 1
                g = 1 << (28 + 9 * j)
                ph = O.rlwe_phase64(keys.K2, keys.bsk_mb2_row(q, r))
                m = np.zeros(p.N, np.uint64)
                m[0] = ind * g
                want = m if u == 1 else (-O.nc_mul_exact64(keys.K2, m)).astype(np.uint64)
                err = O.signed64((ph - want).astype(np.uint64))
                assert np.abs(err).max() < 7 * 2 ** 13, (k, pat, r)


def test_mb2_full_key_equals_rows(keys):
    """The whole pattern key (OpenMP over rows) holds the rows of the row generator."""
    p = keys.p
    bsk = keys.bsk_mb2().reshape(p.n // 2 * 3, 2 * p.bk_l, 2 * p.N)
    for q, r in ((0, 0), (5, 3), (3 * 317 + 2, 7)):
        assert np.array_equal(bsk[q, r], keys.bsk_mb2_row(q, r))


def test_mb2_blind_rotation_closed_form(keys):
    """After the exact two-bit blind rotation of acc = (0, TV) by rotations a~ the
    phase is X^(-<a~, s>) * TV within the noise (2^-24 of the torus); pairs with one
    or both rotations 0 are included.  A wrong pattern order, indicator or skip rule
    rotates by another exponent and fails by about the size of TV."""
    p = keys.p
    N = p.N
    rng = np.random.default_rng(8)
    at = rng.integers(0, 2 * N, p.n).astype(np.uint32)
    at[0:2] = 0          # both rotations of pair 0 vanish: the step is skipped
    at[2] = 0            # pair 1: first rotation 0
    at[5] = 0            # pair 2: second rotation 0
    tv = (rng.integers(-2 ** 20, 2 ** 20, N).astype(np.int64) << 40).astype(np.uint64)
    acc = np.zeros(2 * N, np.uint64)
    acc[N:] = tv
    err = keys.blind_rotate_mb2(at, acc)
    assert err < 0.25
    phi = int(np.dot(at.astype(np.int64), keys.s.astype(np.int64))) % (2 * N)
    mono = np.zeros(N, np.int32)
    sh = (2 * N - phi) % (2 * N)
    if sh < N:
        mono[sh] = 1
    else:
        mono[sh - N] = -1
    want = O.nc_mul_exact64(mono, tv)
    ph = O.rlwe_phase64(keys.K2, acc)
    e = O.signed64((ph - want).astype(np.uint64))
    assert np.abs(e).max() < 2 ** 40


@pytest.mark.parametrize("bit", [0, 1])
def test_pbs_manylut_mb2_levels(keys, bit):
    """The two-bit PBS gives both circuit-bootstrapping levels like the one-bit PBS:
    LWE_K2(bit * 2^48) and LWE_K2(bit * 2^56), noise below 2^-24 of the torus."""
    ct = keys.encrypt_bit(bit, 140 + bit)
    out, err = keys.pbs_mb2(ct)
    assert err < 0.25
    for j in range(2):
        e = centered64(keys.lwe64_phase(out[j]) - bit * 2 ** (48 + 8 * j))
        assert abs(e) < 2 ** 40


def _many_lut_message(lut, ph, N):
    """LUT X^-ph (negacyclic) plus Gpow_j / 2 = 2^47, 2^55 on coefficients 0 and 1."""
    m = (np.arange(N) + ph) % (2 * N)
    msg = np.where(m < N, lut[m % N], (0 - lut[m % N].astype(np.uint64))).astype(np.uint64)
    msg[0] += np.uint64(1 << 47)
    msg[1] += np.uint64(1 << 55)
    return msg


@pytest.mark.parametrize("bw", [0, 1, 3])
def test_many_lut_zero_mask_is_the_rotated_polynomial(keys, bw):
    """gpu_BootstrapManyLUT restated: with a zero mask no rotation step acts, so the
    output is exactly (0, LUT X^-b~ + offsets) with b~ = round(b 2N / 2^(32+bw)) 2^bw."""
    N, n = keys.p.N, keys.p.n
    rng = np.random.default_rng(90 + bw)
    lut = rng.integers(0, 2 ** 64, N, dtype=np.uint64)
    for b in [0, 1 << 31, int(rng.integers(0, 2 ** 32)), (1 << 32) - 1]:
        ct = np.zeros(n + 1, np.uint32)
        ct[n] = b
        out, _ = keys.bootstrap_many_lut(ct, lut, bw)
        bt = (((b + (1 << (19 + bw))) >> (20 + bw)) << bw) % (2 * N)
        assert not out[0].any()
        assert np.array_equal(out[1], _many_lut_message(lut, bt, N))


@pytest.mark.parametrize("mb", [1, 2])
def test_many_lut_phase_is_lut_at_the_switched_phase(keys, mb):
    """Decryption pins the rotation direction and the test-polynomial placement: the
    phase of every coefficient of the output equals LUT X^-phase~ + offsets, with
    phase~ = b~ - <a~, s> mod 2N from the plaintext secret, within the PBS noise
    (asserted 2^40 of 2^64); one and two key bits per step."""
    N, n = keys.p.N, keys.p.n
    rng = np.random.default_rng(95 + mb)
    lut = rng.integers(0, 2 ** 64, N, dtype=np.uint64)
    s = keys.s.astype(np.int64)
    for k in range(2):
        ct = keys.encrypt_bit(k, 700 + k)
        ct[n] = (int(ct[n]) + int(rng.integers(0, 2 ** 32))) % 2 ** 32  # a random phase
        out, err = keys.bootstrap_many_lut(ct, lut, 1, mb=mb)
        assert err < 0.25
        at = np.array([O.special_ms(int(a), N, 1) for a in ct[:n]], np.int64)
        ph = (O.special_ms(int(ct[n]), N, 1) - int(np.dot(at, s))) % (2 * N)
        e = (O.rlwe_phase64(keys.K2, out.reshape(-1)) - _many_lut_message(lut, ph, N)).view(np.int64)
        assert np.abs(e).max() < 2 ** 40


@pytest.mark.parametrize("bit", [0, 1])
def test_many_lut_with_the_cb_polynomial_gives_the_levels(keys, bit):
    """With the circuit bootstrap's polynomial (-2^47, -2^55 alternating) and b + 2^30,
    coefficient j of gpu_BootstrapManyLUT's output decrypts to bit * 2^(48 + 8j), as the
    multi-value PBS's level j does."""
    N, n = keys.p.N, keys.p.n
    lut = np.array([(0 - (1 << (47 + 8 * (k & 1)))) % 2 ** 64 for k in range(N)], np.uint64)
    ct = keys.encrypt_bit(bit, 800 + bit)
    ct[n] = (int(ct[n]) + (1 << 30)) % 2 ** 32
    out, _ = keys.bootstrap_many_lut(ct, lut, 1)
    ph = O.rlwe_phase64(keys.K2, out.reshape(-1))
    for j in range(2):
        e = centered64(int(ph[j]) - bit * (1 << (48 + 8 * j)))
        assert abs(e) < 2 ** 40


# ---------------------------------------------------------------- HomTrace + SchemeSwitch
@pytest.fixture(scope="module")
def hkeys():
    """Keys of the HomTrace + SchemeSwitch form: PBS 2^10 x 4, BSK2 / HT / SS sigma 2^-54."""
    return O.CbsKeys(5, p=O.cbs_params_htss())


def _centered64(v):
    return np.asarray(v, np.uint64).view(np.int64)


def test_automorphism_is_a_ring_map():
    """sigma_k(x y) = sigma_k(x) sigma_k(y) in Z/2^64[X]/(X^N + 1) (checked against the
    exact schoolbook product), sigma_k sigma_k' = sigma_(k k' mod 2N), sigma_1 = id."""
    N = 2048
    rng = np.random.default_rng(4)
    x = rng.integers(-8, 8, N).astype(np.int64).astype(np.uint64)
    y = rng.integers(0, 2 ** 64, N, dtype=np.uint64)
    xy = O.nc_mul_exact64(x.view(np.int64).astype(np.int32), y)
    for k in (3, 5, 1025, 2049, 4095):
        sx, sy = O.automorphism64(x, k), O.automorphism64(y, k)
        assert np.array_equal(O.automorphism64(xy, k), O.nc_mul_exact64(sx.view(np.int64).astype(np.int32), sy))
    assert np.array_equal(O.automorphism64(O.automorphism64(y, 3), 5), O.automorphism64(y, 15))
    assert np.array_equal(O.automorphism64(y, 1), y)


def test_homtrace_of_a_trivial_ciphertext_is_the_trace(hkeys):
    """With a zero mask every key switch of the trace adds exactly zero, so HomTrace of
    (0, m) is exactly (0, N m_0 at coefficient 0): the trace identity, which pins the
    automorphism sequence X -> X^(N/2^k + 1), k = 0 .. 10."""
    N = hkeys.p.N
    m = np.random.default_rng(8).integers(0, 2 ** 64, N, dtype=np.uint64)
    out, _ = hkeys.homtrace(np.concatenate([np.zeros(N, np.uint64), m]))
    want = np.zeros(N, np.uint64)
    want[0] = np.uint64((int(m[0]) * N) % 2 ** 64)
    assert not out[0].any() and np.array_equal(out[1], want)


def test_ht_and_ss_key_rows_decrypt(hkeys):
    """HT key entry (k, d) decrypts to -sigma_k(K2) 2^(24 + 8d), SS entry d to K2^2 2^(24 + 8d),
    within the key noise (Irwin-Hall, sigma 2^10 of 2^64, bounded at 6 sigma)."""
    N = hkeys.p.N
    hkeys._htss_keys()
    K64 = hkeys.K2.astype(np.int64).astype(np.uint64)
    sq = O.nc_mul_exact64(hkeys.K2, K64)
    for k, d in [(0, 0), (3, 2), (10, 4)]:
        ent = hkeys.ht_key[((k * O.HT_L + d) * 2) * N:((k * O.HT_L + d) * 2 + 2) * N]
        want = (0 - O.automorphism64(K64, (N >> k) + 1) * np.uint64(1 << (24 + 8 * d))).astype(np.uint64)
        assert np.abs(_centered64(O.rlwe_phase64(hkeys.K2, ent) - want)).max() <= 6 * 2 ** 10
    for d in range(O.HT_L):
        ent = hkeys.ss_key[d * 2 * N:(d + 1) * 2 * N]
        want = (sq * np.uint64(1 << (24 + 8 * d))).astype(np.uint64)
        assert np.abs(_centered64(O.rlwe_phase64(hkeys.K2, ent) - want)).max() <= 6 * 2 ** 10


def _rlwe_encrypt64(keys, m, seed):
    rng = np.random.default_rng(seed)
    N = keys.p.N
    a = rng.integers(0, 2 ** 64, N, dtype=np.uint64)
    e = rng.integers(-2 ** 10, 2 ** 10, N).astype(np.int64).astype(np.uint64)
    return np.concatenate([a, (O.nc_mul_exact64(keys.K2, a) + m + e).astype(np.uint64)])


def test_homtrace_and_schemeswitch_decrypt(hkeys):
    """HomTrace of an RLWE of m decrypts to N m_0 at coefficient 0 and 0 elsewhere;
    SchemeSwitch of an RLWE of m decrypts to -K2 m; both within 2^40 of 2^64 (the
    trace doubles the input noise of coefficient 0 per round, 2^10 x 2^11, and adds the
    key switches' rounding, 2^-41 of the torus through K2)."""
    N = hkeys.p.N
    rng = np.random.default_rng(12)
    m = rng.integers(0, 2 ** 64, N, dtype=np.uint64)
    ht, err = hkeys.homtrace(_rlwe_encrypt64(hkeys, m, 1))
    assert err < 0.25
    want = np.zeros(N, np.uint64)
    want[0] = np.uint64((int(m[0]) * N) % 2 ** 64)
    e = _centered64(O.rlwe_phase64(hkeys.K2, ht.reshape(-1)) - want)
    assert np.abs(e).max() < 2 ** 40
    ss, err = hkeys.schemeswitch(_rlwe_encrypt64(hkeys, m, 2))
    assert err < 0.25
    want = (0 - O.nc_mul_exact64(hkeys.K2, m)).astype(np.uint64)
    e = _centered64(O.rlwe_phase64(hkeys.K2, ss.reshape(-1)) - want)
    assert np.abs(e).max() < 2 ** 40


@pytest.mark.parametrize("bit", [0, 1])
def test_cbs_htss_rows_and_cmux(hkeys, bit):
    """The circuit bootstrap through HomTrace + SchemeSwitch: row 2j + 1 (HT) decrypts to
    bit 2^(16 + 8j) (constant), row 2j (SS) to -K2 bit 2^(16 + 8j), on the 32-bit torus,
    within 2^14 at coefficient 0 and 2^10 elsewhere (HT), 2^15 (SS); the RGSW then selects
    in a CMux: the SS rows' noise (about 2^-21 per coefficient, the HT noise times K2)
    meets digits of up to 2^7 over 2 rows x N coefficients, about 2^23.2 of 2^32 rms per
    CMux; asserted below 2^26 (the decision margin is 2^30)."""
    p = hkeys.p
    N = p.N
    ct = hkeys.encrypt_bit(bit, 40 + bit)
    rgsw, err = hkeys.cbs_htss(ct)
    assert err < 0.25
    K = hkeys.K2.astype(np.int64)
    for j in range(p.cb_l):
        g = bit << (16 + 8 * j)
        ht = O.rlwe_phase32(hkeys.K2, rgsw[(2 * j + 1) * 2 * N:(2 * j + 2) * 2 * N]).view(np.int32).astype(np.int64)
        want = np.zeros(N, np.int64)
        want[0] = g
        e = ht - want
        assert abs(e[0]) < 2 ** 14 and np.abs(e[1:]).max() < 2 ** 10
        ss = O.rlwe_phase32(hkeys.K2, rgsw[2 * j * 2 * N:(2 * j + 1) * 2 * N]).view(np.int32).astype(np.int64)
        e = ((ss + K * g + 2 ** 31) % 2 ** 32) - 2 ** 31
        assert np.abs(e).max() < 2 ** 15
    C = hkeys.rgsw_fourier(rgsw)
    rng = np.random.default_rng(60 + bit)
    msgs = [(rng.integers(0, 2, N).astype(np.uint32) << np.uint32(31)).astype(np.uint32) for _ in range(2)]
    ds = []
    for msg in msgs:
        a = rng.integers(0, 2 ** 32, N, dtype=np.uint64).astype(np.uint32)
        ds.append(np.concatenate([a, (O.nc_mul_exact32(hkeys.K2, a) + msg).astype(np.uint32)]))
    out = np.zeros(2 * N, np.uint32)
    assert O._cbs_lib().orc_cmux32(N, p.cb_logB, p.cb_l, C, ds[0], ds[1], out) < 0.25
    e = ((O.rlwe_phase32(hkeys.K2, out) - msgs[bit]).astype(np.uint32)).view(np.int32)
    assert np.abs(e).max() < 2 ** 26
