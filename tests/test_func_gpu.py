"""GPU parity of the gate context's programmable bootstrap and large-precision functions
(reference bitwise/src/binfhe-base-scheme.cu:255-698, binfhe-base-scheme-batch.cu:215-698)
against the CPU oracle, through the C ABI of libtfhe_amd.so: every output word equal to the
oracle's on the same inputs, plus decryption checks on a full batch."""
import numpy as np
import pytest

import oracle as O
from conftest import from_dev, to_dev

pytestmark = pytest.mark.gpu

Q, BETA = O.FUNC_Q, O.FUNC_BETA
P = Q // (2 * BETA)


@pytest.fixture(scope="module")
def gk():
    return O.GateKeys(31)


@pytest.fixture(scope="module")
def fctx(gpu, gk):
    import tfhe_amd
    c = tfhe_amd.GateContext(max_batch=1024)
    c.set_keys(gk.bsk, gk.ksk)
    yield c
    c.close()


def _out(batch, digits=None):
    import torch
    shape = (batch, O.gate_params().n + 1) if digits is None else (digits, batch, O.gate_params().n + 1)
    return torch.empty(shape, dtype=torch.int32, device="cuda")


def _get(t):
    import torch
    torch.cuda.synchronize()
    return from_dev(t)


def _cts(gk, values, mod, id0):
    return np.stack([gk.encrypt_mod(int(v), mod, id0 + k) for k, v in enumerate(values)])


def _lut(F):
    return np.array([(F[(x * P) // Q] % P) * (Q // P) for x in range(Q)], np.uint32)


def test_func_property_matches_oracle(gpu):
    import tfhe_amd
    q = 16
    neg = [(3 * x + 1) % q for x in range(q // 2)]
    neg = neg + [(-v) % q for v in neg]
    for t in (neg, [(5 * x) % q for x in range(q // 2)] * 2, [(3 * x + 1) % q for x in range(q)], [0] * q):
        assert tfhe_amd.func_property(t, q) == O.func_property(t, q)


@pytest.mark.parametrize("ctmod,fmod", [(1024, 1024), (2048, 2 ** 16), (256, 2 ** 32), (2, 4)])
def test_bootstrap_func_bit_exact(fctx, gk, ctmod, fmod):
    rng = np.random.default_rng(ctmod * 7 + fmod % 997)
    f = rng.integers(0, fmod, ctmod, dtype=np.uint64).astype(np.uint32)
    batch = 5
    cts = _cts(gk, rng.integers(0, ctmod, batch), ctmod, 40)
    d_out = _out(batch)
    fctx.bootstrap_func(to_dev(cts), ctmod, f, fmod, d_out)
    got = _get(d_out)
    for k in range(batch):
        assert np.array_equal(got[k], gk.bootstrap_func(cts[k], ctmod, f, fmod)), k


@pytest.mark.parametrize("kind", ["negacyclic", "periodic", "arbitrary"])
def test_eval_func_bit_exact(fctx, gk, kind):
    import tfhe_amd
    F = {"negacyclic": [1, 2, 3, 0, 7, 6, 5, 0], "periodic": [3, 1, 4, 1, 3, 1, 4, 1],
         "arbitrary": [(3 * m + 1) % 8 for m in range(8)]}[kind]
    lut = _lut(F)
    assert tfhe_amd.func_property(lut, Q) == {"negacyclic": 0, "periodic": 1, "arbitrary": 2}[kind]
    msgs = np.arange(P)
    cts = _cts(gk, msgs * (Q // P), Q, 60)
    d_out = _out(P)
    fctx.eval_func(to_dev(cts), Q, lut, d_out)
    got = _get(d_out)
    for k in range(P):
        assert np.array_equal(got[k], gk.eval_func(cts[k], Q, lut)), (kind, k)
        ph = gk.phase_mod(got[k], Q)
        assert ((ph + Q // (2 * P)) // (Q // P)) % P == F[k], (kind, k)


@pytest.mark.parametrize("mod,roundbits", [(2 ** 16, 0), (2 ** 16, 1), (2 ** 32, 0)])
def test_eval_floor_bit_exact(fctx, gk, mod, roundbits):
    rng = np.random.default_rng(mod % 1000 + roundbits)
    batch = 4
    cts = _cts(gk, rng.integers(0, mod, batch), mod, 80)
    d_out = _out(batch)
    fctx.eval_floor(to_dev(cts), mod, d_out, roundbits=roundbits)
    got = _get(d_out)
    for k in range(batch):
        assert np.array_equal(got[k], gk.eval_floor(cts[k], mod, roundbits)), k


@pytest.mark.parametrize("mod,scheme_switch", [(2 ** 16, False), (2 ** 22, True), (2 ** 32, False)])
def test_eval_sign_bit_exact(fctx, gk, mod, scheme_switch):
    rng = np.random.default_rng(mod % 991)
    batch = 3
    vals = [int(rng.integers(mod // 16, mod // 2 - mod // 16)), int(rng.integers(mod // 2 + mod // 16, mod - mod // 16)),
            int(rng.integers(0, mod))]
    cts = _cts(gk, vals, mod, 100)
    d_out = _out(batch)
    fctx.eval_sign(to_dev(cts), mod, d_out, scheme_switch=scheme_switch)
    got = _get(d_out)
    for k in range(batch):
        assert np.array_equal(got[k], gk.eval_sign(cts[k], mod, scheme_switch)), k
    for k in range(2):  # the two inputs far from the boundaries decrypt to their top bit
        ph = gk.phase_mod(got[k], Q)
        if scheme_switch:
            want = Q - Q // 4 if vals[k] < mod // 2 else Q // 4
        else:
            want = 0 if vals[k] < mod // 2 else Q // 2
        assert abs((ph - want + Q // 2) % Q - Q // 2) < Q // 8, (k, ph, want)


@pytest.mark.parametrize("mod", [2 ** 13, 2 ** 32])
def test_eval_decomp_bit_exact(fctx, gk, mod):
    rng = np.random.default_rng(mod % 977)
    batch = 3
    cts = _cts(gk, rng.integers(0, mod, batch), mod, 120)
    digits_want = [gk.eval_decomp(cts[k], mod) for k in range(batch)]
    nd = len(digits_want[0])
    d_out = _out(batch, digits=9)
    mods = fctx.eval_decomp(to_dev(cts), mod, d_out, 9)
    assert len(mods) == nd and all(m == Q for m in mods[:-1]) and mods[-1] <= Q
    got = _get(d_out)
    for k in range(batch):
        for d in range(nd):
            assert np.array_equal(got[d, k], digits_want[k][d]), (k, d)


def test_eval_func_full_batch_decrypts(fctx, gk):
    """1024 EvalFunc ciphertexts over Z_1024 (messages of Z_8, the arbitrary table: two
    bootstraps through the modulus 2048): every output decrypts to F(m) (the bootstrap noise,
    about 2 units at this modulus, is far inside beta = 64); sampled rows equal the oracle."""
    batch = 1024
    F = [(3 * m + 1) % 8 for m in range(8)]
    lut = _lut(F)
    rng = np.random.default_rng(5)
    msgs = rng.integers(0, P, batch)
    cts = _cts(gk, msgs * (Q // P), Q, 2000)
    d_out = _out(batch)
    fctx.eval_func(to_dev(cts), Q, lut, d_out)
    got = _get(d_out)
    s = gk.s.astype(np.uint64)
    ph = ((got[:, -1].astype(np.uint64) - got[:, :-1].astype(np.uint64) @ s) % Q).astype(np.int64)
    assert np.array_equal(((ph + Q // (2 * P)) // (Q // P)) % P, np.array(F)[msgs])
    for k in (0, 511, 1023):
        assert np.array_equal(got[k], gk.eval_func(cts[k], Q, lut)), k


def test_sign_full_batch(fctx, gk):
    """1024 signs over Z_2^16 (two floors and the sign bootstrap, five bootstraps each), inputs
    2^12 away from 0 and 2^15.  Sampled rows equal the oracle word for word.  The gate keys'
    bootstrap noise (2^-8.1 of the output modulus) makes a floor at 2^16 and 2^13 move by a
    level now and then, so a few percent of signs come out wrong (the oracle errs on 2 of 150);
    the bound here is 5%."""
    mod, batch = 2 ** 16, 1024
    rng = np.random.default_rng(5)
    top = rng.integers(0, 2, batch)
    vals = top * (mod // 2) + rng.integers(2 ** 12, mod // 2 - 2 ** 12, batch)
    cts = _cts(gk, vals, mod, 4000)
    d_out = _out(batch)
    fctx.eval_sign(to_dev(cts), mod, d_out)
    got = _get(d_out)
    s = gk.s.astype(np.uint64)
    ph = ((got[:, -1].astype(np.uint64) - got[:, :-1].astype(np.uint64) @ s) % Q).astype(np.int64)
    dec = ((ph + Q // 4) // (Q // 2)) % 2
    assert np.mean(dec != top) < 0.05
    for k in (0, 511, 1023):
        assert np.array_equal(got[k], gk.eval_sign(cts[k], mod)), k


def test_func_argument_errors(fctx, gk):
    import tfhe_amd
    cts = _cts(gk, [1], 2 ** 16, 1)
    d_in, d_out = to_dev(cts), _out(1)
    with pytest.raises(tfhe_amd.TfheError) as e:  # ctmod above 2N
        fctx.bootstrap_func(d_in, 4096, np.zeros(4096, np.uint32), 4096, d_out)
    assert e.value.code == tfhe_amd.TFHE_E_PARAM
    with pytest.raises(tfhe_amd.TfheError) as e:  # arbitrary table above N
        fctx.eval_func(d_in, 2048, np.arange(2048, dtype=np.uint32), d_out)
    assert e.value.code == tfhe_amd.TFHE_E_PARAM
    with pytest.raises(tfhe_amd.TfheError) as e:  # sign needs mod > q_small
        fctx.eval_sign(d_in, Q, d_out)
    assert e.value.code == tfhe_amd.TFHE_E_PARAM
    with pytest.raises(tfhe_amd.TfheError) as e:  # 2^32 takes 9 digits
        fctx.eval_decomp(d_in, 2 ** 32, _out(1, digits=8), 8)
    assert e.value.code == tfhe_amd.TFHE_E_CAPACITY
    with pytest.raises(tfhe_amd.TfheError) as e:
        fctx.eval_floor(d_in, 2 ** 16, d_out, batch=fctx.max_batch + 1)
    assert e.value.code in (tfhe_amd.TFHE_E_CAPACITY, tfhe_amd.TFHE_E_SHAPE)


def test_func_empty_batches_leave_outputs_untouched(fctx, gk):
    """A batch of 0 returns OK and launches nothing in every flow (outputs keep a sentinel);
    EvalDecomp still reports its digit moduli (2^32: 8 digits of q_small, then 2^8)."""
    import torch
    SENT = 0x5A5A5A5A
    d_in = to_dev(_cts(gk, [1], 2 ** 16, 1))
    out = torch.full((1, O.gate_params().n + 1), SENT, dtype=torch.int32, device="cuda")
    dig = torch.full((9, 1, O.gate_params().n + 1), SENT, dtype=torch.int32, device="cuda")
    lut = _lut([(3 * m + 1) % 8 for m in range(8)])
    fctx.bootstrap_func(d_in, Q, lut, Q, out, batch=0)
    fctx.eval_func(d_in, Q, lut, out, batch=0)
    fctx.eval_floor(d_in, 2 ** 16, out, batch=0)
    fctx.eval_sign(d_in, 2 ** 16, out, batch=0)
    mods = fctx.eval_decomp(d_in, 2 ** 32, dig, 9, batch=0)
    torch.cuda.synchronize()
    assert mods == [Q] * 8 + [2 ** 8]
    assert (from_dev(out) == SENT).all() and (from_dev(dig) == SENT).all()
