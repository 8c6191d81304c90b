"""GPU parity of the gate-bootstrapping path (N=1024, torus-32) against the CPU
oracle, through the C ABI of libtfhe_amd.so."""
import numpy as np
import pytest

import oracle as O
from conftest import from_dev, to_dev

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def okeys():
    return O.GateKeys(3)


@pytest.fixture(scope="module")
def ctx(gpu, okeys):
    import tfhe_amd
    c = tfhe_amd.GateContext(max_batch=4096)
    c.set_keys(okeys.bsk, okeys.ksk)
    yield c
    c.close()


def _pairs(okeys, batch, seed, bits=None):
    rng = np.random.default_rng(seed)
    a = rng.integers(0, 2, batch) if bits is None else bits[0]
    b = rng.integers(0, 2, batch) if bits is None else bits[1]
    c1 = np.stack([okeys.encrypt(int(a[k]), 2 * k, seed=seed) for k in range(batch)])
    c2 = np.stack([okeys.encrypt(int(b[k]), 2 * k + 1, seed=seed) for k in range(batch)])
    return a, b, c1, c2


def test_keygen_parity_with_oracle(gpu, okeys):
    """GPU key generators reproduce the oracle's BSK and KSK bit for bit."""
    import torch
    import tfhe_amd
    p = tfhe_amd.gate_params()
    d_s, d_K = to_dev(okeys.s), to_dev(okeys.K)
    d_bsk = torch.empty(okeys.bsk.size, dtype=torch.int32, device="cuda")
    d_ksk = torch.empty(okeys.ksk.size, dtype=torch.int32, device="cuda")
    tfhe_amd.check(tfhe_amd.lib().tfhe_gen_bsk32(p, 3, tfhe_amd._ptr(d_s), tfhe_amd._ptr(d_K), tfhe_amd._ptr(d_bsk), None))
    tfhe_amd.check(tfhe_amd.lib().tfhe_gen_ksk32(p, 3, tfhe_amd._ptr(d_s), tfhe_amd._ptr(d_K), tfhe_amd._ptr(d_ksk), None))
    torch.cuda.synchronize()
    assert np.array_equal(from_dev(d_bsk), okeys.bsk)
    assert np.array_equal(from_dev(d_ksk), okeys.ksk)


def test_polymul1024_exact(gpu):
    """Minimum slice: the wavefront FFT negacyclic product equals the schoolbook product."""
    import torch
    import tfhe_amd
    rng = np.random.default_rng(17)
    batch = 9
    x = rng.integers(-64, 64, (batch, 1024)).astype(np.int32)
    y = rng.integers(0, 2 ** 32, (batch, 1024), dtype=np.uint64).astype(np.uint32)
    d_out = torch.empty((batch, 1024), dtype=torch.int32, device="cuda")
    tfhe_amd.polymul1024_batch(to_dev(x), to_dev(y), d_out)
    torch.cuda.synchronize()
    got = from_dev(d_out)
    for k in range(batch):
        assert np.array_equal(got[k], O.nc_mul_exact32(x[k], y[k])), k


def test_external_product_bit_exact(ctx, okeys):
    import ctypes
    import torch
    p = okeys.p
    N, M, rows = p.N, p.N // 2, 2 * p.bk_l
    rng = np.random.default_rng(23)
    acc = rng.integers(0, 2 ** 32, 2 * N, dtype=np.uint64).astype(np.uint32)
    for i in (0, 5, p.n - 1):
        d_out = torch.empty(2 * N, dtype=torch.int32, device="cuda")
        ctx.external_product(i, to_dev(acc), d_out)
        torch.cuda.synchronize()
        key = np.ascontiguousarray(okeys.bsk_f.reshape(p.n, rows * 2 * 2 * M)[i])
        want = np.zeros(2 * N, np.uint32)
        err = O.lib().orc_external_product32(N, p.bk_logB, p.bk_l, key, acc, want)
        assert err < 0.25
        assert np.array_equal(from_dev(d_out), want), i


@pytest.mark.parametrize("gate", [O.NAND, O.XOR, O.AND, O.OR, O.NOR, O.XNOR])
def test_blind_rotate_extract_bit_exact(ctx, okeys, gate):
    """Blind rotation + sample extraction equals the oracle's LWE_N bit for bit."""
    import torch
    batch = 6
    a, b, c1, c2 = _pairs(okeys, batch, 100 + gate)
    d1, d2 = to_dev(c1), to_dev(c2)
    d_lweN = torch.empty((batch, okeys.p.N + 1), dtype=torch.int32, device="cuda")
    ctx.blind_rotate_batch(gate, d1, d2, d_lweN)
    torch.cuda.synchronize()
    got = from_dev(d_lweN)
    for k in range(batch):
        want, err = O.gate_blind_rotate_extract(okeys, gate, c1[k], c2[k])
        assert err < 0.25
        assert np.array_equal(got[k], want), (gate, k)


def test_keyswitch_bit_exact(ctx, okeys):
    import torch
    rng = np.random.default_rng(7)
    batch = 37  # ragged against the 16-ciphertext tile
    lweN = rng.integers(0, 2 ** 32, (batch, okeys.p.N + 1), dtype=np.uint64).astype(np.uint32)
    d_out = torch.empty((batch, okeys.p.n + 1), dtype=torch.int32, device="cuda")
    ctx.keyswitch_batch(to_dev(lweN), d_out)
    torch.cuda.synchronize()
    got = from_dev(d_out)
    for k in range(batch):
        assert np.array_equal(got[k], O.keyswitch32(okeys, lweN[k])), k


def test_keyswitch_uneven_slices_bit_exact(ctx, okeys):
    """1100 ciphertexts: 9 groups x 5 column tiles, so the coefficient range splits
    into 11 slices of 94 with a short last one (84); sampled outputs equal the oracle."""
    import torch
    rng = np.random.default_rng(71)
    batch = 1100
    lweN = rng.integers(0, 2 ** 32, (batch, okeys.p.N + 1), dtype=np.uint64).astype(np.uint32)
    d_out = torch.empty((batch, okeys.p.n + 1), dtype=torch.int32, device="cuda")
    ctx.keyswitch_batch(to_dev(lweN), d_out)
    torch.cuda.synchronize()
    got = from_dev(d_out)
    for k in (0, 127, 128, 600, 1099):
        assert np.array_equal(got[k], O.keyswitch32(okeys, lweN[k])), k


@pytest.mark.parametrize("gate", [O.AND, O.OR, O.NAND, O.NOR, O.XOR, O.XNOR, O.XOR_FAST, O.XNOR_FAST])
def test_gate_eval_batch_truth_table_and_parity(ctx, okeys, gate):
    """Whole gate (rotation + key switch) on all four input pairs: decrypts to the
    truth table and matches the oracle's output ciphertext exactly."""
    import torch
    bits = (np.array([0, 0, 1, 1]), np.array([0, 1, 0, 1]))
    a, b, c1, c2 = _pairs(okeys, 4, 200 + gate, bits)
    d_out = torch.empty((4, okeys.p.n + 1), dtype=torch.int32, device="cuda")
    ctx.eval_batch(gate, to_dev(c1), to_dev(c2), d_out)
    torch.cuda.synchronize()
    got = from_dev(d_out)
    for k in range(4):
        assert okeys.decrypt(got[k]) == O.GATE_FUNCS[gate](int(a[k]), int(b[k]))
        want, _ = okeys.bootstrap(gate, c1[k], c2[k])
        assert np.array_equal(got[k], want)


def test_gate_eval_full_batch_decrypts(ctx, okeys):
    """4096 NAND bootstraps (the batched workload) all decrypt correctly; the
    output phases stay within the +-Q/8 decision margin."""
    import torch
    batch = 4096
    rng = np.random.default_rng(99)
    a = rng.integers(0, 2, batch)
    b = rng.integers(0, 2, batch)
    import tfhe_amd
    c1 = np.stack([tfhe_amd.lwe_encrypt(okeys.s, 77, 2 * k, int(a[k])) for k in range(batch)])
    c2 = np.stack([tfhe_amd.lwe_encrypt(okeys.s, 77, 2 * k + 1, int(b[k])) for k in range(batch)])
    d_out = torch.empty((batch, okeys.p.n + 1), dtype=torch.int32, device="cuda")
    ctx.eval_batch(O.NAND, to_dev(c1), to_dev(c2), d_out)
    torch.cuda.synchronize()
    got = from_dev(d_out)
    s = okeys.s.astype(np.uint64)
    ph = (got[:, -1].astype(np.uint64) - (got[:, :-1].astype(np.uint64) @ s)) % 2 ** 32
    dec = ((ph + (1 << 29)) >> 30) & 3
    assert np.array_equal(dec, 1 - (a & b))
    err = ((ph - dec * (1 << 30) + 2 ** 31) % 2 ** 32).astype(np.int64) - 2 ** 31
    assert np.abs(err).max() < 2 ** 29  # margin Q/8


def test_plain_form_equals_batched(ctx, okeys):
    import torch
    a, b, c1, c2 = _pairs(okeys, 2, 300)
    d1, d2 = to_dev(c1), to_dev(c2)
    d_batch = torch.empty((2, okeys.p.n + 1), dtype=torch.int32, device="cuda")
    ctx.eval_batch(O.NAND, d1, d2, d_batch)
    torch.cuda.synchronize()
    for k in range(2):
        d_one = torch.empty((1, okeys.p.n + 1), dtype=torch.int32, device="cuda")
        ctx.eval_plain(O.NAND, d1[k:k + 1].contiguous(), d2[k:k + 1].contiguous(), d_one)
        torch.cuda.synchronize()
        assert np.array_equal(from_dev(d_one)[0], from_dev(d_batch)[k])


def test_gpu_keygen_context_gates(gpu):
    """Keys generated by tfhe_gate_bt_keygen (GPUBTKeyGen) bootstrap correctly."""
    import torch
    import tfhe_amd
    c = tfhe_amd.GateContext(max_batch=64)
    sk = tfhe_amd.lwe_keygen(636, 5)
    c.bt_keygen(sk, 5)
    K = c.rlwe_sk()
    K2 = np.zeros(1024, np.int32)
    O.lib().orc_gen_rlwe_sk(5, O.DOM_RLWE1_SK, 1024, K2)
    assert np.array_equal(K, K2)
    rng = np.random.default_rng(1)
    a = rng.integers(0, 2, 64)
    b = rng.integers(0, 2, 64)
    c1 = np.stack([tfhe_amd.lwe_encrypt(sk, 8, 2 * k, int(a[k])) for k in range(64)])
    c2 = np.stack([tfhe_amd.lwe_encrypt(sk, 8, 2 * k + 1, int(b[k])) for k in range(64)])
    d_out = torch.empty((64, 637), dtype=torch.int32, device="cuda")
    c.eval_batch(O.XOR, to_dev(c1), to_dev(c2), d_out)
    torch.cuda.synchronize()
    got = from_dev(d_out)
    assert [tfhe_amd.lwe_decrypt(sk, got[k]) for k in range(64)] == list(a ^ b)
    c.close()


def test_empty_batches_leave_outputs_untouched(ctx):
    """A count of 0 is a no-op for the batched gate entry points: OK, nothing launched,
    the sentinel in the outputs kept."""
    import torch
    import tfhe_amd
    L, P, C = tfhe_amd.lib(), tfhe_amd._ptr, tfhe_amd.check
    ct1 = torch.full((2, 637), 0x5A5A5A5A, dtype=torch.int32, device="cuda")
    ct2, out = ct1.clone(), ct1.clone()
    lweN = torch.full((2, 1025), 0x5A5A5A5A, dtype=torch.int32, device="cuda")
    before = (ct1.clone(), lweN.clone())
    C(L.tfhe_gate_eval_batch(ctx.h, tfhe_amd.NAND, P(ct1), P(ct2), 0, P(out), None))
    C(L.tfhe_gate_blind_rotate_batch(ctx.h, tfhe_amd.XOR, P(ct1), P(ct2), 0, P(lweN), None))
    C(L.tfhe_gate_keyswitch_batch(ctx.h, P(lweN), 0, P(out), None))
    torch.cuda.synchronize()
    assert torch.equal(out, before[0]) and torch.equal(lweN, before[1])


def test_error_paths(ctx, okeys):
    import torch
    import tfhe_amd
    d = torch.zeros((2, okeys.p.n + 1), dtype=torch.int32, device="cuda")
    out = torch.zeros_like(d)
    with pytest.raises(tfhe_amd.TfheError) as e:
        ctx.eval_batch(O.NAND, d, d, out)  # ct1 == ct2
    assert e.value.code == tfhe_amd.TFHE_E_PARAM
    with pytest.raises(tfhe_amd.TfheError) as e:
        ctx.eval_batch(6, d, d.clone(), out)  # MAJORITY is not a bootstrapped two-input gate
    assert e.value.code == tfhe_amd.TFHE_E_PARAM
    big = torch.zeros((4097, okeys.p.n + 1), dtype=torch.int32, device="cuda")
    with pytest.raises(tfhe_amd.TfheError) as e:
        ctx.eval_batch(O.NAND, big, big.clone(), big.clone())
    assert e.value.code == tfhe_amd.TFHE_E_CAPACITY
    with pytest.raises(tfhe_amd.TfheError) as e:  # key switch beyond the context's scratch capacity
        tfhe_amd.check(tfhe_amd.lib().tfhe_gate_keyswitch_batch(ctx.h, tfhe_amd._ptr(big), 4097,
                                                                tfhe_amd._ptr(big), None))
    assert e.value.code == tfhe_amd.TFHE_E_CAPACITY
    ctx.eval_batch(O.NAND, d, d.clone(), out, batch=0)  # empty batch is a no-op

