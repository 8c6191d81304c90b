"""tfhe_gate_eval_batch_noalloc (the batched gate bootstrap with the library's fixed device
scratch) against tfhe_gate_eval_batch, through the C ABI: the same words on two calls in a row,
the second smaller, the capacity limit, and a context whose own scratch is smaller than the batch."""
import numpy as np
import pytest

import oracle as O
from conftest import from_dev, to_dev

pytestmark = pytest.mark.gpu

TRUTH = {3: lambda a, b: 1 - (a & b), 11: lambda a, b: a ^ b, 1: lambda a, b: a & b}


@pytest.fixture(scope="module")
def okeys():
    return O.GateKeys(5)


def _pairs(okeys, batch, seed):
    rng = np.random.default_rng(seed)
    a = rng.integers(0, 2, batch)
    b = rng.integers(0, 2, batch)
    c1 = np.stack([okeys.encrypt(int(a[k]), 2 * k, seed=seed) for k in range(batch)])
    c2 = np.stack([okeys.encrypt(int(b[k]), 2 * k + 1, seed=seed) for k in range(batch)])
    return a, b, c1, c2


def test_noalloc_equals_eval_batch(gpu, okeys):
    import torch
    import tfhe_amd
    big = tfhe_amd.GateContext(max_batch=tfhe_amd.NOALLOC_MAX_GATES)
    small = tfhe_amd.GateContext(max_batch=1)  # its own scratch holds one gate
    try:
        big.set_keys(okeys.bsk, okeys.ksk)
        small.set_keys(okeys.bsk, okeys.ksk)
        # the full scratch first, then a smaller ragged batch that reuses it
        for batch, gate, seed in ((tfhe_amd.NOALLOC_MAX_GATES, tfhe_amd.NAND, 41), (300, tfhe_amd.XOR, 42)):
            print(f"noalloc case: batch {batch}, gate {gate}, seed {seed}", flush=True)
            a, b, c1, c2 = _pairs(okeys, batch, seed)
            d1, d2 = to_dev(c1), to_dev(c2)
            ref = torch.empty_like(d1)
            out = torch.full_like(d1, 0x5A5A5A5A)
            big.eval_batch(gate, d1, d2, ref)
            ctx = big if batch == tfhe_amd.NOALLOC_MAX_GATES else small
            ctx.eval_batch_noalloc(gate, d1, d2, out)
            torch.cuda.synchronize()
            got, want = from_dev(out), from_dev(ref)
            assert np.array_equal(got, want), f"batch {batch}: {np.count_nonzero(got != want)} words differ"
            bits = np.array([okeys.decrypt(got[k]) for k in range(batch)])
            assert np.array_equal(bits, TRUTH[gate](a, b))
        # the same context again: the array keeps the last call's contents, the output must not depend on them
        a, b, c1, c2 = _pairs(okeys, 64, 43)
        d1, d2 = to_dev(c1), to_dev(c2)
        ref = torch.empty_like(d1)
        out = torch.empty_like(d1)
        big.eval_batch(tfhe_amd.AND, d1, d2, ref)
        big.eval_batch_noalloc(tfhe_amd.AND, d1, d2, out)
        torch.cuda.synchronize()
        assert np.array_equal(from_dev(out), from_dev(ref))
    finally:
        big.close()
        small.close()


def test_noalloc_rejects_batches_beyond_its_scratch(gpu, okeys):
    import torch
    import tfhe_amd
    c = tfhe_amd.GateContext(max_batch=8)
    try:
        c.set_keys(okeys.bsk, okeys.ksk)
        n1 = c.p.n + 1
        over = tfhe_amd.NOALLOC_MAX_GATES + 1
        d1 = torch.zeros((over, n1), dtype=torch.int32, device="cuda")
        d2 = torch.ones((over, n1), dtype=torch.int32, device="cuda")
        out = torch.empty_like(d1)
        with pytest.raises(tfhe_amd.TfheError) as e:
            c.eval_batch_noalloc(tfhe_amd.NAND, d1, d2, out)
        assert e.value.code == tfhe_amd.TFHE_E_CAPACITY
        with pytest.raises(tfhe_amd.TfheError) as e:  # the same input twice is rejected, as in eval_batch
            c.eval_batch_noalloc(tfhe_amd.NAND, d1, d1, out, batch=4)
        assert e.value.code == tfhe_amd.TFHE_E_PARAM
    finally:
        c.close()
