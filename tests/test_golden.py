"""Frozen golden fixtures (tests/golden/*.npz, written by
tests/golden/make_golden.py from the oracle): the oracle must keep reproducing
them on the CPU, and the C ABI must reproduce them on the GPU with keys
generated from the same seeds by the library's own key generators."""
import os

import numpy as np
import pytest

import oracle as O
from conftest import from_dev, to_dev

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def lut_poly(table):
    poly = np.zeros(2048, np.uint32)
    for x in range(256):
        for j in range(8):
            poly[8 * x + j] = np.uint32(((int(table[x]) >> j) & 1) << 31)
    return poly


# ---------------------------------------------------------------- CPU (oracle)
def test_oracle_gate_fixture():
    g = load("gate_seed3.npz")
    k = O.GateKeys(int(g["seed"]))
    for q in range(len(g["gates"])):
        out, err = k.bootstrap(int(g["gates"][q]), g["ct1"][q], g["ct2"][q])
        assert err < 0.25
        assert np.array_equal(out, g["out"][q]), q
        assert k.decrypt(out) == O.GATE_FUNCS[int(g["gates"][q])](int(g["a"][q]), int(g["b"][q]))
    lweN, _ = O.gate_blind_rotate_extract(k, int(g["gates"][0]), g["ct1"][0], g["ct2"][0])
    assert np.array_equal(lweN, g["lweN0"])


def test_oracle_cbs_fixture():
    g = load("cbs_seed21.npz")
    k = O.CbsKeys(int(g["seed"]))
    for q in range(2):
        out, _ = k.pbs(g["ct"][q])
        assert np.array_equal(out, g["pbs"][q]), q
        for u in (0, 1):
            assert np.array_equal(k.pfks(u, g["lwe64"][q]), g["pfks"][q, u]), (q, u)
        assert np.array_equal(k.ks2(g["lwe64"][q]), g["ks2"][q]), q
    rgsw0, _ = k.cbs(g["ct"][0])
    assert np.array_equal(rgsw0, g["rgsw0"].reshape(-1))
    x, id0 = int(g["lut_x"]), int(g["lut_rgsw_id0"])
    fs = [k.rgsw_fourier(k.rgsw_encrypt((x >> i) & 1, id0 + 2 * i)) for i in range(8)]
    acc, _ = k.vertical_packing(fs, lut_poly(g["table"]))
    for j in range(8):
        assert np.array_equal(k.extract32(acc, j), g["lut"][j]), j
        ph = (int(g["lut"][j][-1]) - int(np.dot(g["lut"][j][:-1].astype(np.uint64), k.K2.astype(np.uint64)))) % 2 ** 32
        assert ((ph + 2 ** 30) >> 31) & 1 == (int(g["table"][x]) >> j) & 1


def test_oracle_cbs_linear_fixture():
    """Linear-key PFKS fixture: the oracle reproduces the switched RLWEs, the circuit
    bootstrap and the key entries; each switched RLWE decrypts to f_u(phase)."""
    g = load("cbs_linear_seed21.npz")
    k = O.CbsKeys(int(g["seed"]))
    for q in range(2):
        for u in (0, 1):
            assert np.array_equal(k.pfks_linear(u, g["lwe64"][q]), g["pfkl"][q, u]), (q, u)
    for (u, i, t), words in zip(g["entries"], g["entry_words"]):
        assert np.array_equal(k.pfkl_entry(int(u), int(i), int(t)), words), (u, i, t)
    rgsw0, _ = k.cbs(g["ct"][0], linear=True)
    assert np.array_equal(rgsw0, g["rgsw0"].reshape(-1))
    N = k.p.N
    for q in range(2):
        ph = k.lwe64_phase(g["lwe64"][q]) >> 32
        m = np.zeros(N, np.uint32)
        m[0] = ph
        err = ((O.rlwe_phase32(k.K2, g["pfkl"][q, 1]) - m).astype(np.uint32)).view(np.int32)
        assert np.abs(err).max() < 2 ** 13, q


def test_aes_plain_fixture_oracle_and_library():
    import tfhe_amd
    g = load("aes128_plain.npz")
    for q in range(len(g["key"])):
        key, blk = bytes(g["key"][q]), bytes(g["block"][q])
        assert O.aes128_encrypt(key, blk) == bytes(g["out"][q])
        assert tfhe_amd.aes128_encrypt_plain(key, blk) == bytes(g["out"][q])


# ---------------------------------------------------------------- GPU (C ABI)
@pytest.mark.gpu
def test_gpu_gate_fixture(gpu):
    """tfhe_gate_bt_keygen from the fixture's seed, then every fixture gate."""
    import torch
    import tfhe_amd
    g = load("gate_seed3.npz")
    seed = int(g["seed"])
    s = np.zeros(636, np.int32)
    O.lib().orc_gen_lwe_sk(seed, 636, s)
    c = tfhe_amd.GateContext(max_batch=64)
    try:
        c.bt_keygen(s, seed)
        for gate in np.unique(g["gates"]):
            sel = np.nonzero(g["gates"] == gate)[0]
            d_out = torch.empty((len(sel), 637), dtype=torch.int32, device="cuda")
            c.eval_batch(int(gate), to_dev(g["ct1"][sel]), to_dev(g["ct2"][sel]), d_out)
            torch.cuda.synchronize()
            assert np.array_equal(from_dev(d_out), g["out"][sel]), int(gate)
    finally:
        c.close()


@pytest.mark.gpu
def test_gpu_cbs_fixture(gpu):
    """tfhe_cbs_keygen from the fixture's seed; PBS, whole CBS, PFKS, KS2 and the
    LUT reproduce the fixture bit for bit."""
    import torch
    import tfhe_amd
    g = load("cbs_seed21.npz")
    seed = int(g["seed"])
    okeys = O.CbsKeys(seed, with_bsk=False)
    c = tfhe_amd.CbsContext(max_blocks=1)
    try:
        c.keygen(okeys.s, seed)
        d_ct = to_dev(g["ct"])
        d_pbs = torch.empty((2, 2, 2049), dtype=torch.int64, device="cuda")
        c.pbs_batch(d_ct, d_pbs)
        d_rgsw = torch.empty((2, 4, 2, 2048), dtype=torch.int32, device="cuda")
        c.cbs_batch(d_ct, d_rgsw)
        d_l64 = torch.from_numpy(g["lwe64"].view(np.int64).copy()).to("cuda")
        d_pf = torch.empty((2, 2, 4096), dtype=torch.int32, device="cuda")
        for u in (0, 1):
            c.pfks_batch(u, d_l64, 2, 2 * 4096, d_pf[:, u])
        l32 = ((g["lwe64"] + np.uint64(1 << 31)) >> np.uint64(32)).astype(np.uint32)
        d_ks = torch.empty((2, 637), dtype=torch.int32, device="cuda")
        c.keyswitch_batch(to_dev(l32), 2, d_ks)
        x, id0 = int(g["lut_x"]), int(g["lut_rgsw_id0"])
        rg = np.concatenate([okeys.rgsw_encrypt((x >> i) & 1, id0 + 2 * i) for i in range(8)])
        d_f = torch.empty(8 * 8 * 1024 * 2, dtype=torch.float64, device="cuda")
        c.rgsw_to_fourier(to_dev(rg), 8, d_f)
        d_lut = torch.empty((1, 8, 2049), dtype=torch.int32, device="cuda")
        c.lut8(d_f, 1, to_dev(lut_poly(g["table"])), 0, 8, d_lut)
        torch.cuda.synchronize()
        assert np.array_equal(from_dev(d_pbs, np.uint64), g["pbs"])
        assert np.array_equal(from_dev(d_rgsw)[0], g["rgsw0"])
        assert np.array_equal(from_dev(d_pf), g["pfks"])
        assert np.array_equal(from_dev(d_ks), g["ks2"])
        assert np.array_equal(from_dev(d_lut)[0], g["lut"])
    finally:
        c.close()


@pytest.mark.gpu
def test_gpu_cbs_linear_fixture(gpu):
    """tfhe_cbs_keygen with pfks_linear = 1 from the fixture's seed: the key entries,
    both switched RLWEs of each input and a whole circuit bootstrap reproduce the
    fixture bit for bit."""
    import torch
    import tfhe_amd
    g = load("cbs_linear_seed21.npz")
    seed = int(g["seed"])
    okeys = O.CbsKeys(seed, with_bsk=False)
    c = tfhe_amd.CbsContext(max_blocks=1, params=tfhe_amd.cbs_params(pfks_linear=1))
    try:
        c.keygen(okeys.s, seed)
        for (u, i, t), words in zip(g["entries"], g["entry_words"]):
            assert np.array_equal(c.pfkl_entry(int(u), int(i), int(t)), words), (u, i, t)
        d_l64 = torch.from_numpy(g["lwe64"].view(np.int64).copy()).to("cuda")
        d_pf = torch.empty((2, 2, 4096), dtype=torch.int32, device="cuda")
        for u in (0, 1):
            c.pfks_batch(u, d_l64, 2, 2 * 4096, d_pf[:, u])
        d_rgsw = torch.empty((1, 4, 2, 2048), dtype=torch.int32, device="cuda")
        c.cbs_batch(to_dev(g["ct"][:1]), d_rgsw)
        torch.cuda.synchronize()
        assert np.array_equal(from_dev(d_pf), g["pfkl"])
        assert np.array_equal(from_dev(d_rgsw)[0], g["rgsw0"])
    finally:
        c.close()
