"""Writes the golden fixtures of this directory from the CPU oracle.

Each fixture holds seeded inputs and the expected outputs of one piece of the
bootstrapping path; keys are not stored (they are regenerated from the seed by
the oracle on the CPU and by the library's key generators on the GPU, which
are bit-identical).  Run from the repository root:

    python tests/golden/make_golden.py            (all fixtures)
    python tests/golden/make_golden.py cbs_linear (one of them)

The fixtures are then frozen: tests/test_golden.py checks that the oracle still
reproduces them (CPU) and that the C ABI reproduces them (GPU).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402

GATE_SEED, CBS_SEED = 3, 21
GATES = [O.NAND, O.XOR, O.AND, O.XNOR]
LUT_X = 0x53
LUT_RGSW_ID0 = 4000


def lut_table():
    """A seeded random byte table packed for vertical packing (coefficient
    8x + j = bit j of table[x], times 2^31)."""
    table = np.random.default_rng(1234).integers(0, 256, 256)
    poly = np.zeros(2048, np.uint32)
    for x in range(256):
        for j in range(8):
            poly[8 * x + j] = np.uint32(((int(table[x]) >> j) & 1) << 31)
    return table.astype(np.uint8), poly


def lut_rgsws(okeys):
    return [okeys.rgsw_encrypt((LUT_X >> i) & 1, LUT_RGSW_ID0 + 2 * i) for i in range(8)]


def make_gate():
    k = O.GateKeys(GATE_SEED)
    a = np.array([0, 0, 1, 1] * len(GATES))
    b = np.array([0, 1, 0, 1] * len(GATES))
    gates = np.repeat(np.array(GATES, np.int32), 4)
    c1 = np.stack([k.encrypt(int(a[q]), 2 * q, seed=GATE_SEED + 100) for q in range(len(a))])
    c2 = np.stack([k.encrypt(int(b[q]), 2 * q + 1, seed=GATE_SEED + 100) for q in range(len(a))])
    out = np.stack([k.bootstrap(int(gates[q]), c1[q], c2[q])[0] for q in range(len(a))])
    lweN, _ = O.gate_blind_rotate_extract(k, int(gates[0]), c1[0], c2[0])
    np.savez(os.path.join(HERE, "gate_seed3.npz"), seed=np.uint64(GATE_SEED), gates=gates, a=a.astype(np.uint8),
             b=b.astype(np.uint8), ct1=c1, ct2=c2, out=out, lweN0=lweN)


def make_cbs():
    k = O.CbsKeys(CBS_SEED)
    bits = np.array([1, 0], np.uint8)
    cts = np.stack([k.encrypt_bit(int(bits[q]), 900 + q) for q in range(2)])
    pbs = np.stack([k.pbs(cts[q])[0] for q in range(2)])
    rgsw0, _ = k.cbs(cts[0])
    rng = np.random.default_rng(2024)
    lwe64 = rng.integers(0, 2 ** 64, (2, 2049), dtype=np.uint64)
    pfks = np.stack([np.stack([k.pfks(u, lwe64[q]) for u in (0, 1)]) for q in range(2)])
    ks2 = np.stack([k.ks2(lwe64[q]) for q in range(2)])
    table, poly = lut_table()
    fs = [k.rgsw_fourier(r) for r in lut_rgsws(k)]
    acc, _ = k.vertical_packing(fs, poly)
    lut = np.stack([k.extract32(acc, j) for j in range(8)])
    np.savez(os.path.join(HERE, "cbs_seed21.npz"), seed=np.uint64(CBS_SEED), bits=bits, ct=cts, pbs=pbs,
             rgsw0=rgsw0.reshape(4, 2, 2048), lwe64=lwe64, pfks=pfks, ks2=ks2, table=table,
             lut_x=np.uint8(LUT_X), lut_rgsw_id0=np.uint32(LUT_RGSW_ID0), lut=lut)


def make_cbs_linear():
    """The linear-key PFKS (40-bit key, signed digits) on the inputs of make_cbs:
    the switched RLWEs of both functions, a whole circuit bootstrap, key entries."""
    k = O.CbsKeys(CBS_SEED)
    bits = np.array([1, 0], np.uint8)
    cts = np.stack([k.encrypt_bit(int(bits[q]), 900 + q) for q in range(2)])
    rng = np.random.default_rng(2024)
    lwe64 = rng.integers(0, 2 ** 64, (2, 2049), dtype=np.uint64)
    pfkl = np.stack([np.stack([k.pfks_linear(u, lwe64[q]) for u in (0, 1)]) for q in range(2)])
    rgsw0, _ = k.cbs(cts[0], linear=True)
    entries = np.array([[0, 0, 0], [1, 2048, 7], [0, 1000, 4]], np.int32)
    ent = np.stack([k.pfkl_entry(int(u), int(i), int(t)) for u, i, t in entries])
    np.savez(os.path.join(HERE, "cbs_linear_seed21.npz"), seed=np.uint64(CBS_SEED), ct=cts, lwe64=lwe64, pfkl=pfkl,
             rgsw0=rgsw0.reshape(4, 2, 2048), entries=entries, entry_words=ent)


def make_aes():
    rng = np.random.default_rng(77)
    keys = rng.integers(0, 256, (3, 16), dtype=np.uint8)
    blocks = rng.integers(0, 256, (3, 16), dtype=np.uint8)
    out = np.stack([np.frombuffer(O.aes128_encrypt(bytes(keys[q]), bytes(blocks[q])), np.uint8) for q in range(3)])
    np.savez(os.path.join(HERE, "aes128_plain.npz"), key=keys, block=blocks, out=out)


if __name__ == "__main__":
    makers = {"gate": make_gate, "cbs": make_cbs, "cbs_linear": make_cbs_linear, "aes": make_aes}
    for name in (sys.argv[1:] or list(makers)):
        makers[name]()
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))
