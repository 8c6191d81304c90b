"""The C-ABI library loads on a machine without a GPU and exports every entry
point declared in include/*.h (no compute calls)."""
import ctypes
import glob
import os
import re

import pytest

import tfhe_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    syms = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"\b(tfhe_[a-z0-9_]+)\s*\(", txt):
            syms.add(m.group(1))
    return sorted(syms)


def test_headers_declare_entry_points():
    assert len(declared_symbols()) >= 10


@pytest.mark.parametrize("sym", declared_symbols())
def test_library_exports(sym):
    if not os.path.exists(tfhe_amd.LIB_PATH):
        pytest.skip("library not built")
    lib = ctypes.CDLL(tfhe_amd.LIB_PATH)
    assert hasattr(lib, sym), sym


def test_binding_signatures_cover_header():
    missing = [s for s in declared_symbols() if s not in tfhe_amd.SIGNATURES]
    assert not missing, missing


def test_host_encrypt_matches_oracle():
    """The library's host Encrypt/Decrypt stand-ins agree with the oracle's sampler."""
    import numpy as np
    import oracle as O
    if not os.path.exists(tfhe_amd.LIB_PATH):
        pytest.skip("library not built")
    sk = tfhe_amd.lwe_keygen(636, 9)
    s2 = np.zeros(636, np.int32)
    O.lib().orc_gen_lwe_sk(9, 636, s2)
    assert np.array_equal(sk, s2)
    for m in range(4):
        ct = tfhe_amd.lwe_encrypt(sk, 9, 100 + m, m)
        ct2 = np.zeros(637, np.uint32)
        O.lib().orc_lwe_encrypt32(9, O.DOM_ENC, 100 + m, 636, s2, m << 30, 15, ct2)
        assert np.array_equal(ct, ct2)
        assert tfhe_amd.lwe_decrypt(sk, ct) == m


def test_func_property_matches_oracle_on_cpu():
    """tfhe_func_property (host code, the test EvalFunc branches on) classifies tables as the
    oracle does: hand-made negacyclic / periodic / arbitrary tables, the all-zero table
    (both properties: negacyclic wins), random tables over several moduli, and bad moduli."""
    import numpy as np
    import oracle as O
    if not os.path.exists(tfhe_amd.LIB_PATH):
        pytest.skip("library not built")
    rng = np.random.default_rng(3)
    for q in (2, 16, 1024, 2048):
        half = rng.integers(0, q, q // 2)
        neg = np.concatenate([half, (-half) % q]).astype(np.uint32)
        per = np.concatenate([half, half]).astype(np.uint32)
        arb = rng.integers(0, q, q).astype(np.uint32)
        for t in (neg, per, arb, np.zeros(q, np.uint32)):
            assert tfhe_amd.func_property(t, q) == O.func_property(t, q), q
    assert tfhe_amd.func_property(np.zeros(12, np.uint32), 12) == -1  # not a power of two
