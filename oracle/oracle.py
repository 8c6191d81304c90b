"""ctypes bindings for the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product library.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")

DOM_LWE_SK, DOM_RLWE1_SK, DOM_RLWE2_SK = 1, 2, 3
DOM_BSK1, DOM_KSK1, DOM_BSK2, DOM_PFKS, DOM_KSK2, DOM_ENC = 4, 5, 6, 7, 8, 16

# OpenFHE BINGATE numbering
OR, AND, NOR, NAND, XOR_FAST, XNOR_FAST = 0, 1, 2, 3, 4, 5
XOR, XNOR = 11, 12
GATE_FUNCS = {
    OR: lambda a, b: a | b, AND: lambda a, b: a & b, NOR: lambda a, b: 1 - (a | b),
    NAND: lambda a, b: 1 - (a & b), XOR_FAST: lambda a, b: a ^ b, XNOR_FAST: lambda a, b: 1 - (a ^ b),
    XOR: lambda a, b: a ^ b, XNOR: lambda a, b: 1 - (a ^ b),
}


class GateParams(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("N", ctypes.c_int), ("bk_logB", ctypes.c_int),
                ("bk_l", ctypes.c_int), ("ks_logB", ctypes.c_int), ("ks_l", ctypes.c_int),
                ("lwe_nshift", ctypes.c_int), ("rlwe_nshift", ctypes.c_int)]


def gate_params(n=636, N=1024, bk_logB=7, bk_l=3, ks_logB=2, ks_l=8, lwe_nshift=15, rlwe_nshift=25):
    """The TFHE gate parameter set: n=636, N=1024, BSK 2^7 x 3, KS 2^2 x 8,
    LWE sigma 2^-15, RLWE sigma 2^-25 (torus-32)."""
    return GateParams(n, N, bk_logB, bk_l, ks_logB, ks_l, lwe_nshift, rlwe_nshift)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.POINTER(GateParams)
        sig = {
            "orc_philox4x32_10": (None, [_u32p, _u32p, _u32p]),
            "orc_rand32": (ctypes.c_uint32, [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]),
            "orc_rand64": (ctypes.c_uint64, [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]),
            "orc_noise": (ctypes.c_int64, [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int]),
            "orc_nc_mul_exact32": (None, [_i32p, _u32p, _u32p, ctypes.c_int]),
            "orc_nc_mul_exact64": (None, [_i32p, _u64p, _u64p, ctypes.c_int]),
            "orc_nc_mul_fft": (ctypes.c_double, [_i64p, _i64p, _i64p, ctypes.c_int]),
            "orc_decompose32": (None, [ctypes.c_uint32, ctypes.c_int, ctypes.c_int, _i32p]),
            "orc_decompose64": (None, [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, _i32p]),
            "orc_gen_lwe_sk": (None, [ctypes.c_uint64, ctypes.c_int, _i32p]),
            "orc_gen_rlwe_sk": (None, [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int, _i32p]),
            "orc_lwe_encrypt32": (None, [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, _i32p, ctypes.c_uint32, ctypes.c_int, _u32p]),
            "orc_lwe_phase32": (ctypes.c_uint32, [ctypes.c_int, _i32p, _u32p]),
            "orc_gen_bsk32": (None, [ctypes.c_uint64, P, _i32p, _i32p, _u32p]),
            "orc_gen_ksk32": (None, [ctypes.c_uint64, P, _i32p, _i32p, _u32p]),
            "orc_bsk_to_fourier": (None, [P, _u32p, _f64p]),
            "orc_modswitch_2N": (ctypes.c_uint32, [ctypes.c_uint32, ctypes.c_int]),
            "orc_gate_test_vector": (None, [ctypes.c_int, ctypes.c_int, ctypes.c_uint32, _u32p]),
            "orc_external_product32": (ctypes.c_double, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _f64p, _u32p, _u32p]),
            "orc_monic_update32": (None, [ctypes.c_int, ctypes.c_uint32, _u32p, _u32p]),
            "orc_blind_rotate32": (ctypes.c_double, [P, _f64p, _u32p, _u32p]),
            "orc_sample_extract32": (None, [ctypes.c_int, _u32p, ctypes.c_uint32, _u32p]),
            "orc_keyswitch32": (None, [P, _u32p, _u32p, _u32p]),
            "orc_gate_prep": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _u32p, _u32p, _u32p]),
            "orc_gate_bootstrap": (ctypes.c_double, [P, _f64p, _u32p, ctypes.c_int, _u32p, _u32p, _u32p]),
            "orc_func_property": (ctypes.c_int, [_u32p, ctypes.c_uint64]),
            "orc_bootstrap_func": (ctypes.c_int, [P, _f64p, _u32p, _u32p, ctypes.c_uint64, _u32p, ctypes.c_uint64,
                                                  _u32p]),
            "orc_eval_func": (ctypes.c_int, [P, _f64p, _u32p, _u32p, ctypes.c_uint64, _u32p, ctypes.c_uint32, _u32p]),
            "orc_eval_floor": (ctypes.c_int, [P, _f64p, _u32p, _u32p, ctypes.c_uint64, ctypes.c_uint32,
                                              ctypes.c_uint32, ctypes.c_uint64, _u32p]),
            "orc_eval_sign": (ctypes.c_int, [P, _f64p, _u32p, _u32p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64,
                                             ctypes.c_int, _u32p]),
            "orc_eval_decomp": (ctypes.c_int, [P, _f64p, _u32p, _u32p, ctypes.c_uint64, ctypes.c_uint32,
                                               ctypes.c_uint64, _u32p, ctypes.c_int]),
            "orc_aes_sbox": (ctypes.c_uint8, [ctypes.c_uint8]),
            "orc_aes128_expand": (None, [_u8p, _u8p]),
            "orc_aes128_encrypt": (None, [_u8p, _u8p, _u8p]),
            "orc_aes_inv_sbox": (ctypes.c_uint8, [ctypes.c_uint8]),
            "orc_aes_ttable": (ctypes.c_uint32, [ctypes.c_int, ctypes.c_int, ctypes.c_uint8]),
            "orc_aes128_decrypt": (None, [_u8p, _u8p, _u8p]),
            "orc_aes128_dec_schedule": (None, [_u8p, _u8p]),
            "orc_aes128_decrypt_ttable": (None, [_u8p, _u8p, _u8p]),
            "orc_aes128_encrypt_ttable": (None, [_u8p, _u8p, _u8p]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


# ---------------------------------------------------------------- helpers
def philox(key, ctr):
    out = np.zeros(4, np.uint32)
    lib().orc_philox4x32_10(np.asarray(key, np.uint32), np.asarray(ctr, np.uint32), out)
    return out


def decompose32(t, logB, l):
    d = np.zeros(l, np.int32)
    lib().orc_decompose32(int(t) & 0xFFFFFFFF, logB, l, d)
    return d


def decompose64(t, logB, l):
    d = np.zeros(l, np.int32)
    lib().orc_decompose64(int(t) & 0xFFFFFFFFFFFFFFFF, logB, l, d)
    return d


def nc_mul_exact32(x, y):
    x = np.ascontiguousarray(x, np.int32)
    y = np.ascontiguousarray(y, np.uint32)
    out = np.zeros(len(y), np.uint32)
    lib().orc_nc_mul_exact32(x, y, out, len(y))
    return out


def nc_mul_fft(x, y):
    x = np.ascontiguousarray(x, np.int64)
    y = np.ascontiguousarray(y, np.int64)
    out = np.zeros(len(y), np.int64)
    err = lib().orc_nc_mul_fft(x, y, out, len(y))
    return out, err


# the large-precision functions' small LWE modulus q and offset beta (this build's choice:
# q = N, so that arbitrary functions fit the doubled modulus 2q = 2N; beta = q / 16, about 10
# sigma of the bootstrap output noise at q; the plaintext space of EvalFunc is q / (2 beta) = 8)
FUNC_Q = 1024
FUNC_BETA = 64


def func_property(lut, q):
    """0 negacyclic, 1 periodic, 2 arbitrary (orc_func_property)."""
    return int(lib().orc_func_property(np.ascontiguousarray(lut, np.uint32), int(q)))


class GateKeys:
    """Gate-level key material generated by the oracle from a seed."""

    def __init__(self, seed, p=None, with_fourier=True):
        self.p = p if p is not None else gate_params()
        p = self.p
        L = lib()
        self.seed = seed
        self.s = np.zeros(p.n, np.int32)
        self.K = np.zeros(p.N, np.int32)
        L.orc_gen_lwe_sk(seed, p.n, self.s)
        L.orc_gen_rlwe_sk(seed, DOM_RLWE1_SK, p.N, self.K)
        rows = 2 * p.bk_l
        self.bsk = np.zeros(p.n * rows * 2 * p.N, np.uint32)
        L.orc_gen_bsk32(seed, ctypes.byref(p), self.s, self.K, self.bsk)
        B = 1 << p.ks_logB
        self.ksk = np.zeros(p.N * p.ks_l * B * (p.n + 1), np.uint32)
        L.orc_gen_ksk32(seed, ctypes.byref(p), self.s, self.K, self.ksk)
        self.bsk_f = None
        if with_fourier:
            self.bsk_f = np.zeros(p.n * rows * 2 * p.N, np.float64)
            L.orc_bsk_to_fourier(ctypes.byref(p), self.bsk, self.bsk_f)

    def encrypt(self, bit, id_, dom=DOM_ENC, seed=None):
        """LWE encryption of bit * 2^30 (OpenFHE p = 4 encoding)."""
        p = self.p
        ct = np.zeros(p.n + 1, np.uint32)
        lib().orc_lwe_encrypt32(self.seed if seed is None else seed, dom, id_, p.n, self.s,
                                (int(bit) << 30) & 0xFFFFFFFF, p.lwe_nshift, ct)
        return ct

    def phase(self, ct):
        return int(lib().orc_lwe_phase32(self.p.n, self.s, np.ascontiguousarray(ct, np.uint32)))

    def decrypt(self, ct):
        """round(phase * 4 / 2^32) mod 4, the OpenFHE p = 4 decryption."""
        ph = self.phase(ct)
        return ((ph + (1 << 29)) >> 30) & 3

    def bootstrap(self, gate, ct1, ct2):
        p = self.p
        out = np.zeros(p.n + 1, np.uint32)
        err = lib().orc_gate_bootstrap(ctypes.byref(p), self.bsk_f, self.ksk, gate,
                                       np.ascontiguousarray(ct1, np.uint32),
                                       np.ascontiguousarray(ct2, np.uint32), out)
        return out, err

    # ---- programmable bootstrap and large-precision functions (func_oracle.c) ----
    def encrypt_mod(self, value, mod, id_, seed=None):
        """LWE over Z_mod (a power of two <= 2^32) of `value` (in Z_mod): the 2^32-torus
        encryption of value * 2^32 / mod, every element rounded to Z_mod."""
        p = self.p
        ct = np.zeros(p.n + 1, np.uint32)
        sc = (1 << 32) // int(mod)
        lib().orc_lwe_encrypt32(self.seed if seed is None else seed, DOM_ENC, id_, p.n, self.s,
                                (int(value) * sc) & 0xFFFFFFFF, p.lwe_nshift, ct)
        if int(mod) == 1 << 32:
            return ct
        return (((ct.astype(np.uint64) + sc // 2) // sc) % int(mod)).astype(np.uint32)

    def phase_mod(self, ct, mod):
        ct = np.asarray(ct, np.uint64)
        return int((int(ct[-1]) - int(np.dot(ct[:-1], self.s.astype(np.uint64)))) % int(mod))

    def bootstrap_func(self, ct, ctmod, f, fmod):
        out = np.zeros(self.p.n + 1, np.uint32)
        rc = lib().orc_bootstrap_func(ctypes.byref(self.p), self.bsk_f, self.ksk, np.ascontiguousarray(ct, np.uint32),
                                      int(ctmod), np.ascontiguousarray(f, np.uint32), int(fmod), out)
        assert rc == 0
        return out

    def eval_func(self, ct, q, lut, beta=FUNC_BETA):
        out = np.zeros(self.p.n + 1, np.uint32)
        rc = lib().orc_eval_func(ctypes.byref(self.p), self.bsk_f, self.ksk, np.ascontiguousarray(ct, np.uint32),
                                 int(q), np.ascontiguousarray(lut, np.uint32), beta, out)
        assert rc == 0
        return out

    def eval_floor(self, ct, mod, roundbits=0, beta=FUNC_BETA, qsmall=FUNC_Q):
        out = np.zeros(self.p.n + 1, np.uint32)
        rc = lib().orc_eval_floor(ctypes.byref(self.p), self.bsk_f, self.ksk, np.ascontiguousarray(ct, np.uint32),
                                  int(mod), beta, roundbits, qsmall, out)
        assert rc == 0
        return out

    def eval_sign(self, ct, mod, scheme_switch=False, beta=FUNC_BETA, qsmall=FUNC_Q):
        out = np.zeros(self.p.n + 1, np.uint32)
        rc = lib().orc_eval_sign(ctypes.byref(self.p), self.bsk_f, self.ksk, np.ascontiguousarray(ct, np.uint32),
                                 int(mod), beta, qsmall, int(bool(scheme_switch)), out)
        assert rc == 0
        return out

    def eval_decomp(self, ct, mod, beta=FUNC_BETA, qsmall=FUNC_Q, max_digits=32):
        out = np.zeros((max_digits, self.p.n + 1), np.uint32)
        cnt = lib().orc_eval_decomp(ctypes.byref(self.p), self.bsk_f, self.ksk, np.ascontiguousarray(ct, np.uint32),
                                    int(mod), beta, qsmall, out, max_digits)
        assert cnt > 0
        return out[:cnt]


def aes_sbox(x):
    return int(lib().orc_aes_sbox(x))


def aes128_expand(key):
    rk = np.zeros(176, np.uint8)
    lib().orc_aes128_expand(np.frombuffer(bytes(key), np.uint8).copy(), rk)
    return rk


def aes_inv_sbox(x):
    return int(lib().orc_aes_inv_sbox(x))


def aes_ttable(dec, r, x):
    """Te_r[x] (dec = 0) or Td_r[x] (dec = 1) as a big-endian 32-bit word."""
    return int(lib().orc_aes_ttable(int(dec), r, x))


def _u8(b):
    return np.frombuffer(bytes(b), np.uint8).copy()


def aes128_decrypt(key, block):
    out = np.zeros(16, np.uint8)
    lib().orc_aes128_decrypt(aes128_expand(key), _u8(block), out)
    return bytes(out)


def aes128_dec_schedule(key):
    dk = np.zeros(176, np.uint8)
    lib().orc_aes128_dec_schedule(aes128_expand(key), dk)
    return dk


def aes128_decrypt_ttable(key, block):
    out = np.zeros(16, np.uint8)
    lib().orc_aes128_decrypt_ttable(aes128_dec_schedule(key), _u8(block), out)
    return bytes(out)


def aes128_encrypt_ttable(key, block):
    out = np.zeros(16, np.uint8)
    lib().orc_aes128_encrypt_ttable(aes128_expand(key), _u8(block), out)
    return bytes(out)


def aes128_encrypt(key, block):
    rk = aes128_expand(key)
    out = np.zeros(16, np.uint8)
    lib().orc_aes128_encrypt(rk, np.frombuffer(bytes(block), np.uint8).copy(), out)
    return bytes(out)


def gate_blind_rotate_extract(keys, gate, ct1, ct2):
    """Oracle blind rotation + sample extraction (+Q/8+1), i.e. the LWE under the
    RLWE key before key switching; returns (lweN, max FFT rounding error)."""
    L = lib()
    p = keys.p
    n, N = p.n, p.N
    prep = np.zeros(n + 1, np.uint32)
    if L.orc_gate_prep(gate, n, np.ascontiguousarray(ct1, np.uint32), np.ascontiguousarray(ct2, np.uint32), prep):
        raise ValueError("unsupported gate")
    at = np.array([L.orc_modswitch_2N(int(v), N) for v in prep[:n]], np.uint32)
    bt = L.orc_modswitch_2N(int(prep[n]), N)
    acc = np.zeros(2 * N, np.uint32)
    tv = np.zeros(N, np.uint32)
    L.orc_gate_test_vector(N, gate, bt, tv)
    acc[N:] = tv
    err = L.orc_blind_rotate32(ctypes.byref(p), keys.bsk_f, at, acc)
    lwe = np.zeros(N + 1, np.uint32)
    L.orc_sample_extract32(N, acc, 0, lwe)
    lwe[N] = (int(lwe[N]) + (1 << 29) + 1) & 0xFFFFFFFF
    return lwe, err


def keyswitch32(keys, lweN):
    out = np.zeros(keys.p.n + 1, np.uint32)
    lib().orc_keyswitch32(ctypes.byref(keys.p), keys.ksk, np.ascontiguousarray(lweN, np.uint32), out)
    return out


# ====================================================================== level 2
class CbsParams(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("N", ctypes.c_int), ("bk_logB", ctypes.c_int), ("bk_l", ctypes.c_int),
                ("cb_logB", ctypes.c_int), ("cb_l", ctypes.c_int), ("pf_logB", ctypes.c_int), ("pf_l", ctypes.c_int),
                ("ks_logB", ctypes.c_int), ("ks_l", ctypes.c_int), ("lwe_nshift", ctypes.c_int),
                ("rlwe_nshift", ctypes.c_int), ("pf_nshift", ctypes.c_int)]


def cbs_params(n=636, N=2048, bk_logB=9, rlwe_nshift=19):
    """Circuit-bootstrapping level: N=2048, PBS 2^9 x 4 (torus-64, sigma 2^-51), PFKS 2^4 x 8
    (torus-32 key, sigma 2^-30), CB 2^8 x 2 (torus-32 RGSW), KS 2^2 x 8 to the level-0 key
    (sigma 2^-15)."""
    return CbsParams(n, N, bk_logB, 4, 8, 2, 4, 8, 2, 8, 15, rlwe_nshift, 30)


def cbs_params_htss(n=636, N=2048):
    """The HomTrace + SchemeSwitch form: PBS 2^10 x 4 with BSK2 sigma 2^-54 (PBS noise near
    2^-32, which the trace multiplies by N at the constant coefficient), HT / SS gadget
    2^8 x 5 with the same key noise."""
    return cbs_params(n, N, bk_logB=10, rlwe_nshift=22)


HT_LOGB, HT_L = 8, 5


_cbs_sigs_done = False


def _cbs_lib():
    global _cbs_sigs_done
    L = lib()
    if not _cbs_sigs_done:
        P = ctypes.POINTER(CbsParams)
        G = ctypes.POINTER(GateParams)
        sig = {
            "orc_to_fourier64": (None, [_u64p, ctypes.c_int, _f64p]),
            "orc_rlwe_encrypt64": (None, [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, _i32p,
                                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, _u64p]),
            "orc_rgsw_encrypt64": (None, [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, _i32p,
                                          ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _u64p]),
            "orc_gen_bsk64": (None, [ctypes.c_uint64, P, _i32p, _i32p, _u64p]),
            "orc_bsk64_to_fourier": (None, [P, _u64p, _f64p]),
            "orc_external_product64": (ctypes.c_double, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _f64p, _u64p, _u64p]),
            "orc_monic_update64": (None, [ctypes.c_int, ctypes.c_uint32, _u64p, _u64p]),
            "orc_blind_rotate64": (ctypes.c_double, [P, _f64p, _u32p, _u64p]),
            "orc_sample_extract64": (None, [ctypes.c_int, _u64p, ctypes.c_uint32, _u64p]),
            "orc_special_ms": (ctypes.c_uint32, [ctypes.c_uint32, ctypes.c_int, ctypes.c_int]),
            "orc_pbs_manylut": (ctypes.c_double, [P, _f64p, _u32p, _u64p]),
            "orc_pbs_manylut_mb": (ctypes.c_double, [P, _f64p, ctypes.c_int, _u32p, _u64p]),
            "orc_bootstrap_many_lut": (ctypes.c_double, [P, _f64p, ctypes.c_int, _u32p, _u64p, ctypes.c_int,
                                                         _u64p]),
            "orc_blind_rotate_lut64": (ctypes.c_double, [P, _f64p, ctypes.c_int, _u32p, _u64p, ctypes.c_int,
                                                         _u64p]),
            "orc_automorphism64": (None, [ctypes.c_int, _u64p, ctypes.c_uint32, _u64p]),
            "orc_gen_ht_key": (None, [ctypes.c_uint64, P, _i32p, _u64p]),
            "orc_gen_ss_key": (None, [ctypes.c_uint64, P, _i32p, _u64p]),
            "orc_ksk64_to_fourier": (None, [P, ctypes.c_int, _u64p, _f64p]),
            "orc_homtrace64": (ctypes.c_double, [P, _f64p, _u64p]),
            "orc_schemeswitch64": (ctypes.c_double, [P, _f64p, _u64p, _u64p]),
            "orc_cbs_htss": (ctypes.c_double, [P, _f64p, _f64p, _f64p, _u32p, _u32p]),
            "orc_mb2_indicator": (ctypes.c_uint32, [_i32p, ctypes.c_int, ctypes.c_int]),
            "orc_gen_bsk64_mb2_row": (None, [ctypes.c_uint64, P, _i32p, _i32p, ctypes.c_uint32, ctypes.c_int,
                                             _u64p]),
            "orc_gen_bsk64_mb2": (None, [ctypes.c_uint64, P, _i32p, _i32p, _u64p]),
            "orc_rgsw64_to_fourier": (None, [P, ctypes.c_long, _u64p, _f64p]),
            "orc_blind_rotate64_mb2": (ctypes.c_double, [P, _f64p, _u32p, _u64p]),
            "orc_pfks_id": (ctypes.c_uint32, [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
            "orc_gen_pfks_entry": (None, [ctypes.c_uint64, P, _i32p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, _u32p]),
            "orc_pfks": (None, [ctypes.c_uint64, P, _i32p, ctypes.c_int, _u64p, _u32p]),
            "orc_cbs": (ctypes.c_double, [ctypes.c_uint64, P, _i32p, _f64p, _u32p, _u32p]),
            "orc_gen_pfkl_entry": (None, [ctypes.c_uint64, P, _i32p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          _u64p]),
            "orc_pfks_linear": (None, [ctypes.c_uint64, P, _i32p, ctypes.c_int, _u64p, _u32p]),
            "orc_cbs_linear": (ctypes.c_double, [ctypes.c_uint64, P, _i32p, _f64p, _u32p, _u32p]),
            "orc_rgsw_to_fourier32": (None, [ctypes.c_int, ctypes.c_int, _u32p, _f64p]),
            "orc_cmux32": (ctypes.c_double, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _f64p, _u32p, _u32p, _u32p]),
            "orc_vertical_packing": (ctypes.c_double, [P, ctypes.c_int, _f64p, _u32p, _u32p]),
            "orc_ks2": (None, [P, _u32p, _u64p, _u32p]),
            "orc_gen_ksk2": (None, [ctypes.c_uint64, P, _i32p, _i32p, _u32p]),
            "orc_gen_kskl_entry": (None, [ctypes.c_uint64, P, _i32p, _i32p, ctypes.c_int, ctypes.c_int, _u32p]),
            "orc_gen_kskl": (None, [ctypes.c_uint64, P, _i32p, _i32p, _u32p]),
            "orc_ks2_linear32": (None, [P, _u32p, _u32p, _u32p]),
            "orc_ks2_linear": (None, [P, _u32p, _u64p, _u32p]),
            "orc_gen_ksk32_dom": (None, [ctypes.c_uint64, ctypes.c_uint32, G, _i32p, _i32p, _u32p]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _cbs_sigs_done = True
    return L


def nc_mul_exact64(x, y):
    x = np.ascontiguousarray(x, np.int32)
    y = np.ascontiguousarray(y, np.uint64)
    out = np.zeros(len(y), np.uint64)
    lib().orc_nc_mul_exact64(x, y, out, len(y))
    return out


def rlwe_phase32(K, ct):
    """b - a*K over Z/2^32 for ct = [a | b]."""
    N = len(K)
    ct = np.asarray(ct, np.uint32)
    return (ct[N:] - nc_mul_exact32(K, ct[:N])).astype(np.uint32)


def special_ms(v, N, bw):
    """round(v 2N / (2^32 2^bw)) 2^bw mod 2N (orc_special_ms)."""
    return int(_cbs_lib().orc_special_ms(int(v) & 0xFFFFFFFF, N, bw))


def automorphism64(poly, k):
    """poly(X^k) mod X^N + 1 (k odd), uint64 coefficients."""
    poly = np.ascontiguousarray(poly, np.uint64)
    out = np.zeros_like(poly)
    _cbs_lib().orc_automorphism64(len(poly), poly, k, out)
    return out


def rlwe_phase64(K, ct):
    """b - a*K over Z/2^64 for ct = [a | b]."""
    N = len(K)
    return (ct[N:] - nc_mul_exact64(K, ct[:N])).astype(np.uint64)


def to_fourier64(poly):
    N = len(poly)
    out = np.zeros(2 * N, np.float64)
    _cbs_lib().orc_to_fourier64(np.ascontiguousarray(poly, np.uint64), N, out)
    return out


def signed64(v):
    """uint64 array -> signed int64 array (two's complement view) as Python-friendly ints."""
    return np.asarray(v, np.uint64).view(np.int64)


class CbsKeys:
    """Circuit-bootstrapping key material from a seed (the PFKS key is never
    materialised: its entries are regenerated on demand by orc_pfks)."""

    def __init__(self, seed, p=None, s=None, with_bsk=True):
        self.p = p if p is not None else cbs_params()
        p = self.p
        L = _cbs_lib()
        self.seed = seed
        self.s = np.zeros(p.n, np.int32)
        if s is None:
            L.orc_gen_lwe_sk(seed, p.n, self.s)
        else:
            self.s[:] = s
        self.K2 = np.zeros(p.N, np.int32)
        L.orc_gen_rlwe_sk(seed, DOM_RLWE2_SK, p.N, self.K2)
        self.bsk = self.bsk_f = None
        if with_bsk:
            rows = 2 * p.bk_l
            self.bsk = np.zeros(p.n * rows * 2 * p.N, np.uint64)
            L.orc_gen_bsk64(seed, ctypes.byref(p), self.s, self.K2, self.bsk)
            self.bsk_f = np.zeros(p.n * rows * 2 * 2 * p.N, np.float64)
            L.orc_bsk64_to_fourier(ctypes.byref(p), self.bsk, self.bsk_f)
        self._ksk2 = None
        self._bsk_mb2 = self._bsk_mb2_f = None

    def bsk_mb2(self):
        """Pattern key of two key bits per step: [n/2][3][2l][2][N] u64 (RGSW q = 3k + p of
        the indicator of pattern p of the key-bit pair k: (1,0), (0,1), (1,1))."""
        if self._bsk_mb2 is None:
            p = self.p
            self._bsk_mb2 = np.zeros((p.n // 2) * 3 * 2 * p.bk_l * 2 * p.N, np.uint64)
            _cbs_lib().orc_gen_bsk64_mb2(self.seed, ctypes.byref(p), self.s, self.K2, self._bsk_mb2)
        return self._bsk_mb2

    def bsk_mb2_fourier(self):
        if self._bsk_mb2_f is None:
            p = self.p
            bsk = self.bsk_mb2()
            self._bsk_mb2_f = np.zeros(bsk.size * 2, np.float64)
            _cbs_lib().orc_rgsw64_to_fourier(ctypes.byref(p), (p.n // 2) * 3, bsk, self._bsk_mb2_f)
        return self._bsk_mb2_f

    def bsk_mb2_row(self, q, r):
        p = self.p
        row = np.zeros(2 * p.N, np.uint64)
        _cbs_lib().orc_gen_bsk64_mb2_row(self.seed, ctypes.byref(p), self.s, self.K2, q, r, row)
        return row

    def pbs_mb2(self, ct32):
        """Exact multi-value PBS with two key bits per step (the pattern key)."""
        p = self.p
        out = np.zeros(2 * (p.N + 1), np.uint64)
        err = _cbs_lib().orc_pbs_manylut_mb(ctypes.byref(p), self.bsk_mb2_fourier(), 2,
                                            np.ascontiguousarray(ct32, np.uint32), out)
        return out.reshape(2, p.N + 1), err

    def bootstrap_many_lut(self, ct32, lut, bw, mb=1):
        """The reference's gpu_BootstrapManyLUT: RLWE [2][N] (mask, body) of LUT X^-phase~
        plus Gpow_j / 2 on body coefficient j < cb_l (mb = 2: the pattern key)."""
        p = self.p
        out = np.zeros(2 * p.N, np.uint64)
        bskf = self.bsk_mb2_fourier() if mb == 2 else self.bsk_f
        err = _cbs_lib().orc_bootstrap_many_lut(ctypes.byref(p), bskf, mb, np.ascontiguousarray(ct32, np.uint32),
                                                np.ascontiguousarray(lut, np.uint64), int(bw), out)
        return out.reshape(2, p.N), err

    def _htss_keys(self):
        if getattr(self, "_htk", None) is None:
            p, L = self.p, _cbs_lib()
            R = int(np.log2(p.N))
            self.ht_key = np.zeros(R * HT_L * 2 * p.N, np.uint64)
            L.orc_gen_ht_key(self.seed, ctypes.byref(p), self.K2, self.ht_key)
            self.ss_key = np.zeros(HT_L * 2 * p.N, np.uint64)
            L.orc_gen_ss_key(self.seed, ctypes.byref(p), self.K2, self.ss_key)
            self._htk = np.zeros(R * 2 * HT_L * 2 * 2 * p.N, np.float64)
            L.orc_ksk64_to_fourier(ctypes.byref(p), R, self.ht_key, self._htk)
            self._ssk = np.zeros(2 * HT_L * 2 * 2 * p.N, np.float64)
            L.orc_ksk64_to_fourier(ctypes.byref(p), 1, self.ss_key, self._ssk)
        return self._htk, self._ssk

    def homtrace(self, ct):
        """HomTrace of an RLWE [2][N] (torus-64): all log2 N rounds; returns (ct', err)."""
        htk, _ = self._htss_keys()
        out = np.ascontiguousarray(ct, np.uint64).reshape(-1).copy()
        err = _cbs_lib().orc_homtrace64(ctypes.byref(self.p), htk, out)
        return out.reshape(2, self.p.N), err

    def schemeswitch(self, ct):
        """RLWE of -K2 m from an RLWE of m (torus-64)."""
        _, ssk = self._htss_keys()
        out = np.zeros(2 * self.p.N, np.uint64)
        err = _cbs_lib().orc_schemeswitch64(ctypes.byref(self.p), ssk, np.ascontiguousarray(ct, np.uint64).reshape(-1),
                                            out)
        return out.reshape(2, self.p.N), err

    def cbs_htss(self, ct32):
        """Circuit bootstrap through HomTrace + SchemeSwitch: torus-32 RGSW [4][2][N] (row 2j SS,
        2j + 1 HT)."""
        htk, ssk = self._htss_keys()
        p = self.p
        rgsw = np.zeros(2 * p.cb_l * 2 * p.N, np.uint32)
        err = _cbs_lib().orc_cbs_htss(ctypes.byref(p), self.bsk_f, htk, ssk, np.ascontiguousarray(ct32, np.uint32),
                                      rgsw)
        return rgsw, err

    def blind_rotate_mb2(self, a_tilde, acc):
        """Exact two-bit blind rotation of acc [2N] u64 (in place) by the rotations a_tilde [n]."""
        err = _cbs_lib().orc_blind_rotate64_mb2(ctypes.byref(self.p), self.bsk_mb2_fourier(),
                                                np.ascontiguousarray(a_tilde, np.uint32), acc)
        return err

    @property
    def ksk2(self):
        if self._ksk2 is None:
            p = self.p
            B = 1 << p.ks_logB
            self._ksk2 = np.zeros(p.N * p.ks_l * B * (p.n + 1), np.uint32)
            _cbs_lib().orc_gen_ksk2(self.seed, ctypes.byref(p), self.s, self.K2, self._ksk2)
        return self._ksk2

    @property
    def kskl(self):
        """Linear N -> n key switching key [N][ks_l][n+1] u32 (one row per (i, j))."""
        if getattr(self, "_kskl", None) is None:
            p = self.p
            self._kskl = np.zeros(p.N * p.ks_l * (p.n + 1), np.uint32)
            _cbs_lib().orc_gen_kskl(self.seed, ctypes.byref(p), self.s, self.K2, self._kskl)
        return self._kskl

    def kskl_entry(self, i, j):
        p = self.p
        row = np.zeros(p.n + 1, np.uint32)
        _cbs_lib().orc_gen_kskl_entry(self.seed, ctypes.byref(p), self.s, self.K2, int(i), int(j), row)
        return row

    def encrypt_bit(self, bit, id_, seed=None):
        """Level-0 LWE of bit * 2^31 (p = 2)."""
        p = self.p
        ct = np.zeros(p.n + 1, np.uint32)
        lib().orc_lwe_encrypt32(self.seed if seed is None else seed, DOM_ENC, id_, p.n, self.s,
                                (int(bit) << 31) & 0xFFFFFFFF, p.lwe_nshift, ct)
        return ct

    def phase32(self, ct):
        return int(lib().orc_lwe_phase32(self.p.n, self.s, np.ascontiguousarray(ct, np.uint32)))

    def decrypt_bit(self, ct):
        return ((self.phase32(ct) + (1 << 30)) >> 31) & 1

    def lwe64_phase(self, lwe):
        N = self.p.N
        return (int(lwe[N]) - int(np.dot(np.asarray(lwe[:N], np.uint64), self.K2.astype(np.uint64)))) % 2 ** 64

    def pbs(self, ct32):
        p = self.p
        out = np.zeros(2 * (p.N + 1), np.uint64)
        err = _cbs_lib().orc_pbs_manylut(ctypes.byref(p), self.bsk_f, np.ascontiguousarray(ct32, np.uint32), out)
        return out.reshape(2, p.N + 1), err

    def pfks(self, u, lwe64):
        """Torus-32 RLWE_K2 of f_u(phase of lwe64)."""
        p = self.p
        out = np.zeros(2 * p.N, np.uint32)
        _cbs_lib().orc_pfks(self.seed, ctypes.byref(p), self.K2, u, np.ascontiguousarray(lwe64, np.uint64), out)
        return out

    def pfks_entry(self, u, i, t, v):
        p = self.p
        out = np.zeros(2 * p.N, np.uint32)
        _cbs_lib().orc_gen_pfks_entry(self.seed, ctypes.byref(p), self.K2, u, i, t, v, out)
        return out

    def pfks_linear(self, u, lwe64):
        """Torus-32 RLWE_K2 of f_u(phase of lwe64) through the linear 40-bit key."""
        p = self.p
        out = np.zeros(2 * p.N, np.uint32)
        _cbs_lib().orc_pfks_linear(self.seed, ctypes.byref(p), self.K2, u, np.ascontiguousarray(lwe64, np.uint64),
                                   out)
        return out

    def pfkl_entry(self, u, i, t):
        """Linear-key entry (u, i, t): [2][N] values mod 2^40."""
        p = self.p
        out = np.zeros(2 * p.N, np.uint64)
        _cbs_lib().orc_gen_pfkl_entry(self.seed, ctypes.byref(p), self.K2, u, i, t, out)
        return out

    def cbs(self, ct32, linear=False):
        p = self.p
        rgsw = np.zeros(2 * p.cb_l * 2 * p.N, np.uint32)
        fn = _cbs_lib().orc_cbs_linear if linear else _cbs_lib().orc_cbs
        err = fn(self.seed, ctypes.byref(p), self.K2, self.bsk_f, np.ascontiguousarray(ct32, np.uint32), rgsw)
        return rgsw, err

    def rgsw_encrypt(self, m, id0, logB=None, l=None):
        """Torus-32 RGSW_K2(m): a torus-64 encryption rounded to its top 32 bits."""
        p = self.p
        logB = p.cb_logB if logB is None else logB
        l = p.cb_l if l is None else l
        rgsw = np.zeros(2 * l * 2 * p.N, np.uint64)
        _cbs_lib().orc_rgsw_encrypt64(self.seed, DOM_ENC, id0, p.N, self.K2, int(m), logB, l, p.rlwe_nshift, rgsw)
        return ((rgsw + np.uint64(1 << 31)) >> np.uint64(32)).astype(np.uint32)

    def rgsw_fourier(self, rgsw):
        p = self.p
        rows = len(rgsw) // (2 * p.N)
        out = np.zeros(rows * 2 * p.N, np.float64)
        _cbs_lib().orc_rgsw_to_fourier32(p.N, rows, np.ascontiguousarray(rgsw, np.uint32), out)
        return out

    def vertical_packing(self, rgsw_fs, table_poly, stride=8):
        p = self.p
        acc = np.zeros(2 * p.N, np.uint32)
        err = _cbs_lib().orc_vertical_packing(ctypes.byref(p), stride, np.ascontiguousarray(np.concatenate(rgsw_fs)),
                                              np.ascontiguousarray(table_poly, np.uint32), acc)
        return acc, err

    def ks2(self, lwe64):
        out = np.zeros(self.p.n + 1, np.uint32)
        _cbs_lib().orc_ks2(ctypes.byref(self.p), self.ksk2, np.ascontiguousarray(lwe64, np.uint64), out)
        return out

    def ks2_linear(self, lwe64):
        """The N -> n key switch of a torus-64 LWE_K2 with the linear key and signed digits."""
        out = np.zeros(self.p.n + 1, np.uint32)
        _cbs_lib().orc_ks2_linear(ctypes.byref(self.p), self.kskl, np.ascontiguousarray(lwe64, np.uint64), out)
        return out

    def ks2_linear32(self, lwe32):
        """The same for a torus-32 LWE_K2 [N+1] u32 (the GPU key switch's input)."""
        out = np.zeros(self.p.n + 1, np.uint32)
        _cbs_lib().orc_ks2_linear32(ctypes.byref(self.p), self.kskl, np.ascontiguousarray(lwe32, np.uint32), out)
        return out

    def extract32(self, acc, idx):
        lwe = np.zeros(self.p.N + 1, np.uint32)
        lib().orc_sample_extract32(self.p.N, np.ascontiguousarray(acc, np.uint32), idx, lwe)
        return lwe

    def extract64(self, acc, idx):
        lwe = np.zeros(self.p.N + 1, np.uint64)
        _cbs_lib().orc_sample_extract64(self.p.N, np.ascontiguousarray(acc, np.uint64), idx, lwe)
        return lwe
