"""ctypes bindings of libtfhe_amd.so (the gfx950 TFHE bootstrapping library).

Device buffers are passed as raw pointers; torch tensors on ``cuda`` are the
convenient allocator (``tensor.data_ptr()``).  The library is loaded from this
directory; a missing library raises instead of falling back to anything.
"""
import ctypes
import os
import subprocess

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libtfhe_amd.so")

TFHE_OK, TFHE_E_PARAM, TFHE_E_SHAPE, TFHE_E_HIP, TFHE_E_CAPACITY, TFHE_E_STATE = range(6)
OR, AND, NOR, NAND, XOR_FAST, XNOR_FAST = 0, 1, 2, 3, 4, 5
XOR, XNOR = 11, 12


class TfheError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"tfhe error {code}: {msg}")
        self.code = code


class GateParams(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("N", ctypes.c_int), ("bk_logB", ctypes.c_int),
                ("bk_l", ctypes.c_int), ("ks_logB", ctypes.c_int), ("ks_l", ctypes.c_int),
                ("lwe_nshift", ctypes.c_int), ("rlwe_nshift", ctypes.c_int)]


def gate_params():
    """n=636, N=1024, BSK base 2^7 l=3, KS base 2^2 l=8, LWE sigma 2^-15, RLWE sigma 2^-25."""
    return GateParams(636, 1024, 7, 3, 2, 8, 15, 25)


def build():
    """Compile libtfhe_amd.so for gfx950 in-tree (hipcc cross-compiles without a GPU)."""
    subprocess.run(["make", "-s", "-j8", "-C", PKG_DIR], check=True)


_lib = None
_vp = ctypes.c_void_p
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")

# exported symbol -> (restype, argtypes)
SIGNATURES = {
    "tfhe_last_error": (ctypes.c_char_p, []),
    "tfhe_gate_ctx_create": (ctypes.c_int, [ctypes.POINTER(GateParams), ctypes.c_uint32, ctypes.POINTER(_vp)]),
    "tfhe_gate_ctx_destroy": (ctypes.c_int, [_vp]),
    "tfhe_gate_bt_keygen": (ctypes.c_int, [_vp, _i32p, ctypes.c_uint64]),
    "tfhe_gate_set_keys": (ctypes.c_int, [_vp, _u32p, _u32p]),
    "tfhe_gate_get_rlwe_sk": (ctypes.c_int, [_vp, _i32p]),
    "tfhe_gate_eval_batch": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, ctypes.c_uint32, _vp, _vp]),
    "tfhe_gate_eval_plain": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp, _vp]),
    "tfhe_gate_blind_rotate_batch": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, ctypes.c_uint32, _vp, _vp]),
    "tfhe_gate_keyswitch_batch": (ctypes.c_int, [_vp, _vp, ctypes.c_uint32, _vp, _vp]),
    "tfhe_polymul1024_batch": (ctypes.c_int, [_vp, _vp, ctypes.c_uint32, _vp, _vp]),
    "tfhe_gate_external_product": (ctypes.c_int, [_vp, ctypes.c_uint32, _vp, _vp, _vp]),
    "tfhe_gen_bsk32": (ctypes.c_int, [ctypes.POINTER(GateParams), ctypes.c_uint64, _vp, _vp, _vp, _vp]),
    "tfhe_gen_ksk32": (ctypes.c_int, [ctypes.POINTER(GateParams), ctypes.c_uint64, _vp, _vp, _vp, _vp]),
    "tfhe_lwe_keygen": (None, [ctypes.c_int, ctypes.c_uint64, _i32p]),
    "tfhe_rlwe_keygen": (None, [ctypes.c_int, ctypes.c_uint64, ctypes.c_uint32, _i32p]),
    "tfhe_lwe_encrypt": (None, [ctypes.c_int, _i32p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                                ctypes.c_uint32, ctypes.c_int, _u32p]),
    "tfhe_lwe_decrypt": (ctypes.c_uint32, [ctypes.c_int, _i32p, _u32p, ctypes.c_uint32]),
}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise TfheError(TFHE_E_STATE, f"{LIB_PATH} is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != TFHE_OK:
        raise TfheError(rc, lib().tfhe_last_error().decode())


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream(s):
    return ctypes.c_void_p(0 if s is None else s)


def polymul1024_batch(d_x, d_y, d_out, stream=None):
    check(lib().tfhe_polymul1024_batch(_ptr(d_x), _ptr(d_y), d_x.shape[0], _ptr(d_out), _stream(stream)))


# ------------------------------------------------------------------ host helpers
def lwe_keygen(n, seed):
    sk = np.zeros(n, np.int32)
    lib().tfhe_lwe_keygen(n, seed, sk)
    return sk


def lwe_encrypt(sk, seed, id_, m, p=4, nshift=15):
    ct = np.zeros(len(sk) + 1, np.uint32)
    lib().tfhe_lwe_encrypt(len(sk), sk, seed, id_, m, p, nshift, ct)
    return ct


def lwe_decrypt(sk, ct, p=4):
    return int(lib().tfhe_lwe_decrypt(len(sk), sk, np.ascontiguousarray(ct, np.uint32), p))


class GateContext:
    """Gate bootstrapping context (GPUBinFHEContext counterpart)."""

    def __init__(self, max_batch=4096, params=None):
        self.p = params if params is not None else gate_params()
        h = ctypes.c_void_p()
        check(lib().tfhe_gate_ctx_create(ctypes.byref(self.p), max_batch, ctypes.byref(h)))
        self.h = h
        self.max_batch = max_batch

    def close(self):
        if self.h:
            lib().tfhe_gate_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def bt_keygen(self, lwe_sk, seed):
        check(lib().tfhe_gate_bt_keygen(self.h, np.ascontiguousarray(lwe_sk, np.int32), seed))

    def set_keys(self, bsk, ksk):
        check(lib().tfhe_gate_set_keys(self.h, np.ascontiguousarray(bsk, np.uint32),
                                       np.ascontiguousarray(ksk, np.uint32)))

    def rlwe_sk(self):
        K = np.zeros(self.p.N, np.int32)
        check(lib().tfhe_gate_get_rlwe_sk(self.h, K))
        return K

    def eval_batch(self, gate, d_ct1, d_ct2, d_out, batch=None, stream=None):
        b = d_ct1.shape[0] if batch is None else batch
        check(lib().tfhe_gate_eval_batch(self.h, gate, _ptr(d_ct1), _ptr(d_ct2), b, _ptr(d_out), _stream(stream)))

    def eval_plain(self, gate, d_ct1, d_ct2, d_out, stream=None):
        check(lib().tfhe_gate_eval_plain(self.h, gate, _ptr(d_ct1), _ptr(d_ct2), _ptr(d_out), _stream(stream)))

    def blind_rotate_batch(self, gate, d_ct1, d_ct2, d_lweN, stream=None):
        check(lib().tfhe_gate_blind_rotate_batch(self.h, gate, _ptr(d_ct1), _ptr(d_ct2), d_ct1.shape[0],
                                                 _ptr(d_lweN), _stream(stream)))

    def external_product(self, key_index, d_acc, d_out, stream=None):
        check(lib().tfhe_gate_external_product(self.h, key_index, _ptr(d_acc), _ptr(d_out), _stream(stream)))

    def keyswitch_batch(self, d_lweN, d_out, stream=None):
        check(lib().tfhe_gate_keyswitch_batch(self.h, _ptr(d_lweN), d_lweN.shape[0], _ptr(d_out), _stream(stream)))
