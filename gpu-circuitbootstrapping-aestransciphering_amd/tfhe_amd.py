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


class RgswBatchView(ctypes.Structure):
    """tfhe_rgsw_batch_view: borrowed circuit-bootstrap outputs held in the context."""
    _fields_ = [("d_rgsw", ctypes.c_void_p), ("batch", ctypes.c_uint32), ("digits", ctypes.c_uint32),
                ("N", ctypes.c_uint32), ("per_ciphertext_elems", ctypes.c_size_t), ("stream", ctypes.c_void_p)]


class CbsParams(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("N", ctypes.c_int), ("bk_logB", ctypes.c_int), ("bk_l", ctypes.c_int),
                ("cb_logB", ctypes.c_int), ("cb_l", ctypes.c_int), ("pf_logB", ctypes.c_int), ("pf_l", ctypes.c_int),
                ("ks_logB", ctypes.c_int), ("ks_l", ctypes.c_int), ("lwe_nshift", ctypes.c_int),
                ("rlwe_nshift", ctypes.c_int), ("pf_nshift", ctypes.c_int), ("pbs_approx", ctypes.c_int),
                ("pbs_multibit", ctypes.c_int), ("pfks_linear", ctypes.c_int), ("lwe_to_rgsw", ctypes.c_int),
                ("ks_linear", ctypes.c_int)]


def cbs_params(pbs_approx=0, pbs_multibit=1, pfks_linear=0, lwe_to_rgsw=0, ks_linear=0):
    """n=636, N=2048; PBS 2^9 x 4 on the 64-bit torus (sigma 2^-51), PFKS 2^4 x 8 with
    a torus-32 key (sigma 2^-30), CB 2^8 x 2 (torus-32 RGSW), KS 2^2 x 8 to the
    level-0 key (sigma 2^-15).  pbs_approx=1 selects the one-product FP64 PBS
    (rounding error about 2^-33 of the torus) instead of the exact two-limb one;
    pbs_multibit=2 (with pbs_approx=1) blind-rotates two key bits per step; pfks_linear=1
    switches keys through the linear 40-bit PFKS key on the matrix cores.  lwe_to_rgsw=1
    converts by HomTrace + SchemeSwitch (the reference's conversion): exact PBS with a
    2^10 x 4 gadget and BSK2 / HT / SS keys at sigma 2^-54, no PFKS key.  ks_linear=1 runs
    the N -> n key switch with a linear key and signed digits as an int8 GEMM on the
    matrix cores."""
    if lwe_to_rgsw:
        return CbsParams(636, 2048, 10, 4, 8, 2, 4, 8, 2, 8, 15, 22, 30, 0, min(int(pbs_multibit), 1), 0, 1,
                         int(ks_linear))
    return CbsParams(636, 2048, 9, 4, 8, 2, 4, 8, 2, 8, 15, 19, 30, int(pbs_approx), int(pbs_multibit),
                     int(pfks_linear), 0, int(ks_linear))


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
    "tfhe_gate_eval_batch_noalloc": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, ctypes.c_uint32, _vp, _vp]),
    "tfhe_gate_eval_plain": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp, _vp]),
    "tfhe_gate_blind_rotate_batch": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, ctypes.c_uint32, _vp, _vp]),
    "tfhe_gate_keyswitch_batch": (ctypes.c_int, [_vp, _vp, ctypes.c_uint32, _vp, _vp]),
    "tfhe_func_property": (ctypes.c_int, [_u32p, ctypes.c_uint64]),
    "tfhe_gate_bootstrap_func_batch": (ctypes.c_int, [_vp, _vp, ctypes.c_uint32, ctypes.c_uint64, _u32p,
                                                      ctypes.c_uint64, _vp, _vp]),
    "tfhe_gate_eval_func_batch": (ctypes.c_int, [_vp, _vp, ctypes.c_uint32, ctypes.c_uint64, _u32p, _vp, _vp]),
    "tfhe_gate_eval_floor_batch": (ctypes.c_int, [_vp, _vp, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint32,
                                                  _vp, _vp]),
    "tfhe_gate_eval_sign_batch": (ctypes.c_int, [_vp, _vp, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, _vp,
                                                 _vp]),
    "tfhe_gate_eval_decomp_batch": (ctypes.c_int, [_vp, _vp, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint32, _vp,
                                                   ctypes.POINTER(ctypes.c_uint32),
                                                   np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS"),
                                                   _vp]),
    "tfhe_polymul1024_batch": (ctypes.c_int, [_vp, _vp, ctypes.c_uint32, _vp, _vp]),
    "tfhe_gate_external_product": (ctypes.c_int, [_vp, ctypes.c_uint32, _vp, _vp, _vp]),
    "tfhe_gen_bsk32": (ctypes.c_int, [ctypes.POINTER(GateParams), ctypes.c_uint64, _vp, _vp, _vp, _vp]),
    "tfhe_gen_ksk32": (ctypes.c_int, [ctypes.POINTER(GateParams), ctypes.c_uint64, _vp, _vp, _vp, _vp]),
    "tfhe_lwe_keygen": (None, [ctypes.c_int, ctypes.c_uint64, _i32p]),
    "tfhe_rlwe_keygen": (None, [ctypes.c_int, ctypes.c_uint64, ctypes.c_uint32, _i32p]),
    "tfhe_lwe_encrypt": (None, [ctypes.c_int, _i32p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                                ctypes.c_uint32, ctypes.c_int, _u32p]),
    "tfhe_lwe_decrypt": (ctypes.c_uint32, [ctypes.c_int, _i32p, _u32p, ctypes.c_uint32]),
    # circuit-bootstrapping level and AES
    "tfhe_cbs_ctx_create": (ctypes.c_int, [ctypes.POINTER(CbsParams), ctypes.c_uint32, ctypes.POINTER(_vp)]),
    "tfhe_cbs_ctx_destroy": (ctypes.c_int, [_vp]),
    "tfhe_cbs_keygen": (ctypes.c_int, [_vp, _i32p, ctypes.c_uint64]),
    "tfhe_cbs_get_rlwe_sk": (ctypes.c_int, [_vp, _i32p]),
    "tfhe_cbs_get_pfks_entry": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                               np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")]),
    "tfhe_gen_bsk64": (ctypes.c_int, [ctypes.POINTER(CbsParams), ctypes.c_uint64, _vp, _vp, _vp, _vp]),
    "tfhe_gen_pfkl40": (ctypes.c_int, [ctypes.POINTER(CbsParams), ctypes.c_uint64, _vp, _vp, _vp]),
    "tfhe_cbs_get_pfkl_entry": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                               np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")]),
    "tfhe_gen_pfks32": (ctypes.c_int, [ctypes.POINTER(CbsParams), ctypes.c_uint64, _vp, _vp, _vp]),
    "tfhe_gen_htss_keys": (ctypes.c_int, [ctypes.POINTER(CbsParams), ctypes.c_uint64, _vp, _vp, _vp]),
    "tfhe_cbs_set_keys": (ctypes.c_int, [_vp, _i32p, _vp, _vp, _vp]),
    "tfhe_cbs_batch": (ctypes.c_int, [_vp, _vp, ctypes.c_uint32, _vp, _vp]),
    "tfhe_cbs_pbs_batch": (ctypes.c_int, [_vp, _vp, ctypes.c_uint32, _vp, _vp]),
    "tfhe_cbs_pfks_batch": (ctypes.c_int, [_vp, ctypes.c_int, _vp, ctypes.c_uint32, ctypes.c_size_t, _vp, _vp]),
    "tfhe_cbs_batch_from_device_lwe": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_uint32, ctypes.c_uint32,
                                                      ctypes.POINTER(RgswBatchView), _vp]),
    "tfhe_cbs_write_keys": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(CbsParams), ctypes.c_uint64, _i32p, _vp,
                                           _vp, _vp]),
    "tfhe_cbs_save_keys": (ctypes.c_int, [_vp, ctypes.c_char_p]),
    "tfhe_cbs_load_keys": (ctypes.c_int, [_vp, ctypes.c_char_p]),
    "tfhe_cbs_key_file_info": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(CbsParams),
                                              ctypes.POINTER(ctypes.c_uint64)]),
    "tfhe_cbs_batch_to_device": (ctypes.c_int, [_vp, _u32p, ctypes.c_uint32, ctypes.POINTER(RgswBatchView), _vp]),
    "tfhe_cbs_bootstrap_many_lut": (ctypes.c_int, [_vp, _vp, ctypes.c_uint32, _vp, ctypes.c_uint32, _vp, _vp]),
    "tfhe_rgsw_to_fourier": (ctypes.c_int, [_vp, _vp, ctypes.c_uint32, _vp, _vp]),
    "tfhe_lut8_eval_batch": (ctypes.c_int, [_vp, _vp, ctypes.c_uint32, _vp, ctypes.c_size_t, ctypes.c_int, _vp, _vp]),
    "tfhe_lut8x32_eval_batch": (ctypes.c_int, [_vp, _vp, ctypes.c_uint32, _vp, _vp, _vp]),
    "tfhe_cbs_keyswitch_batch": (ctypes.c_int, [_vp, _vp, ctypes.c_uint32, _vp, _vp]),
    "tfhe_cbs_autotune": (ctypes.c_int, [_vp, ctypes.c_uint32, _i32p, _vp]),
    "tfhe_cbs_set_pbs_form": (ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.c_int]),
    "tfhe_cbs_timing_enable": (ctypes.c_int, [_vp, ctypes.c_int]),
    "tfhe_cbs_timing_read": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_uint32),
                                            ctypes.POINTER(ctypes.c_uint64)]),
    "tfhe_polymul2048_batch": (ctypes.c_int, [_vp, _vp, ctypes.c_uint32, _vp, _vp]),
    "tfhe_polymul2048_approx_batch": (ctypes.c_int, [_vp, _vp, ctypes.c_uint32, _vp, _vp]),
    "tfhe_polymul2048_split32_batch": (ctypes.c_int, [_vp, _vp, ctypes.c_uint32, ctypes.c_int, _vp, _vp]),
    "tfhe_aes_encrypt_key_bits": (None, [ctypes.c_int, _i32p, ctypes.c_uint64,
                                         np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS"),
                                         ctypes.c_int, _u32p]),
    "tfhe_aes128_encrypt_plain": (None, [np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")] * 3),
    "tfhe_aes_set_round_keys": (ctypes.c_int, [_vp, _u32p]),
    "tfhe_aes_round_batch": (ctypes.c_int, [_vp, ctypes.c_int, _vp, ctypes.c_uint32, _vp, _vp]),
    "tfhe_aes_ctr_keystream_batch": (ctypes.c_int, [_vp, np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS"),
                                                    ctypes.c_uint32, _vp, _vp]),
    "tfhe_aes_ctr_keystream_graph": (ctypes.c_int, [_vp, np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS"),
                                                    ctypes.c_uint32, _vp, _vp]),
    "tfhe_aes_ctr_keystream_plain": (ctypes.c_int, [_vp, np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS"),
                                                    ctypes.c_uint32, _vp, _vp]),
    "tfhe_lwe_xor_public": (ctypes.c_int, [_vp, _vp, ctypes.c_uint32, ctypes.c_int, _vp]),
    "tfhe_aes_set_round_form": (ctypes.c_int, [_vp, ctypes.c_int]),
    "tfhe_aes_encrypt_dec_key_bits": (None, [ctypes.c_int, _i32p, ctypes.c_uint64,
                                             np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS"),
                                             ctypes.c_int, _u32p]),
    "tfhe_aes_set_dec_round_keys": (ctypes.c_int, [_vp, _u32p]),
    "tfhe_aes_dec_round_batch": (ctypes.c_int, [_vp, ctypes.c_int, _vp, ctypes.c_uint32, _vp, _vp]),
    "tfhe_aes_decrypt_batch": (ctypes.c_int, [_vp, np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS"),
                                              ctypes.c_uint32, _vp, _vp]),
    "tfhe_aes_ctr_transcipher_batch": (ctypes.c_int, [_vp, ctypes.c_int,
                                                      np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS"),
                                                      np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS"),
                                                      ctypes.c_uint32, _vp, _vp]),
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


def _need(t, nbytes, what):
    """Raise unless t is a contiguous device tensor holding at least nbytes: the C ABI
    takes raw pointers, so an undersized buffer would be read or written past its end."""
    if t is None:
        raise TfheError(TFHE_E_PARAM, f"{what}: buffer is None")
    if not getattr(t, "is_cuda", False) or not t.is_contiguous():
        raise TfheError(TFHE_E_PARAM, f"{what}: expected a contiguous device tensor")
    have = t.numel() * t.element_size()
    if have < nbytes:
        raise TfheError(TFHE_E_SHAPE, f"{what}: {have} bytes, {nbytes} needed")


def _need_span(t, nbytes, what):
    """Raise unless t is a device tensor whose storage holds nbytes from t's first
    element: for outputs written at a stride (rows out_stride words apart), where t may
    be a strided view into a larger buffer."""
    if t is None:
        raise TfheError(TFHE_E_PARAM, f"{what}: buffer is None")
    if not getattr(t, "is_cuda", False):
        raise TfheError(TFHE_E_PARAM, f"{what}: expected a device tensor")
    st = t.untyped_storage()
    have = st.nbytes() - (t.data_ptr() - st.data_ptr())
    if have < nbytes:
        raise TfheError(TFHE_E_SHAPE, f"{what}: {have} bytes from the first element, {nbytes} needed")


def _blocks16(blocks, what):
    """Join 16-byte AES blocks into one host buffer, rejecting any other length."""
    for k, b in enumerate(blocks):
        if len(b) != 16:
            raise TfheError(TFHE_E_SHAPE, f"{what}: block {k} has {len(b)} bytes, not 16")
    return np.ascontiguousarray(np.frombuffer(b"".join(bytes(b) for b in blocks), np.uint8))


def _stream(s):
    return ctypes.c_void_p(0 if s is None else s)


def polymul1024_batch(d_x, d_y, d_out, stream=None):
    b = d_x.shape[0]
    for t, w in ((d_x, "x"), (d_y, "y"), (d_out, "out")):
        _need(t, b * 1024 * 4, f"polymul1024 {w}")
    check(lib().tfhe_polymul1024_batch(_ptr(d_x), _ptr(d_y), b, _ptr(d_out), _stream(stream)))


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
        row = 4 * (self.p.n + 1)
        for t, w in ((d_ct1, "ct1"), (d_ct2, "ct2"), (d_out, "out")):
            _need(t, b * row, f"eval_batch {w}")
        check(lib().tfhe_gate_eval_batch(self.h, gate, _ptr(d_ct1), _ptr(d_ct2), b, _ptr(d_out), _stream(stream)))

    def eval_batch_noalloc(self, gate, d_ct1, d_ct2, d_out, batch=None, stream=None):
        """tfhe_gate_eval_batch_noalloc: the same words as eval_batch, with the library's fixed
        device scratch (at most NOALLOC_MAX_GATES gates; the call allocates nothing)."""
        b = d_ct1.shape[0] if batch is None else batch
        row = 4 * (self.p.n + 1)
        for t, w in ((d_ct1, "ct1"), (d_ct2, "ct2"), (d_out, "out")):
            _need(t, b * row, f"eval_batch_noalloc {w}")
        check(lib().tfhe_gate_eval_batch_noalloc(self.h, gate, _ptr(d_ct1), _ptr(d_ct2), b, _ptr(d_out),
                                                 _stream(stream)))

    def eval_plain(self, gate, d_ct1, d_ct2, d_out, stream=None):
        row = 4 * (self.p.n + 1)
        for t, w in ((d_ct1, "ct1"), (d_ct2, "ct2"), (d_out, "out")):
            _need(t, row, f"eval_plain {w}")
        check(lib().tfhe_gate_eval_plain(self.h, gate, _ptr(d_ct1), _ptr(d_ct2), _ptr(d_out), _stream(stream)))

    def blind_rotate_batch(self, gate, d_ct1, d_ct2, d_lweN, stream=None):
        b = d_ct1.shape[0]
        _need(d_ct2, b * 4 * (self.p.n + 1), "blind_rotate ct2")
        _need(d_lweN, b * 4 * (self.p.N + 1), "blind_rotate out")
        check(lib().tfhe_gate_blind_rotate_batch(self.h, gate, _ptr(d_ct1), _ptr(d_ct2), b,
                                                 _ptr(d_lweN), _stream(stream)))

    def external_product(self, key_index, d_acc, d_out, stream=None):
        _need(d_acc, 8 * self.p.N, "external_product acc")
        _need(d_out, 8 * self.p.N, "external_product out")
        check(lib().tfhe_gate_external_product(self.h, key_index, _ptr(d_acc), _ptr(d_out), _stream(stream)))

    def keyswitch_batch(self, d_lweN, d_out, stream=None):
        _need(d_out, d_lweN.shape[0] * 4 * (self.p.n + 1), "keyswitch out")
        check(lib().tfhe_gate_keyswitch_batch(self.h, _ptr(d_lweN), d_lweN.shape[0], _ptr(d_out), _stream(stream)))

    # ---- programmable bootstrap and large-precision functions ----
    def _func_io(self, d_ct, d_out, batch, what, digits=1):
        b = d_ct.shape[0] if batch is None else batch
        row = 4 * (self.p.n + 1)
        _need(d_ct, b * row, f"{what} ct")
        _need(d_out, digits * b * row, f"{what} out")
        return b

    def bootstrap_func(self, d_ct, ctmod, lut, fmod, d_out, batch=None, stream=None):
        """BootstrapFunc on a batch over Z_ctmod with a table over Z_fmod (lut: ctmod entries)."""
        b = self._func_io(d_ct, d_out, batch, "bootstrap_func")
        lut = np.ascontiguousarray(lut, np.uint32)
        if lut.size != int(ctmod):
            raise ValueError("bootstrap_func: the table needs ctmod entries")
        check(lib().tfhe_gate_bootstrap_func_batch(self.h, _ptr(d_ct), b, int(ctmod), lut, int(fmod), _ptr(d_out),
                                                   _stream(stream)))

    def eval_func(self, d_ct, q, lut, d_out, batch=None, stream=None):
        """EvalFunc on a batch over Z_q with a table of q entries over Z_q."""
        b = self._func_io(d_ct, d_out, batch, "eval_func")
        lut = np.ascontiguousarray(lut, np.uint32)
        if lut.size != int(q):
            raise ValueError("eval_func: the table needs q entries")
        check(lib().tfhe_gate_eval_func_batch(self.h, _ptr(d_ct), b, int(q), lut, _ptr(d_out), _stream(stream)))

    def eval_floor(self, d_ct, mod, d_out, roundbits=0, batch=None, stream=None):
        b = self._func_io(d_ct, d_out, batch, "eval_floor")
        check(lib().tfhe_gate_eval_floor_batch(self.h, _ptr(d_ct), b, int(mod), roundbits, _ptr(d_out),
                                               _stream(stream)))

    def eval_sign(self, d_ct, mod, d_out, scheme_switch=False, batch=None, stream=None):
        b = self._func_io(d_ct, d_out, batch, "eval_sign")
        check(lib().tfhe_gate_eval_sign_batch(self.h, _ptr(d_ct), b, int(mod), int(bool(scheme_switch)),
                                              _ptr(d_out), _stream(stream)))

    def eval_decomp(self, d_ct, mod, d_out, max_digits, batch=None, stream=None):
        """EvalDecomp into d_out [max_digits][batch][n+1]; returns the digit moduli (one per digit)."""
        b = self._func_io(d_ct, d_out, batch, "eval_decomp", digits=max_digits)
        cnt = ctypes.c_uint32(0)
        mods = np.zeros(max(int(max_digits), 1), np.uint64)
        check(lib().tfhe_gate_eval_decomp_batch(self.h, _ptr(d_ct), b, int(mod), max_digits, _ptr(d_out),
                                                ctypes.byref(cnt), mods, _stream(stream)))
        return [int(m) for m in mods[:cnt.value]]


def func_property(lut, q):
    """0 negacyclic, 1 periodic, 2 arbitrary (tfhe_func_property)."""
    return lib().tfhe_func_property(np.ascontiguousarray(lut, np.uint32), int(q))


FUNC_Q = 1024
FUNC_BETA = 64


def polymul2048_batch(d_x, d_y, d_out, stream=None, approx=False):
    """Negacyclic products x * y (N = 2048, torus-64 y): exact, or in the PBS's
    one-product form (approx=True)."""
    b = d_x.shape[0]
    _need(d_x, b * 2048 * 4, "polymul2048 x")
    _need(d_y, b * 2048 * 8, "polymul2048 y")
    _need(d_out, b * 2048 * 8, "polymul2048 out")
    fn = lib().tfhe_polymul2048_approx_batch if approx else lib().tfhe_polymul2048_batch
    check(fn(_ptr(d_x), _ptr(d_y), b, _ptr(d_out), _stream(stream)))


def write_keys(path, params, seed, rlwe_sk, bsk, pfks, ksk2):
    """A key file from keys made elsewhere (tfhe_cbs_write_keys); bsk / pfks / ksk2 are
    numpy arrays (host) or device tensors."""
    def p_(a):
        return ctypes.c_void_p(a.ctypes.data) if isinstance(a, np.ndarray) else _ptr(a)
    check(lib().tfhe_cbs_write_keys(os.fsencode(path), ctypes.byref(params), seed,
                                    np.ascontiguousarray(rlwe_sk, np.int32), p_(bsk), p_(pfks), p_(ksk2)))


def key_file_info(path):
    """(CbsParams, seed) of a key file (tfhe_cbs_key_file_info)."""
    prm, seed = CbsParams(), ctypes.c_uint64()
    check(lib().tfhe_cbs_key_file_info(os.fsencode(path), ctypes.byref(prm), ctypes.byref(seed)))
    return prm, seed.value


SPLIT32_FORMS = {"fp32_3x11": 0, "fp32_exact_top": 1, "fp64": 2}


def polymul2048_split32_batch(d_x, d_y, d_out, form, stream=None):
    """Negacyclic products x * y mod 2^32 (N = 2048, torus-32 y) in the split-precision
    FP32 forms ("fp32_3x11": limbs 11/11/10, "fp32_exact_top": limbs 11/6/5/5/5) or one
    FP64 product ("fp64"); tfhe_polymul2048_split32_batch."""
    b = d_x.shape[0]
    _need(d_x, b * 2048 * 4, "split32 x")
    _need(d_y, b * 2048 * 4, "split32 y")
    _need(d_out, b * 2048 * 4, "split32 out")
    check(lib().tfhe_polymul2048_split32_batch(_ptr(d_x), _ptr(d_y), b, SPLIT32_FORMS[form], _ptr(d_out),
                                               _stream(stream)))


def aes_encrypt_dec_key_bits(sk, seed, key, nshift=15):
    """The equivalent inverse cipher's round-key bits encrypted under sk, [11][128][n+1]."""
    n = len(sk)
    out = np.zeros(11 * 128 * (n + 1), np.uint32)
    lib().tfhe_aes_encrypt_dec_key_bits(n, np.ascontiguousarray(sk, np.int32), seed,
                                        np.frombuffer(bytes(key), np.uint8).copy(), nshift, out)
    return out.reshape(11, 128, n + 1)


def aes_encrypt_key_bits(sk, seed, key, nshift=15):
    n = len(sk)
    out = np.zeros(11 * 128 * (n + 1), np.uint32)
    lib().tfhe_aes_encrypt_key_bits(n, np.ascontiguousarray(sk, np.int32), seed,
                                    np.frombuffer(bytes(key), np.uint8).copy(), nshift, out)
    return out.reshape(11, 128, n + 1)


def aes128_encrypt_plain(key, block):
    out = np.zeros(16, np.uint8)
    lib().tfhe_aes128_encrypt_plain(np.frombuffer(bytes(key), np.uint8).copy(),
                                    np.frombuffer(bytes(block), np.uint8).copy(), out)
    return bytes(out)


class CbsContext:
    """Circuit-bootstrapping + AES transciphering context (GPUCirBTSContext counterpart)."""

    def __init__(self, max_blocks=1, params=None):
        self.p = params if params is not None else cbs_params()
        h = ctypes.c_void_p()
        check(lib().tfhe_cbs_ctx_create(ctypes.byref(self.p), max_blocks, ctypes.byref(h)))
        self.h = h
        self.max_blocks = max_blocks

    def close(self):
        if self.h:
            lib().tfhe_cbs_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def keygen(self, lwe_sk, seed):
        check(lib().tfhe_cbs_keygen(self.h, np.ascontiguousarray(lwe_sk, np.int32), seed))

    def rlwe_sk(self):
        K = np.zeros(self.p.N, np.int32)
        check(lib().tfhe_cbs_get_rlwe_sk(self.h, K))
        return K

    def set_keys(self, rlwe_sk, bsk, pfks, ksk2):
        """Upload keys made elsewhere (tfhe_cbs_set_keys); bsk / pfks / ksk2 are numpy
        arrays (host) or device tensors."""
        def p_(a):
            return a.ctypes.data if isinstance(a, np.ndarray) else _ptr(a)
        check(lib().tfhe_cbs_set_keys(self.h, np.ascontiguousarray(rlwe_sk, np.int32), p_(bsk), p_(pfks), p_(ksk2)))

    def save_keys(self, path):
        """Write the context's keys (made by keygen) to a key file (tfhe_cbs_save_keys)."""
        check(lib().tfhe_cbs_save_keys(self.h, os.fsencode(path)))

    def load_keys(self, path):
        """Check a key file against the context's parameter set and upload it (tfhe_cbs_load_keys)."""
        check(lib().tfhe_cbs_load_keys(self.h, os.fsencode(path)))

    def pfkl_entry(self, u, i, t):
        """Linear-key entry (u, i, t): 2N words mod 2^40."""
        out = np.zeros(2 * self.p.N, np.uint64)
        check(lib().tfhe_cbs_get_pfkl_entry(self.h, u, i, t, out))
        return out

    def pfks_entry(self, u, i, t, v):
        out = np.zeros(2 * self.p.N, np.uint32)
        check(lib().tfhe_cbs_get_pfks_entry(self.h, u, i, t, v, out))
        return out

    def pbs_batch(self, d_lwe, d_out, batch=None, stream=None):
        b = d_lwe.shape[0] if batch is None else batch
        _need(d_lwe, b * 4 * (self.p.n + 1), "pbs in")
        _need(d_out, b * 2 * 8 * (self.p.N + 1), "pbs out")
        check(lib().tfhe_cbs_pbs_batch(self.h, _ptr(d_lwe), b, _ptr(d_out), _stream(stream)))

    def pfks_batch(self, u, d_lwe64, count, out_stride, d_out, stream=None):
        if count:
            _need(d_lwe64, count * 8 * (self.p.N + 1), "pfks in")
            _need_span(d_out, ((count - 1) * out_stride + 2 * self.p.N) * 4, "pfks out")
        check(lib().tfhe_cbs_pfks_batch(self.h, u, _ptr(d_lwe64), count, out_stride, _ptr(d_out), _stream(stream)))

    def cbs_batch(self, d_lwe, d_rgsw, batch=None, stream=None):
        b = d_lwe.shape[0] if batch is None else batch
        _need(d_lwe, b * 4 * (self.p.n + 1), "cbs in")
        _need(d_rgsw, b * 4 * 2 * self.p.N * 4, "cbs out")
        check(lib().tfhe_cbs_batch(self.h, _ptr(d_lwe), b, _ptr(d_rgsw), _stream(stream)))

    def cbs_batch_from_device_lwe(self, d_a, d_b, batch, a_stride, stream=None):
        """Circuit bootstraps of LWEs given as u64 masks (rows a_stride apart) and u64
        bodies; returns the RgswBatchView of the outputs held in the context."""
        v = RgswBatchView()
        check(lib().tfhe_cbs_batch_from_device_lwe(self.h, _ptr(d_a) if d_a is not None else None,
                                                   _ptr(d_b) if d_b is not None else None, batch, a_stride,
                                                   ctypes.byref(v), _stream(stream)))
        return v

    def cbs_batch_to_device(self, h_lwe, stream=None):
        """Circuit bootstraps of host LWE ciphertexts [batch][n+1] (uint32); returns the
        RgswBatchView of the outputs held in the context (tfhe_cbs_batch_to_device)."""
        h = np.ascontiguousarray(h_lwe, np.uint32).reshape(-1, self.p.n + 1)
        v = RgswBatchView()
        check(lib().tfhe_cbs_batch_to_device(self.h, h.reshape(-1), h.shape[0], ctypes.byref(v), _stream(stream)))
        return v

    def bootstrap_many_lut(self, d_lwe, d_lut, bitwidth, d_out, batch=None, stream=None):
        """gpu_BootstrapManyLUT over a batch: d_lwe [batch][n+1] uint32, d_lut [N] (torus-64),
        d_out [batch][2][N] (torus-64 RLWE, mask then body)."""
        b = d_lwe.shape[0] if batch is None else batch
        _need(d_lwe, b * 4 * (self.p.n + 1), "many_lut in")
        _need(d_lut, 8 * self.p.N, "many_lut lut")
        _need(d_out, b * 2 * 8 * self.p.N, "many_lut out")
        check(lib().tfhe_cbs_bootstrap_many_lut(self.h, _ptr(d_lwe), b, _ptr(d_lut), bitwidth, _ptr(d_out),
                                                _stream(stream)))

    def rgsw_to_fourier(self, d_rgsw, count, d_rgsw_f, stream=None):
        _need(d_rgsw, count * 8 * self.p.N * 4, "rgsw_to_fourier in")
        _need(d_rgsw_f, count * 8 * 1024 * 16, "rgsw_to_fourier out")
        check(lib().tfhe_rgsw_to_fourier(self.h, _ptr(d_rgsw), count, _ptr(d_rgsw_f), _stream(stream)))

    def lut8(self, d_rgsw_f, nlut, d_table, table_stride, out_bits, d_out32, stream=None):
        if nlut:
            _need(d_rgsw_f, nlut * 8 * 8 * 1024 * 16, "lut8 rgsw")
            _need(d_table, ((nlut - 1) * table_stride + self.p.N) * 4, "lut8 table")
            _need(d_out32, nlut * out_bits * (self.p.N + 1) * 4, "lut8 out")
        check(lib().tfhe_lut8_eval_batch(self.h, _ptr(d_rgsw_f), nlut, _ptr(d_table), table_stride, out_bits,
                                         _ptr(d_out32), _stream(stream)))

    def lut8x32(self, d_rgsw_f, nbytes, d_tables, d_out32, stream=None):
        """8 -> 32-bit LUT (T-table lookup) of nbytes bytes; d_tables [4][N] (tfhe_lut8x32_eval_batch)."""
        if nbytes:
            _need(d_rgsw_f, nbytes * 8 * 8 * 1024 * 16, "lut8x32 rgsw")
            _need(d_tables, 4 * self.p.N * 4, "lut8x32 tables")
            _need(d_out32, nbytes * 32 * (self.p.N + 1) * 4, "lut8x32 out")
        check(lib().tfhe_lut8x32_eval_batch(self.h, _ptr(d_rgsw_f), nbytes, _ptr(d_tables), _ptr(d_out32),
                                            _stream(stream)))

    def keyswitch_batch(self, d_lwe32, count, d_out, stream=None):
        _need(d_lwe32, count * (self.p.N + 1) * 4, "keyswitch in")
        _need(d_out, count * (self.p.n + 1) * 4, "keyswitch out")
        check(lib().tfhe_cbs_keyswitch_batch(self.h, _ptr(d_lwe32), count, _ptr(d_out), _stream(stream)))

    def autotune(self, blocks, stream=None):
        """Time and keep the fastest launch shapes for AES rounds of `blocks` blocks
        (tfhe_cbs_autotune); returns the choices as a dict."""
        ch = np.full(5, -1, np.int32)
        check(lib().tfhe_cbs_autotune(self.h, blocks, ch, _stream(stream)))
        return {"pbs_form": [int(ch[0]), int(ch[1])], "pfks_slices": [int(ch[2]), int(ch[3])],
                "ks_slices": int(ch[4])}

    def set_pbs_form(self, inputs, form):
        """Pin the PBS launch form for launches of `inputs` ciphertexts (-1: back to the
        size model); tfhe_cbs_set_pbs_form."""
        check(lib().tfhe_cbs_set_pbs_form(self.h, int(inputs), int(form)))

    def timing_enable(self, on=True):
        """Bracket every PBS launch with HIP events from now on (tfhe_cbs_timing_enable)."""
        check(lib().tfhe_cbs_timing_enable(self.h, int(bool(on))))

    def timing_read(self):
        """(summed PBS launch ms, launches, PBS inputs) since enabling or the last read."""
        ms, n, inp = ctypes.c_double(), ctypes.c_uint32(), ctypes.c_uint64()
        check(lib().tfhe_cbs_timing_read(self.h, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(inp)))
        return ms.value, n.value, inp.value

    def set_round_form(self, form):
        """0 = S-box LUT + LWE linear layers, 1 = T-table rounds (tfhe_aes_set_round_form)."""
        check(lib().tfhe_aes_set_round_form(self.h, int(form)))

    def set_dec_round_keys(self, dk_bits):
        check(lib().tfhe_aes_set_dec_round_keys(self.h, np.ascontiguousarray(dk_bits, np.uint32).reshape(-1)))

    def aes_dec_round(self, rnd, d_in, blocks, d_out, stream=None):
        row = 128 * (self.p.n + 1) * 4
        _need(d_in, blocks * row, "aes_dec_round in")
        _need(d_out, blocks * row, "aes_dec_round out")
        check(lib().tfhe_aes_dec_round_batch(self.h, rnd, _ptr(d_in), blocks, _ptr(d_out), _stream(stream)))

    def aes_decrypt(self, cipher_blocks, d_out, stream=None):
        """LWE bits [blocks][128][n+1] of AES-128^-1 of public ciphertext blocks (tfhe_aes_decrypt_batch)."""
        ct = _blocks16(cipher_blocks, "aes_decrypt ciphertext")
        _need(d_out, len(cipher_blocks) * 128 * (self.p.n + 1) * 4, "aes_decrypt out")
        check(lib().tfhe_aes_decrypt_batch(self.h, ct, len(cipher_blocks), _ptr(d_out), _stream(stream)))

    def set_round_keys(self, rk_bits):
        check(lib().tfhe_aes_set_round_keys(self.h, np.ascontiguousarray(rk_bits, np.uint32).reshape(-1)))

    def aes_round(self, rnd, d_in, blocks, d_out, stream=None):
        row = 128 * (self.p.n + 1) * 4
        _need(d_in, blocks * row, "aes_round in")
        _need(d_out, blocks * row, "aes_round out")
        check(lib().tfhe_aes_round_batch(self.h, rnd, _ptr(d_in), blocks, _ptr(d_out), _stream(stream)))

    def aes_ctr_keystream(self, ctr_blocks, d_out, stream=None, graph=False, plain=False):
        """CTR keystream bits of AES_k(ctr) for each 16-byte counter block; graph=True
        replays a captured hipGraph (tfhe_aes_ctr_keystream_graph), plain=True launches
        every bootstrap on its own (tfhe_aes_ctr_keystream_plain)."""
        ctr = _blocks16(ctr_blocks, "aes_ctr_keystream counters")
        _need(d_out, len(ctr_blocks) * 128 * (self.p.n + 1) * 4, "aes_ctr_keystream out")
        fn = (lib().tfhe_aes_ctr_keystream_graph if graph else
              lib().tfhe_aes_ctr_keystream_plain if plain else lib().tfhe_aes_ctr_keystream_batch)
        check(fn(self.h, ctr, len(ctr_blocks), _ptr(d_out), _stream(stream)))

    def aes_ctr_transcipher(self, ctr_blocks, cipher_blocks, d_out, stream=None, mode="batch"):
        """LWE bits of the plaintext of public AES-CTR ciphertext blocks: the keystream of
        the counter blocks plus the ciphertext bits (tfhe_aes_ctr_transcipher_batch);
        mode "batch", "graph" or "plain"."""
        if len(cipher_blocks) != len(ctr_blocks):
            raise ValueError("one ciphertext block per counter block")
        ctr = _blocks16(ctr_blocks, "aes_ctr_transcipher counters")
        ct = _blocks16(cipher_blocks, "aes_ctr_transcipher ciphertext")
        _need(d_out, len(ctr_blocks) * 128 * (self.p.n + 1) * 4, "aes_ctr_transcipher out")
        check(lib().tfhe_aes_ctr_transcipher_batch(self.h, AES_MODES[mode], ctr, ct, len(ctr_blocks), _ptr(d_out),
                                                   _stream(stream)))


AES_MODES = {"batch": 0, "graph": 1, "plain": 2}
NOALLOC_MAX_GATES = 1024  # TFHE_NOALLOC_MAX_GATES

