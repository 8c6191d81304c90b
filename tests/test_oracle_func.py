"""CPU pins of the oracle's programmable bootstrap and large-precision functions
(oracle/func_oracle.c; reference bitwise/src/binfhe-base-scheme.cu:255-698): the function
classes, BootstrapFunc against its closed form (the test vector indexed by the exact phase),
EvalFunc on negacyclic, periodic and arbitrary functions of Z_8, EvalFloor, EvalSign and
EvalDecomp against the plaintext recursion they restate."""
import numpy as np
import pytest

import oracle as O

Q, BETA = O.FUNC_Q, O.FUNC_BETA
P = Q // (2 * BETA)  # plaintext space of EvalFunc


@pytest.fixture(scope="module")
def gk():
    return O.GateKeys(31)


def _centered(v, mod):
    v %= mod
    return v - mod if v >= mod // 2 else v


def _lut(F):
    """LUT over Z_q of a function F of Z_P: x -> F(floor(x P / q)) q / P (the floor the beta offset makes exact)."""
    return np.array([(F[(x * P) // Q] % P) * (Q // P) for x in range(Q)], np.uint32)


def test_function_classes():
    q = 16
    neg = [(3 * x + 1) % q for x in range(q // 2)]
    neg = neg + [(-v) % q for v in neg]
    per = [(5 * x) % q for x in range(q // 2)] * 2
    arb = [(3 * x + 1) % q for x in range(q)]
    assert O.func_property(neg, q) == 0
    assert O.func_property(per, q) == 1
    assert O.func_property(arb, q) == 2
    assert O.func_property([0] * q, q) == 0  # both properties: negacyclic is checked first


@pytest.mark.parametrize("ctmod,fmod", [(1024, 1024), (2048, 2 ** 16), (256, 2 ** 32)])
def test_bootstrap_func_closed_form(gk, ctmod, fmod):
    """With u = <a, s> mod ctmod (the rotation, exact for ctmod | 2N) and the phase p = b - u,
    the output encrypts f(p) when u < ctmod / 2 and -f(p + ctmod / 2) otherwise (coefficient 0
    of the test vector rotated by X^-u, negated past N), for a random table f and random
    inputs."""
    rng = np.random.default_rng(ctmod + fmod % 1000)
    f = rng.integers(0, min(fmod, 2 ** 32), ctmod, dtype=np.uint64).astype(np.uint32)
    tol = 2 ** 27 if fmod == 2 ** 32 else max(64, fmod >> 5)  # about 8 sigma of the key-switch and rounding noise
    for k in range(6):
        ct = gk.encrypt_mod(int(rng.integers(0, ctmod)), ctmod, 500 + k)
        ph = gk.phase_mod(ct, ctmod)
        u = int(np.dot(ct[:-1].astype(np.uint64), gk.s.astype(np.uint64)) % ctmod)
        want = int(f[ph]) if u < ctmod // 2 else (-int(f[(ph + ctmod // 2) % ctmod])) % fmod
        out = gk.bootstrap_func(ct, ctmod, f, fmod)
        assert abs(_centered(gk.phase_mod(out, fmod) - want, fmod)) < tol


@pytest.mark.parametrize("kind", ["negacyclic", "periodic", "arbitrary"])
def test_eval_func_plaintext(gk, kind):
    """EvalFunc of F: Z_8 -> Z_8 on m * q / 8 decrypts to F(m) for every m."""
    F = {"negacyclic": [1, 2, 3, 0, 7, 6, 5, 0], "periodic": [3, 1, 4, 1, 3, 1, 4, 1],
         "arbitrary": [(3 * m + 1) % 8 for m in range(8)]}[kind]
    lut = _lut(F)
    assert O.func_property(lut, Q) == {"negacyclic": 0, "periodic": 1, "arbitrary": 2}[kind]
    for m in range(P):
        ct = gk.encrypt_mod(m * (Q // P), Q, 600 + m)
        out = gk.eval_func(ct, Q, lut)
        ph = gk.phase_mod(out, Q)
        assert ((ph + Q // (2 * P)) // (Q // P)) % P == F[m], (kind, m, ph)


def _floor_ref(x, mod, q=Q, beta=BETA):
    return (((x + beta) % mod) // q) * q


def test_eval_floor(gk):
    """EvalFloor over Z_2^16 decrypts to floor((phase + beta) / q) q within the bootstrap noise
    at that modulus; with roundbits = 1 the floor is to multiples of beta 2^2."""
    mod = 2 ** 16
    rng = np.random.default_rng(7)
    for k in range(4):
        ct = gk.encrypt_mod(int(rng.integers(0, mod)), mod, 700 + k)
        ph = gk.phase_mod(ct, mod)
        out = gk.eval_floor(ct, mod)
        assert abs(_centered(gk.phase_mod(out, mod) - _floor_ref(ph, mod), mod)) < 2048
        out = gk.eval_floor(ct, mod, roundbits=1)
        assert abs(_centered(gk.phase_mod(out, mod) - _floor_ref(ph, mod, q=BETA * 4), mod)) < 2048


@pytest.mark.parametrize("mod", [2 ** 16, 2 ** 22])
def test_eval_sign(gk, mod):
    """EvalSign decrypts to 0 (phase in the lower half of Z_mod) or q / 2 (upper half); the
    scheme-switching form to -q/4 or +q/4 before the shift."""
    rng = np.random.default_rng(mod % 97)
    for k in range(4):
        half = int(rng.integers(0, 2))
        v = half * (mod // 2) + int(rng.integers(mod // 16, mod // 2 - mod // 16))
        ct = gk.encrypt_mod(v, mod, 800 + k)
        ph = gk.phase_mod(ct, mod)
        assert (ph >= mod // 2) == bool(half)
        out = gk.eval_sign(ct, mod)
        assert abs(_centered(gk.phase_mod(out, Q) - half * (Q // 2), Q)) < BETA
        out = gk.eval_sign(ct, mod, scheme_switch=True)
        assert abs(_centered(gk.phase_mod(out, Q) - (Q // 4 if half else Q - Q // 4), Q)) < BETA


def test_eval_decomp(gk):
    """EvalDecomp over Z_2^13: two digits, at the moduli 2^13 and 2^10 (the floor keeps
    q / (2 beta) = 8 more levels).  Digit 0 is the input reduced mod q, word for word; and since
    floor(x) = x + beta - ((x + beta) mod q) and x_1 = floor(x_0) / 8, the digits give the input
    phase back through x_0 = 8 x_1 + ((d_0 + beta) mod q) - beta within the bootstrap noise.
    The inputs keep (x + beta) mod q away from the floor's branch points 0 and q / 2 by more
    than 8 sigma of the bootstrap output noise at 2^13 (2^-8.1 of the modulus, from the blind
    rotation and the key switch): nearer, the floor moves by one level (q), which the gate parameters' key-switch
    noise allows at large moduli."""
    mod = 2 ** 13
    rng = np.random.default_rng(11)
    for k in range(4):
        lo = int(rng.integers(300, 450)) + (Q // 2 if k % 2 else 0)  # (x + beta) mod q in [364, 514) or [876, 1026)
        v = int(rng.integers(0, mod // Q)) * Q + lo - BETA
        ct = gk.encrypt_mod(v, mod, 900 + k)
        digits = gk.eval_decomp(ct, mod)
        assert len(digits) == 2
        assert np.array_equal(digits[0], ct % Q)
        m1 = mod // Q * 2 * BETA
        x = gk.phase_mod(digits[1], m1)
        x = (8 * x + (gk.phase_mod(digits[0], Q) + BETA) % Q - BETA) % mod
        assert abs(_centered(x - gk.phase_mod(ct, mod), mod)) < 2 ** 8
