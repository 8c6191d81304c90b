"""Split-precision FP32 products (the configs' split-FP32 external-product arithmetic)
against the oracle's exact schoolbook product mod 2^32 (orc_nc_mul_exact32, pinned
by the oracle's own tests): digit polynomials in [-256, 256] times uniform torus-32
polynomials, N = 2048."""
import numpy as np
import pytest

import oracle as O
from conftest import from_dev, to_dev

pytestmark = pytest.mark.gpu


def _errors(got, want):
    """signed coefficient errors mod 2^32"""
    d = (got.astype(np.int64) - want.astype(np.int64)) % 2 ** 32
    return np.where(d >= 2 ** 31, d - 2 ** 32, d)


@pytest.fixture(scope="module")
def products(gpu):
    rng = np.random.default_rng(41)
    batch = 48
    x = rng.integers(-256, 257, (batch, 2048)).astype(np.int32)
    y = rng.integers(0, 2 ** 32, (batch, 2048), dtype=np.uint64).astype(np.uint32)
    want = np.stack([O.nc_mul_exact32(x[k], y[k]) for k in range(batch)])
    return x, y, want


@pytest.mark.parametrize("form", ["fp64", "fp32_exact_top", "fp32_3x11"])
def test_split32_product_error(products, form):
    """fp64 (one FP64 product): bit-exact.  fp32_exact_top (limbs 11/6/5/5/5): every
    limb above the lowest is exact on these inputs, so no error reaches weight 2^11;
    asserted max |error| < 2^9 of 2^32.  fp32_3x11 (limbs 11/11/10): the top limb's
    products (rms about 2^21, against FP32's 24-bit significand) round wrongly now and
    then, each miss worth 2^22; asserted: max |error| < 2^28 (at most a few misses of
    the top limb per coefficient) and the error printed."""
    import torch
    import tfhe_amd
    x, y, want = products
    d_out = torch.empty(y.shape, dtype=torch.int32, device="cuda")
    tfhe_amd.polymul2048_split32_batch(to_dev(x), to_dev(y), d_out, form)
    torch.cuda.synchronize()
    err = _errors(from_dev(d_out), want)
    worst = int(np.abs(err).max())
    frac_wrong = float(np.mean(err != 0))
    print(f"split32 {form}: max |coefficient error| {worst} = 2^{np.log2(max(worst, 1)):.2f} of 2^32, "
          f"rms 2^{np.log2(max(np.sqrt(np.mean(err.astype(np.float64) ** 2)), 1e-9)):.2f}, "
          f"{100 * frac_wrong:.3f}% of coefficients wrong")
    if form == "fp64":
        assert worst == 0
    elif form == "fp32_exact_top":
        assert worst < 2 ** 9
    else:
        assert worst < 2 ** 28


def test_split32_empty_and_bad_form(gpu):
    import ctypes
    import torch
    import tfhe_amd
    L = tfhe_amd.lib()
    buf = torch.zeros(2048, dtype=torch.int32, device="cuda")
    p = ctypes.c_void_p(buf.data_ptr())
    assert L.tfhe_polymul2048_split32_batch(p, p, 0, 0, p, None) == tfhe_amd.TFHE_OK
    assert L.tfhe_polymul2048_split32_batch(p, p, 1, 3, p, None) == tfhe_amd.TFHE_E_PARAM
