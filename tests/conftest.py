import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "gpu-circuitbootstrapping-aestransciphering_amd")
for p in (os.path.join(ROOT, "oracle"), PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def to_dev(a):
    """numpy uint32/int32 array -> int32 CUDA tensor holding the same bits."""
    import torch
    a = np.ascontiguousarray(a)
    return torch.from_numpy(a.view(np.int32).copy()).to("cuda")


def from_dev(t, dtype=np.uint32):
    return t.cpu().numpy().view(dtype)


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import tfhe_amd
    tfhe_amd.lib()  # raises if the native library is missing
    return torch.device("cuda:0")
