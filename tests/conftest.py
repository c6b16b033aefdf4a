import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "flash-sd-kde_amd")
for p in (PKG, ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP) GPU")
    config.addinivalue_line("markers", "small: small parity case (reference test shapes)")
    config.addinivalue_line("markers", "large: larger parity case")


@pytest.fixture(scope="session")
def golden():
    import numpy as np

    path = os.path.join(ROOT, "tests", "golden", "reference_vectors.npz")
    with np.load(path, allow_pickle=False) as z:
        data = {k: z[k] for k in z.files}

    class G:
        def __call__(self, case, key):
            v = data[f"{case}/{key}"]
            return v.item() if v.ndim == 0 else v

    return G()


@pytest.fixture(scope="session")
def gpu():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    from flash_sd_kde import _native

    _native.load_library()
    return torch.device("cuda:0")
