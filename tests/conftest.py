import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); run with -m gpu")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:  # noqa: BLE001
        return False


@pytest.fixture(scope="session")
def ap():
    """The product package on the GPU; fails (does not skip) if the native
    library is missing, so a GPU run can never pass on a fallback."""
    if not gpu_available():
        pytest.skip("no HIP device")
    import anionpy_amd
    anionpy_amd._lib.load()
    # IEEE flag tracking off by default, so the tests exercise the plain
    # kernels; tests of the flag variants and of np.errstate turn it on
    anionpy_amd.seterr(all="ignore")
    return anionpy_amd
