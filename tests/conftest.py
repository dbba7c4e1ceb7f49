import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP tracking kernels)")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import ptref
    if not (os.path.exists(os.path.join(ROOT, "oracle", "libptref_fp64.so"))
            and os.path.exists(os.path.join(ROOT, "oracle", "libptref_refsem.so"))):
        ptref.build()
    return ptref


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from pentrack_amd import _native
    _native.lib()
    return torch.device("cuda:0")
