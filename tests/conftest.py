import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and libff_amd.so")
    # autotune results of this session go to a file of its own, so that no
    # result of an earlier run is taken from the user's tuning cache
    if "FFK_TUNE_CACHE" not in os.environ:
        import tempfile

        os.environ["FFK_TUNE_CACHE"] = os.path.join(tempfile.mkdtemp(prefix="ffk_tune_"), "autotune.txt")


def golden(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def bits(a):
    """Bit patterns of a float32 array, for exact comparisons (NaNs included)."""
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O

    if not os.path.exists(O.ORACLE_SO):
        O.build()
    O.lib()
    return O


@pytest.fixture(scope="session")
def gpu():
    """The product library on cuda:0 (GPU tests only)."""
    import torch

    from form_factors_amd import ffk

    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a HIP device")
    ffk.lib()
    torch.cuda.set_device(0)
    return ffk
