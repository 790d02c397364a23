import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(REPO, "tests", "golden")
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: longer CPU-side checks")


def golden(name):
    """Load a committed fixture (plain arrays only, no pickles)."""
    path = os.path.join(GOLDEN, name)
    if name.endswith(".npy"):
        return np.load(path, allow_pickle=False)
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def ref_lib():
    """The reference build (oracle/_ref), present only where the reference tree was."""
    from oracle import oracle
    import ctypes
    from gausslectronitegrate_amd import _abi
    if not os.path.exists(oracle.REF_LIB_PATH):
        pytest.skip("reference build oracle/_ref not present")
    return _abi.bind(ctypes.CDLL(oracle.REF_LIB_PATH))
