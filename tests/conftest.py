import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built libwedg.so")


def rel(a, b):
    """max |a - b| / max |a| (the reference's rhs_rel, test/runtests.jl:228-229)."""
    import numpy as np
    a = np.asarray(a); b = np.asarray(b)
    s = np.abs(a).max()
    return float(np.abs(a - b).max() / s) if s > 0 else float(np.abs(a - b).max())


@pytest.fixture(scope="session")
def oracle_built():
    from oracle import batched
    batched.build()
    return batched
