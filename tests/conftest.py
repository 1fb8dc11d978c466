import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def oracle():
    from oracle.loader import load_oracle
    return load_oracle()


def golden(name):
    """A golden fixture (plain arrays only; nothing is unpickled)."""
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def osc_for(params_row, nutype):
    """oracle.loader.Osc for a (th12, th13, th23, dcp, dm12, dm23) row."""
    from oracle.loader import Osc
    t12, t13, t23, d, m12, m23 = [float(x) for x in params_row]
    return Osc(t12, t13, t23, d, m12, m23, -1 if nutype else 1)
