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


def conservation_all_jobs(b, n_jobs, tol):
    """Every point of every job of a batch, one job at a time: finite, within
    [0, 1] up to tol, and both sums of its 3 x 3 matrix P(a -> b) (over the
    final and over the initial flavour) equal to 1 within tol.  Returns the
    largest deviation of a sum."""
    worst = 0.0
    for j in range(n_jobs):
        got = b.result(j)
        assert np.isfinite(got).all(), f"job {j}: non-finite probability"
        assert got.min() > -tol and got.max() < 1 + tol, f"job {j}: out of [0, 1]"
        P = got.reshape(3, 3, *got.shape[1:])
        worst = max(worst, float(np.abs(P.sum(0) - 1).max()),
                    float(np.abs(P.sum(1) - 1).max()))
    assert worst < tol, f"probability sums off by {worst:.3e}"
    return worst


def osc_for(params_row, nutype):
    """oracle.loader.Osc for a (th12, th13, th23, dcp, dm12, dm23) row."""
    from oracle.loader import Osc
    t12, t13, t23, d, m12, m23 = [float(x) for x in params_row]
    return Osc(t12, t13, t23, d, m12, m23, -1 if nutype else 1)
