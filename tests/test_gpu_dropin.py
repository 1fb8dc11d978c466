"""The C++ drop-in classes (include/cudaprob3_amd/hip_propagator.hpp), driven
by tests/cpp/dropin_main.cpp the way the reference examples drive
AtmosCudaPropagatorSingle / BeamCudaPropagatorSingle, against the oracle."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, osc_for
from cudaprob3_amd import workloads as W

pytestmark = pytest.mark.gpu

EXE = os.path.join(ROOT, "tests", "cpp", "build", "dropin_main")
NOM = [W.TH12, W.TH13, W.TH23, W.DCP, W.DM12, W.DM23]


def run(*args):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built (run __graft_entry__.build())")
    r = subprocess.run([EXE, *map(str, args)], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr)
    return r


@pytest.mark.parametrize("nutype", [0, 1])
def test_dropin_atmos_fp64(oracle, tmp_path, nutype):
    out = tmp_path / "a.bin"
    r = run("atmos", out, 0, nutype)
    assert r.returncode == 0
    got = np.fromfile(out, dtype=np.float64).reshape(100, 100, 9)
    e = W.prem4()
    ref = oracle.atmos(nutype, W.cos_grid(100), W.log_grid(100, 0.1, 100.0),
                       e.radii, e.rhos, e.yps, osc_for(NOM, nutype), 22.0,
                       threads=4)
    assert np.abs(got - ref).max() < 1e-10


def test_dropin_atmos_mixed(oracle, tmp_path):
    out = tmp_path / "a.bin"
    assert run("atmos", out, 2, 0).returncode == 0
    got = np.fromfile(out, dtype=np.float64).reshape(100, 100, 9)
    e = W.prem4()
    ref = oracle.atmos(0, W.cos_grid(100), W.log_grid(100, 0.1, 100.0),
                       e.radii, e.rhos, e.yps, osc_for(NOM, 0), 22.0,
                       threads=4)
    assert np.abs(got - ref).max() < 1e-5


@pytest.mark.parametrize("nutype", [0, 1])
def test_dropin_atmos_poly_fp64(oracle, tmp_path, nutype):
    """setDensity(radii, a, b, c, yps) through the C++ class."""
    out = tmp_path / "p.bin"
    assert run("poly", out, 0, nutype).returncode == 0
    got = np.fromfile(out, dtype=np.float64).reshape(100, 100, 9)
    ep = W.prem4_poly()
    ref = oracle.atmos_poly(nutype, W.cos_grid(100), W.log_grid(100, 0.1, 100.0),
                            ep.radii, ep.a, ep.b, ep.c, ep.yps,
                            osc_for(NOM, nutype), 22.0, threads=4)
    assert np.abs(got - ref).max() < 1e-10


def dropin_height_prob(nC, nE, nb):
    """The bin probabilities of dropin_main's avg case."""
    t, f, e, c, b = np.meshgrid(np.arange(2), np.arange(3), np.arange(nE),
                                np.arange(nC), np.arange(nb), indexing="ij")
    w = 1.0 + ((7 * t + 5 * f + 3 * e + c + 11 * b) % 13)
    return w / w.sum(axis=-1, keepdims=True)


@pytest.mark.parametrize("prec,tol", [(0, 1e-10), (2, 1e-5)])
@pytest.mark.parametrize("nutype", [0, 1])
def test_dropin_height_avg(oracle, tmp_path, nutype, prec, tol):
    """Production-height averaging through the C++ class
    (setProductionHeightAveraging, then the reference's
    SetNumberOfProductionHeightBinsForAveraging / setProductionHeightList)."""
    out = tmp_path / "h.bin"
    assert run("avg", out, prec, nutype).returncode == 0
    got = np.fromfile(out, dtype=np.float64).reshape(100, 100, 9)
    e = W.prem4()
    ref = oracle.atmos_pha(nutype, W.cos_grid(100), W.log_grid(100, 0.1, 100.0),
                           e.radii, e.rhos, e.yps, osc_for(NOM, nutype), 22.0,
                           dropin_height_prob(100, 100, 5),
                           [5.0, 10.0, 15.0, 22.0, 30.0, 50.0], threads=4)
    assert np.abs(got - ref).max() < tol


@pytest.mark.parametrize("nutype", [0, 1])
def test_dropin_beam_fp64(oracle, tmp_path, nutype):
    out = tmp_path / "b.bin"
    assert run("beam", out, 0, nutype).returncode == 0
    got = np.fromfile(out, dtype=np.float64).reshape(500, 9)
    ref = oracle.beam(nutype, W.log_grid(500, 0.1, 20.0), 2.6, 1300.0,
                      osc_for(NOM, nutype))
    assert np.abs(got - ref).max() < 1e-10


def test_dropin_errors():
    r = run("errors")
    assert r.returncode == 0, r.stderr


@pytest.mark.parametrize("prec,tol", [(0, 1e-10), (2, 1e-5)])
@pytest.mark.parametrize("nutype", [0, 1])
def test_dropin_through_base_pointers(oracle, tmp_path, nutype, prec, tol):
    """Both classes behind std::unique_ptr<cudaprob3linear::Propagator<double>>
    (the reference's names from cudaprob3linear.hpp), configured through the
    base and the atmospheric/beam interfaces only, against the oracle."""
    out = tmp_path / "v.bin"
    assert run("virtual", out, prec, nutype).returncode == 0
    got = np.fromfile(out, dtype=np.float64)
    atm = got[:100 * 100 * 9].reshape(100, 100, 9)
    beam = got[100 * 100 * 9:].reshape(500, 9)
    e = W.prem4()
    ref = oracle.atmos(nutype, W.cos_grid(100), W.log_grid(100, 0.1, 100.0),
                       e.radii, e.rhos, e.yps, osc_for(NOM, nutype), 22.0,
                       threads=4)
    assert np.abs(atm - ref).max() < tol
    refb = oracle.beam(nutype, W.log_grid(500, 0.1, 20.0), 2.6, 1300.0,
                       osc_for(NOM, nutype))
    assert np.abs(beam - refb).max() < tol


def write_rows(path, rows):
    """A density file laid out like models/PREM_4layer.dat (whitespace
    separated rows, here with runs of spaces)."""
    with open(path, "w") as f:
        for r in rows:
            f.write("   ".join(repr(float(x)) for x in r) + "\n")


@pytest.mark.parametrize("prec,tol", [(0, 1e-10), (2, 1e-5)])
def test_dropin_three_column_file_vs_golden(tmp_path, prec, tol):
    """setDensityFromFile on the rows of the reference's models/PREM_4layer.dat
    (golden g5, written by the compiled reference from that file) through
    the C++ class on the GPU."""
    from conftest import golden
    g = golden("g5_prem4_file")
    model = tmp_path / "prem4.dat"
    write_rows(model, g["model_rows"])
    out = tmp_path / "f.bin"
    r = run("file", out, prec, 0, model)
    assert r.returncode == 0
    assert "5 layer boundaries" in r.stdout
    got = np.fromfile(out, dtype=np.float64).reshape(20, 20, 9)
    err = np.abs(got - g["prob"]).max()
    print("3-column file through the C++ class, precision", prec, "max|d|", err)
    assert err < tol


def test_dropin_two_column_file_refused(tmp_path):
    """A 2-column (legacy) file is refused with an exception (the reference
    ends the process, atmoscpupropagator.hpp:341)."""
    model = tmp_path / "legacy.dat"
    with open(model, "w") as f:
        f.write("0. 13.0\n1220.0 13.0\n6371.0 3.3\n")
    r = run("file", tmp_path / "x.bin", 0, 0, model)
    assert r.returncode == 4
    assert "exception" in r.stderr
