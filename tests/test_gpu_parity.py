"""Parity of the HIP kernels (through the C ABI) with the reference.

FP64: within 1e-10 absolute per probability of the reference CPU propagator
(golden vectors from the compiled reference) and of the oracle on the same
inputs, with unitarity checked.  FP32 and mixed: error bounds against the FP64
oracle (mixed: 1e-5 max absolute).
"""
import numpy as np
import pytest

from conftest import golden, osc_for
from cudaprob3_amd import workloads as W

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-10
TOL_MIXED = 1e-5
TOL_FP32 = 2e-3


def run_atmos(cz, ev, radii, rhos, yps, params, nutype, precision="fp64",
              h_km=22.0, nc=0, w=0, mode="auto"):
    from cudaprob3_amd import AtmosPropagator
    with AtmosPropagator(len(cz), len(ev), precision=precision) as p:
        p.setSolveMode(mode)
        if nc or w:
            p.setTuning(nc, w)
        p.setEnergyList(ev)
        p.setCosineList(cz)
        t12, t13, t23, d, m12, m23 = [float(x) for x in params]
        p.setMNSMatrix(t12, t13, t23, d, -1 if nutype else 1)
        p.setNeutrinoMasses(m12, m23)
        p.setDensity(radii, rhos, yps)
        p.setProductionHeight(h_km)
        p.calculateProbabilities(nutype)
        return p.probabilities()


def run_beam(ev, params, nutype, precision="fp64", rho=2.6, L=1300.0):
    from cudaprob3_amd import BeamPropagator
    with BeamPropagator(len(ev), precision=precision) as p:
        p.setEnergyList(ev)
        t12, t13, t23, d, m12, m23 = [float(x) for x in params]
        p.setMNSMatrix(t12, t13, t23, d, -1 if nutype else 1)
        p.setNeutrinoMasses(m12, m23)
        p.setDensity(rho)
        p.setPathLength(L)
        p.calculateProbabilities(nutype)
        return p.probabilities()


NOM = [W.TH12, W.TH13, W.TH23, W.DCP, W.DM12, W.DM23]


def unitarity(prob):
    P = prob.reshape(-1, 3, 3)
    return max(np.abs(P.sum(-1) - 1).max(), np.abs(P.sum(-2) - 1).max())


@pytest.mark.parametrize("nutype", [0, 1])
def test_prem4_fp64_vs_golden(nutype):
    g = golden(f"g0_prem4_{'nubar' if nutype else 'nu'}")
    out = run_atmos(g["cosv"], g["ev"], g["radii"], g["rhos"], g["yps"], NOM,
                    nutype)
    s = int(g["stride"])
    print("prem4 nutype", nutype, "max|d| vs golden",
          np.abs(out[::s, ::s] - g["prob"]).max())
    assert np.abs(out[::s, ::s] - g["prob"]).max() < TOL_FP64
    assert np.allclose(out.sum(axis=(0, 1)), g["chan_sum"], rtol=0,
                       atol=TOL_FP64 * out.shape[0] * out.shape[1])
    assert unitarity(out) < 1e-12


@pytest.mark.parametrize("nutype", [0, 1])
def test_prem4_fp64_vs_oracle_full(oracle, nutype):
    e = W.prem4()
    cz, ev = W.cos_grid(100), W.log_grid(100, 0.1, 100.0)
    out = run_atmos(cz, ev, e.radii, e.rhos, e.yps, NOM, nutype)
    ref = oracle.atmos(nutype, cz, ev, e.radii, e.rhos, e.yps,
                       osc_for(NOM, nutype), 22.0, threads=4)
    assert np.abs(out - ref).max() < TOL_FP64


@pytest.mark.parametrize("nutype", [0, 1])
def test_shells12_fp64_vs_golden(nutype):
    g = golden(f"g1_shells12_{'nubar' if nutype else 'nu'}")
    out = run_atmos(g["cosv"], g["ev"], g["radii"], g["rhos"], g["yps"], NOM,
                    nutype)
    assert np.abs(out - g["prob"]).max() < TOL_FP64
    assert unitarity(out) < 1e-12


def test_shells59_fp64_vs_golden():
    g = golden("g3_shells59")
    for nt in (0, 1):
        for i, par in enumerate(g["params"]):
            out = run_atmos(g["cosv"], g["ev"], g["radii"], g["rhos"],
                            g["yps"], par, nt)
            assert np.abs(out - g["prob"][nt, i]).max() < TOL_FP64


def test_edges_fp64_vs_golden():
    g = golden("g4_edges")
    for key in ["nominal", "dm12_zero", "dm23_zero", "dcp_zero", "inverted"]:
        for nt in (0, 1):
            out = run_atmos(g["cosv"], g["ev"], g["radii"], g["rhos"],
                            g["yps"], g[f"params_{key}"], nt)
            ref = g[f"prob_{key}_{nt}"]
            assert np.abs(out - ref).max() < TOL_FP64, key


def test_beam_fp64_vs_golden():
    g = golden("g2_beam")
    for nt in (0, 1):
        for i, par in enumerate(g["params"]):
            out = run_beam(g["ev"], par, nt)
            assert np.abs(out - g["prob"][nt, i]).max() < TOL_FP64
            assert unitarity(out) < 1e-12


TUNINGS = ([("shared", nc, w) for w in (1, 2, 4) for nc in (1, 2, 4, 8)]
           + [("table", nc, 1) for nc in (1, 2, 4, 8)] + [("auto", 0, 0)])


@pytest.mark.parametrize("mode,nc,w", TUNINGS)
def test_tunings_identical(mode, nc, w):
    """The solve mode, the cosine block and the waves sharing a solve change
    only where a point's eigen-solutions are made and which lanes share them,
    never a point's arithmetic: bitwise identical to the shared solve with
    one cosine per lane.  37 cosines leave ragged last blocks (and idle
    waves) for every tuning; 70 energies a partial energy tile."""
    e = W.earth_12()
    cz, ev = W.cos_grid(37), W.log_grid(70, 0.1, 100.0)
    for prec in ("fp64", "fp32", "mixed"):
        b = run_atmos(cz, ev, e.radii, e.rhos, e.yps, NOM, 0, prec, nc=1, w=1,
                      mode="shared")
        o = run_atmos(cz, ev, e.radii, e.rhos, e.yps, NOM, 0, prec, nc=nc, w=w,
                      mode=mode)
        assert np.array_equal(b, o), prec


@pytest.mark.parametrize("nutype", [0, 1])
def test_mixed_and_fp32_error(oracle, nutype):
    e = W.earth_12()
    cz, ev = W.cos_grid(120), W.log_grid(120, 0.1, 100.0)
    ref = oracle.atmos(nutype, cz, ev, e.radii, e.rhos, e.yps,
                       osc_for(NOM, nutype), 22.0, threads=4)
    mixed = run_atmos(cz, ev, e.radii, e.rhos, e.yps, NOM, nutype, "mixed")
    fp32 = run_atmos(cz, ev, e.radii, e.rhos, e.yps, NOM, nutype, "fp32")
    dm, d32 = np.abs(mixed - ref), np.abs(fp32 - ref)
    print(f"nutype {nutype}: mixed max {dm.max():.2e} rms "
          f"{np.sqrt((dm**2).mean()):.2e}; fp32 max {d32.max():.2e} rms "
          f"{np.sqrt((d32**2).mean()):.2e}")
    assert dm.max() < TOL_MIXED
    assert d32.max() < TOL_FP32


def test_beam_mixed_error(oracle):
    ev = W.log_grid(3000, 0.1, 20.0)
    for par in W.draw_param_sets(3, 5):
        row = [par.th12, par.th13, par.th23, par.dcp, par.dm12, par.dm23]
        for nt in (0, 1):
            ref = oracle.beam(nt, ev, 2.6, 1300.0, osc_for(row, nt))
            out = run_beam(ev, row, nt, "mixed")
            assert np.abs(out - ref).max() < TOL_MIXED


@pytest.mark.parametrize("nutype", [0, 1])
def test_mixed_error_shells59_full_grid(oracle, nutype):
    """The headline batch's shape: 59 shells, a 400 x 400 grid over
    0.5-50 GeV, seeded parameter sets; mixed within 1e-5 of the FP64 oracle."""
    e = W.earth_59()
    cz, ev = W.cos_grid(400), W.log_grid(400, 0.5, 50.0)
    for p in W.draw_param_sets(2, 1024):
        row = [p.th12, p.th13, p.th23, p.dcp, p.dm12, p.dm23]
        ref = oracle.atmos(nutype, cz, ev, e.radii, e.rhos, e.yps,
                           osc_for(row, nutype), 22.0, threads=8)
        out = run_atmos(cz, ev, e.radii, e.rhos, e.yps, row, nutype, "mixed")
        d = np.abs(out - ref)
        print(f"59 shells 400x400 nutype {nutype}: mixed max {d.max():.2e} "
              f"rms {np.sqrt((d**2).mean()):.2e}")
        assert d.max() < TOL_MIXED


# ------------------------------------------------------ polynomial densities

def run_poly(cz, ev, ep, params, nutype, precision="fp64", path=None, nc=0, w=0,
             mode="auto"):
    from cudaprob3_amd import AtmosPropagator
    with AtmosPropagator(len(cz), len(ev), precision=precision) as p:
        p.setSolveMode(mode)
        if nc or w:
            p.setTuning(nc, w)
        p.setEnergyList(ev)
        p.setCosineList(cz)
        t12, t13, t23, d, m12, m23 = [float(x) for x in params]
        p.setMNSMatrix(t12, t13, t23, d, -1 if nutype else 1)
        p.setNeutrinoMasses(m12, m23)
        if path:
            p.setDensityFromFile(path)
        else:
            p.setDensity(ep.radii, ep.a, ep.b, ep.c, ep.yps)
        p.setProductionHeight(22.0)
        p.calculateProbabilities(nutype)
        return p.probabilities()


@pytest.mark.parametrize("nutype", [0, 1])
def test_poly_fp64_vs_golden(nutype):
    """setDensity(radii, a, b, c, yps): the compiled reference's polynomial
    path (golden g6, cos = -1 included) within 1e-10, unitarity kept."""
    g = golden(f"g6_prem4poly_{'nubar' if nutype else 'nu'}")
    ep = W.prem4_poly()
    out = run_poly(g["cosv"], g["ev"], ep, NOM, nutype)
    print("poly nutype", nutype, "max|d| vs golden", np.abs(out - g["prob"]).max())
    assert np.abs(out - g["prob"]).max() < TOL_FP64
    assert unitarity(out) < 1e-12


def test_poly_against_oracle_file_and_tunings(oracle, tmp_path):
    """A 200 x 200 grid through the polynomial model against the oracle:
    FP64 within 1e-10, mixed within 1e-5, FP32 within 2e-3; the 5-column
    file gives the same bits as the arrays; tunings and the solve mode do
    not change a bit (the model solves per trajectory)."""
    ep = W.prem4_poly()
    cz, ev = W.cos_grid(200), W.log_grid(200, 0.1, 100.0)
    path = str(tmp_path / "poly.dat")
    ep.write(path)
    for nt in (0, 1):
        ref = oracle.atmos_poly(nt, cz, ev, ep.radii, ep.a, ep.b, ep.c, ep.yps,
                                osc_for(NOM, nt), 22.0, threads=8)
        f64 = run_poly(cz, ev, ep, NOM, nt)
        mixed = run_poly(cz, ev, ep, NOM, nt, "mixed")
        f32 = run_poly(cz, ev, ep, NOM, nt, "fp32")
        print("poly 200x200 nutype", nt, "fp64", np.abs(f64 - ref).max(),
              "mixed", np.abs(mixed - ref).max(), "fp32", np.abs(f32 - ref).max())
        assert np.abs(f64 - ref).max() < TOL_FP64
        assert np.abs(mixed - ref).max() < TOL_MIXED
        assert np.abs(f32 - ref).max() < TOL_FP32
        assert np.array_equal(run_poly(cz, ev, ep, NOM, nt, path=path), f64)
        assert np.array_equal(run_poly(cz, ev, ep, NOM, nt, "mixed", nc=4, w=4,
                                       mode="table"), mixed)


def test_poly_then_constant_model():
    """A constant model set after a polynomial one is used as such."""
    from cudaprob3_amd import AtmosPropagator
    e, ep = W.prem4(), W.prem4_poly()
    cz, ev = W.cos_grid(30), W.log_grid(50, 0.1, 100.0)
    base = run_atmos(cz, ev, e.radii, e.rhos, e.yps, NOM, 0)
    with AtmosPropagator(len(cz), len(ev)) as p:
        p.setEnergyList(ev)
        p.setCosineList(cz)
        t12, t13, t23, d, m12, m23 = NOM
        p.setMNSMatrix(t12, t13, t23, d, 1)
        p.setNeutrinoMasses(m12, m23)
        p.setDensity(ep.radii, ep.a, ep.b, ep.c, ep.yps)
        p.setProductionHeight(22.0)
        p.calculateProbabilities(0)
        p.setDensity(e.radii, e.rhos, e.yps)
        p.calculateProbabilities(0)
        assert np.array_equal(p.probabilities(), base)


def test_wide_parameter_sweep(oracle):
    """24 seeded parameter sets from wide ranges (either mass ordering,
    theta13 up to 0.5, delta over the circle) on the 12-shell model: FP64
    within 1e-10 of the oracle, mixed within 1e-5, and FP64 rows and columns
    of P summing to one within 1e-11 (the compiled reference itself reaches
    7.5e-13 on one of these sets: FP64 rounding over the layers)."""
    rng = np.random.default_rng(20261203)
    e = W.earth_12()
    cz, ev = W.cos_grid(48), W.log_grid(48, 0.1, 100.0)
    worst64, worstm = 0.0, 0.0
    for k in range(24):
        par = [rng.uniform(0.1, 1.4), rng.uniform(0.0, 0.5), rng.uniform(0.5, 1.1),
               rng.uniform(-np.pi, np.pi), rng.uniform(1e-5, 2e-4),
               rng.choice([-1.0, 1.0]) * rng.uniform(1e-3, 4e-3)]
        nt = k % 2
        ref = oracle.atmos(nt, cz, ev, e.radii, e.rhos, e.yps, osc_for(par, nt),
                           22.0, threads=8)
        f64 = run_atmos(cz, ev, e.radii, e.rhos, e.yps, par, nt)
        mixed = run_atmos(cz, ev, e.radii, e.rhos, e.yps, par, nt, "mixed")
        worst64 = max(worst64, float(np.abs(f64 - ref).max()))
        worstm = max(worstm, float(np.abs(mixed - ref).max()))
        assert unitarity(f64) < 1e-11, par
    print("wide sweep: fp64", worst64, "mixed", worstm)
    assert worst64 < TOL_FP64
    assert worstm < TOL_MIXED


# ------------------------------------------- production-height averaging

def run_atmos_avg(cz, ev, earth, params, nutype, prob, edges, precision="fp64",
                  h_km=22.0, mode="auto"):
    from cudaprob3_amd import AtmosPropagator
    with AtmosPropagator(len(cz), len(ev), precision=precision) as p:
        p.setSolveMode(mode)
        p.setEnergyList(ev)
        p.setCosineList(cz)
        t12, t13, t23, d, m12, m23 = [float(x) for x in params]
        p.setMNSMatrix(t12, t13, t23, d, -1 if nutype else 1)
        p.setNeutrinoMasses(m12, m23)
        p.setDensity(earth.radii, earth.rhos, earth.yps)
        p.setProductionHeight(h_km)
        p.SetNumberOfProductionHeightBinsForAveraging(len(edges) - 1)
        p.setProductionHeightAveraging(True)
        p.setProductionHeightList(np.asarray(prob).ravel(), edges)
        p.calculateProbabilities(nutype)
        return p.probabilities()


def row_unitarity(prob):
    """Each initial flavour's probabilities sum to 1; with bin weights that
    differ between initial flavours the columns need not."""
    return np.abs(prob.reshape(-1, 3, 3).sum(-1) - 1).max()


def height_prob(nC, nE, nb, seed):
    pr = np.random.default_rng(seed).random((2, 3, nE, nC, nb))
    return pr / pr.sum(axis=-1, keepdims=True)


@pytest.mark.parametrize("nutype", [0, 1])
def test_height_avg_fp64_vs_golden(oracle, nutype):
    """Production-height averaging on the GPU against the golden made by the
    compiled reference's averaging path: within 1e-10 of the oracle with the
    intended sinc, and within 1e-4 of the golden (the reference CPU build's
    integer abs() in its sinc test, math.hpp:50, takes the series for every
    |x| < 1)."""
    g = golden("g7_prem4_height_avg")
    e = W.EarthModel(g["radii"], g["rhos"], g["yps"])
    for mode in ("table", "shared"):
        out = run_atmos_avg(g["cosv"], g["ev"], e, NOM, nutype, g["height_prob"],
                            g["edges_km"], mode=mode)
        ref = oracle.atmos_pha(nutype, g["cosv"], g["ev"], g["radii"], g["rhos"],
                               g["yps"], osc_for(NOM, nutype), 22.0,
                               g["height_prob"], g["edges_km"])
        print("height avg", mode, nutype, "max|d| vs oracle", np.abs(out - ref).max(),
              "vs golden", np.abs(out - g["prob"][nutype]).max())
        assert np.abs(out - ref).max() < TOL_FP64
        assert np.abs(out - g["prob"][nutype]).max() < 1e-4


@pytest.mark.parametrize("precision,tol", [("fp64", TOL_FP64), ("mixed", TOL_MIXED),
                                           ("fp32", TOL_FP32)])
def test_height_avg_vs_oracle(oracle, precision, tol):
    """12 shells, 61 x 70 points (a ragged last energy tile), 7 bins from 0.5
    to 60 km, drawn parameter sets, both neutrino types; the slab-table and
    in-kernel solves give the same bits."""
    e = W.earth_12()
    cz = np.linspace(-1.0, 1.0, 61)
    ev = W.log_grid(70, 0.1, 100.0)
    edges = np.array([0.5, 4.0, 9.0, 15.0, 22.0, 30.0, 41.0, 60.0])
    pr = height_prob(len(cz), len(ev), len(edges) - 1, 8)
    for k, p in enumerate(W.draw_param_sets(2, 88)):
        row = [p.th12, p.th13, p.th23, p.dcp, p.dm12, p.dm23]
        for nt in (0, 1):
            ref = oracle.atmos_pha(nt, cz, ev, e.radii, e.rhos, e.yps,
                                   osc_for(row, nt), 18.0, pr, edges, threads=4)
            a = run_atmos_avg(cz, ev, e, row, nt, pr, edges, precision, 18.0, "table")
            b = run_atmos_avg(cz, ev, e, row, nt, pr, edges, precision, 18.0, "shared")
            print(precision, k, nt, "max|d|", np.abs(a - ref).max())
            assert np.abs(a - ref).max() < tol
            assert np.array_equal(a, b)
            assert row_unitarity(a) < {"fp64": 1e-12, "mixed": TOL_MIXED, "fp32": TOL_FP32}[precision]


def test_height_avg_limits():
    """A zero-width bin at the production height gives the unaveraged
    result, also through a polynomial model (per-trajectory solves); a
    propagator switched back off runs the plain path again."""
    cz = np.linspace(-1.0, 0.95, 40)
    ev = W.log_grid(90, 0.1, 50.0)
    one = np.ones((2, 3, len(ev), len(cz), 1))
    e = W.prem4()
    for nt in (0, 1):
        plain = run_atmos(cz, ev, e.radii, e.rhos, e.yps, NOM, nt)
        avg = run_atmos_avg(cz, ev, e, NOM, nt, one, [22.0, 22.0])
        assert np.abs(avg - plain).max() < 1e-12
    ep = W.prem4_poly()
    from cudaprob3_amd import AtmosPropagator
    with AtmosPropagator(len(cz), len(ev)) as p:
        p.setEnergyList(ev)
        p.setCosineList(cz)
        p.setMNSMatrix(W.TH12, W.TH13, W.TH23, W.DCP, 1)
        p.setNeutrinoMasses(W.DM12, W.DM23)
        p.setDensity(ep.radii, ep.a, ep.b, ep.c, ep.yps)
        p.setProductionHeight(22.0)
        p.calculateProbabilities(0)
        plain = p.probabilities()
        p.SetNumberOfProductionHeightBinsForAveraging(1)
        p.setProductionHeightAveraging(True)
        p.setProductionHeightList(one.ravel(), [22.0, 22.0])
        p.calculateProbabilities(0)
        assert np.abs(p.probabilities() - plain).max() < 1e-12
        p.setProductionHeightAveraging(False)
        p.calculateProbabilities(0)
        assert np.array_equal(p.probabilities(), plain)


def test_height_avg_errors():
    """The reference's call-order and size rules (atmoscpupropagator.hpp:88-108,
    395-427, 453-454), as Cp3Error."""
    from cudaprob3_amd import AtmosPropagator
    from cudaprob3_amd._native import Cp3Error
    with AtmosPropagator(3, 4) as p:
        p.setEnergyList([1.0, 2.0, 3.0, 4.0])
        p.setCosineList([-1.0, -0.5, 0.2])
        p.setMNSMatrix(W.TH12, W.TH13, W.TH23, W.DCP, 1)
        p.setNeutrinoMasses(W.DM12, W.DM23)
        p.setDensity([0.0, 3480.0, 6371.0], [11.3, 11.3, 4.0], [0.5, 0.5, 0.5])
        p.setProductionHeight(15.0)
        with pytest.raises(Cp3Error, match="invalid number of production height bins"):
            p.SetNumberOfProductionHeightBinsForAveraging(21)
        with pytest.raises(Cp3Error, match="not expecting"):
            p.setProductionHeightList(np.ones(72), [1.0, 2.0])
        p.SetNumberOfProductionHeightBinsForAveraging(1)
        p.setProductionHeightAveraging(True)
        with pytest.raises(Cp3Error, match="production height array was not set"):
            p.calculateProbabilities(0)
        with pytest.raises(Cp3Error, match="Prob array is not the expected size"):
            p.setProductionHeightList(np.ones(71), [1.0, 2.0])
        with pytest.raises(Cp3Error, match="ProductionHeightBins array"):
            p.setProductionHeightList(np.ones(72), [1.0, 2.0, 3.0])
        # 1 bin x 2 types x 3 flavours x 4 energies x 3 cosines
        p.setProductionHeightList(np.full(72, 1.0), [10.0, 20.0])
        p.calculateProbabilities(1)
        assert unitarity(p.probabilities()) < 1e-12
        p.SetNumberOfProductionHeightBinsForAveraging(2)  # turns it off again
        with pytest.raises(Cp3Error, match="not expecting"):
            p.setProductionHeightList(np.ones(144), [1.0, 2.0, 3.0])


@pytest.mark.parametrize("precision,tol", [("fp64", TOL_FP64), ("mixed", TOL_MIXED)])
def test_height_avg_poly_vs_oracle(oracle, precision, tol):
    """Averaging through a polynomial model (per-trajectory solves, one
    cosine per wave), against the oracle's averaging through the same model,
    which equals the compiled reference's bitwise with its sinc."""
    from cudaprob3_amd import AtmosPropagator
    ep = W.prem4_poly()
    cz = np.concatenate([W.cos_grid(50), np.linspace(0.0, 0.95, 10)])
    ev = W.log_grid(90, 0.1, 100.0)
    edges = np.array([3.0, 11.0, 22.0, 35.0, 47.0])
    pr = height_prob(len(cz), len(ev), len(edges) - 1, 12)
    for nt in (0, 1):
        ref = oracle.atmos_poly_pha(nt, cz, ev, ep.radii, ep.a, ep.b, ep.c, ep.yps,
                                    osc_for(NOM, nt), 22.0, pr, edges, threads=4)
        with AtmosPropagator(len(cz), len(ev), precision=precision) as p:
            p.setEnergyList(ev)
            p.setCosineList(cz)
            p.setMNSMatrix(W.TH12, W.TH13, W.TH23, W.DCP, -1 if nt else 1)
            p.setNeutrinoMasses(W.DM12, W.DM23)
            p.setDensity(ep.radii, ep.a, ep.b, ep.c, ep.yps)
            p.setProductionHeight(22.0)
            p.SetNumberOfProductionHeightBinsForAveraging(len(edges) - 1)
            p.setProductionHeightAveraging(True)
            p.setProductionHeightList(pr.ravel(), edges)
            p.calculateProbabilities(nt)
            got = p.probabilities()
        print(precision, nt, "max|d|", np.abs(got - ref).max())
        assert np.abs(got - ref).max() < tol


def test_empty_grids_rejected():
    """An empty grid is refused at construction with a status code (the
    reference accepts it and ends the process on the empty kernel launch,
    atmoscudapropagator.cu:284-287 with hpc_helpers.cuh:71-78)."""
    from cudaprob3_amd import AtmosPropagator, BeamPropagator
    from cudaprob3_amd._native import Cp3Error
    for nc, ne in ((0, 10), (10, 0), (0, 0)):
        with pytest.raises(Cp3Error, match="must be positive"):
            AtmosPropagator(nc, ne)
    with pytest.raises(Cp3Error, match="must be positive"):
        BeamPropagator(0)


@pytest.mark.parametrize("mode", ["table", "shared"])
def test_three_column_file_vs_golden(tmp_path, mode):
    """setDensityFromFile with the rows of the reference's
    models/PREM_4layer.dat written as a 3-column file (runs of spaces, a
    trailing comment line): the compiled reference's result on that file
    (golden g5) within 1e-10 in FP64 and 1e-5 mixed, and the same bits as
    setDensity with the same rows; a 2-column file is refused.  As in the
    reference (atmoscpupropagator.hpp:289-296), the rows of a 3-column file
    are read as a stream of numbers up to the first non-number, so comment
    lines may follow the rows but not precede them."""
    from cudaprob3_amd import AtmosPropagator
    from cudaprob3_amd._native import Cp3Error
    g = golden("g5_prem4_file")
    rows = g["model_rows"]
    path = tmp_path / "prem4.dat"
    with open(path, "w") as f:
        for r in rows:
            f.write("   ".join(repr(float(x)) for x in r) + "\n")
        f.write("# radius density Ye\n")
    t12, t13, t23, d, m12, m23 = NOM
    for prec, tol in (("fp64", TOL_FP64), ("mixed", TOL_MIXED)):
        with AtmosPropagator(20, 20, precision=prec) as p:
            p.setSolveMode(mode)
            p.setEnergyList(g["ev"])
            p.setCosineList(g["cosv"])
            p.setMNSMatrix(t12, t13, t23, d, 1)
            p.setNeutrinoMasses(m12, m23)
            p.setDensityFromFile(str(path))
            p.setProductionHeight(22.0)
            p.calculateProbabilities(0)
            got = p.probabilities()
            assert p.getNlayerBoundaries() == 5
        err = np.abs(got - g["prob"]).max()
        print("3-column file", prec, mode, "max|d| vs golden", err)
        assert err < tol
        arrays = run_atmos(g["cosv"], g["ev"], rows[:, 0], rows[:, 1], rows[:, 2],
                           NOM, 0, prec, mode=mode)
        assert np.array_equal(got, arrays)
    legacy = tmp_path / "legacy.dat"
    legacy.write_text("0. 13.0\n1220.0 13.0\n6371.0 3.3\n")
    with AtmosPropagator(20, 20) as p:
        with pytest.raises(Cp3Error, match="2 entries per line"):
            p.setDensityFromFile(str(legacy))


def test_calculate_without_model_refused():
    """calculateProbabilities before any density model is a call-order error
    (checked before any upload or kernel launch), and the object stays
    usable: after setDensity it calculates normally."""
    from cudaprob3_amd import AtmosPropagator
    from cudaprob3_amd._native import Cp3Error, CP3_ERR_STATE
    e = W.prem4()
    cz, ev = W.cos_grid(8), W.log_grid(16, 0.1, 100.0)
    t12, t13, t23, d, m12, m23 = NOM
    for mode in ("table", "shared"):
        with AtmosPropagator(len(cz), len(ev)) as p:
            p.setSolveMode(mode)
            p.setEnergyList(ev)
            p.setCosineList(cz)
            p.setMNSMatrix(t12, t13, t23, d, 1)
            p.setNeutrinoMasses(m12, m23)
            p.setProductionHeight(22.0)
            with pytest.raises(Cp3Error, match="density model was not set") as ex:
                p.calculateProbabilities(0)
            assert ex.value.code == CP3_ERR_STATE
            p.setDensity(e.radii, e.rhos, e.yps)
            p.calculateProbabilities(0)
            ref = run_atmos(cz, ev, e.radii, e.rhos, e.yps, NOM, 0)
            assert np.array_equal(p.probabilities(), ref)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_channel_fetch_follows_recalculation(precision):
    """A channel reaches the host on its first access after a calculation:
    reads of single channels in any order, before and after a recalculation
    with other parameters, equal a fresh propagator's full result bitwise
    (no stale channel survives the recalculation)."""
    from cudaprob3_amd import AtmosPropagator, BeamPropagator
    e = W.earth_12()
    cz, ev = W.cos_grid(40), W.log_grid(56, 0.5, 50.0)
    other = list(NOM)
    other[3], other[5] = 0.7, 2.6e-3
    ref_a = run_atmos(cz, ev, e.radii, e.rhos, e.yps, NOM, 0, precision)
    ref_b = run_atmos(cz, ev, e.radii, e.rhos, e.yps, other, 0, precision)
    assert not np.array_equal(ref_a[:, :, 4], ref_b[:, :, 4])
    with AtmosPropagator(len(cz), len(ev), precision=precision) as p:
        p.setEnergyList(ev)
        p.setCosineList(cz)
        p.setDensity(e.radii, e.rhos, e.yps)
        p.setProductionHeight(22.0)
        for par, ref in ((NOM, ref_a), (other, ref_b)):
            t12, t13, t23, d, m12, m23 = par
            p.setMNSMatrix(t12, t13, t23, d, 1)
            p.setNeutrinoMasses(m12, m23)
            p.calculateProbabilities(0)
            assert p.getProbability(3, 5, 4) == ref[3, 5, 4]
            arr = p.getProbabilityArr(7).reshape(len(ev), len(cz)).T
            assert np.array_equal(arr, ref[:, :, 7])
            assert p.getProbability(39, 55, 0) == ref[39, 55, 0]
        assert np.array_equal(p.probabilities(), ref_b)
    bev = W.log_grid(300, 0.1, 20.0)
    rb_a = run_beam(bev, NOM, 0, precision)
    rb_b = run_beam(bev, other, 0, precision)
    with BeamPropagator(len(bev), precision=precision) as p:
        p.setEnergyList(bev)
        p.setDensity(2.6)
        p.setPathLength(1300.0)
        for par, ref in ((NOM, rb_a), (other, rb_b)):
            t12, t13, t23, d, m12, m23 = par
            p.setMNSMatrix(t12, t13, t23, d, 1)
            p.setNeutrinoMasses(m12, m23)
            p.calculateProbabilities(0)
            assert p.getProbability(17, 2) == ref[17, 2]
            assert np.array_equal(p.getProbabilityArr(8), ref[:, 8])
        assert np.array_equal(p.probabilities(), rb_b)


def test_extreme_energies_and_cosines(oracle):
    """Energies from 1 MeV to 10 TeV and cosines from -1 to 0.99 (above the
    horizon: atmosphere only) on the 12-shell model, the nominal set and two
    wide-range ones, both neutrino types: FP64 within 1e-10 of the oracle and
    mixed within 1e-5.  At 1 MeV a layer's phase reaches 1e4-1e5 radians;
    the host build of the same arithmetic stays within 3.3e-11 (FP64) and
    6.3e-6 (mixed) there."""
    e = W.earth_12()
    cz = np.concatenate([np.linspace(-1.0, -1e-4, 40), np.linspace(0.0, 0.99, 8)])
    ev = W.log_grid(60, 1e-3, 1e4)
    rng = np.random.default_rng(5)
    sets = [NOM] + [[rng.uniform(0.1, 1.4), rng.uniform(0.0, 0.5), rng.uniform(0.5, 1.1),
                     rng.uniform(-np.pi, np.pi), rng.uniform(1e-5, 2e-4),
                     rng.choice([-1.0, 1.0]) * rng.uniform(1e-3, 4e-3)] for _ in range(2)]
    worst64, worstm = 0.0, 0.0
    for par in sets:
        for nt in (0, 1):
            ref = oracle.atmos(nt, cz, ev, e.radii, e.rhos, e.yps, osc_for(par, nt),
                               22.0, threads=8)
            f64 = run_atmos(cz, ev, e.radii, e.rhos, e.yps, par, nt)
            mixed = run_atmos(cz, ev, e.radii, e.rhos, e.yps, par, nt, "mixed")
            worst64 = max(worst64, float(np.abs(f64 - ref).max()))
            worstm = max(worstm, float(np.abs(mixed - ref).max()))
    print(f"extreme E and cos: fp64 {worst64:.2e}, mixed {worstm:.2e}")
    assert worst64 < TOL_FP64
    assert worstm < TOL_MIXED


@pytest.mark.parametrize("rho,L", [(0.0, 295.0), (2.6, 1300.0), (13.0, 12742.0)])
def test_beam_extreme_inputs(oracle, rho, L):
    """The beam over 1 MeV-10 TeV in vacuum, at 2.6 g/cm3 over 1300 km and at
    13 g/cm3 over the Earth's diameter: FP64 within 1e-10 of the oracle,
    mixed within 1e-5 (the host build: 2.3e-11 and 1.1e-6)."""
    ev = W.log_grid(400, 1e-3, 1e4)
    for nt in (0, 1):
        ref = oracle.beam(nt, ev, rho, L, osc_for(NOM, nt), threads=8)
        f64 = run_beam(ev, NOM, nt, "fp64", rho, L)
        mixed = run_beam(ev, NOM, nt, "mixed", rho, L)
        assert np.abs(f64 - ref).max() < TOL_FP64
        assert np.abs(mixed - ref).max() < TOL_MIXED
