"""Pins the oracle (CPU restatement of the reference propagator).

1. Against the golden vectors written by the compiled reference
   (tests/golden/, see make_golden.py): equal to <= 1e-13 (bitwise in
   practice).
2. Against the compiled reference itself (oracle/_ref) where it is built.
3. Against what the mathematics fixes, independently of the reference code:
   PMNS unitarity, matter eigenvalues vs numpy.linalg.eigvalsh, one layer vs
   scipy.linalg.expm, the vacuum formula, chord geometry, path unitarity,
   T-symmetry at delta = 0 and the nu / nu-bar relation.
"""
import os

import numpy as np
import pytest
import scipy.linalg

from conftest import golden, osc_for
from cudaprob3_amd import workloads as W

TOL_GOLD = 1e-13
NOM = [W.TH12, W.TH13, W.TH23, W.DCP, W.DM12, W.DM23]


@pytest.mark.parametrize("nutype", [0, 1])
def test_prem4_golden(oracle, nutype):
    g = golden(f"g0_prem4_{'nubar' if nutype else 'nu'}")
    out = oracle.atmos(nutype, g["cosv"], g["ev"], g["radii"], g["rhos"],
                       g["yps"], osc_for(NOM, nutype), 22.0, threads=4)
    s = int(g["stride"])
    assert np.abs(out[::s, ::s] - g["prob"]).max() <= TOL_GOLD
    assert np.allclose(out.sum(axis=(0, 1)), g["chan_sum"], rtol=1e-13,
                       atol=0)
    assert np.allclose((out * out).sum(axis=(0, 1)), g["chan_sumsq"],
                       rtol=1e-13, atol=0)


@pytest.mark.parametrize("nutype", [0, 1])
def test_shells12_golden(oracle, nutype):
    g = golden(f"g1_shells12_{'nubar' if nutype else 'nu'}")
    out = oracle.atmos(nutype, g["cosv"], g["ev"], g["radii"], g["rhos"],
                       g["yps"], osc_for(NOM, nutype), 22.0, threads=4)
    assert np.abs(out - g["prob"]).max() <= TOL_GOLD


def test_shells59_golden(oracle):
    g = golden("g3_shells59")
    for nt in (0, 1):
        for i, par in enumerate(g["params"]):
            out = oracle.atmos(nt, g["cosv"], g["ev"], g["radii"], g["rhos"],
                               g["yps"], osc_for(par, nt), 22.0, threads=4)
            assert np.abs(out - g["prob"][nt, i]).max() <= TOL_GOLD


def test_beam_golden(oracle):
    g = golden("g2_beam")
    for nt in (0, 1):
        for i, par in enumerate(g["params"]):
            out = oracle.beam(nt, g["ev"], float(g["rho"]), float(g["L_km"]),
                              osc_for(par, nt))
            assert np.abs(out - g["prob"][nt, i]).max() <= TOL_GOLD


def test_edge_golden(oracle):
    g = golden("g4_edges")
    for key in ["nominal", "dm12_zero", "dm23_zero", "dcp_zero", "inverted"]:
        for nt in (0, 1):
            out = oracle.atmos(nt, g["cosv"], g["ev"], g["radii"], g["rhos"],
                               g["yps"], osc_for(g[f"params_{key}"], nt),
                               22.0)
            assert np.abs(out - g[f"prob_{key}_{nt}"]).max() <= TOL_GOLD, key


def test_density_file_golden(oracle, tmp_path):
    g = golden("g5_prem4_file")
    path = tmp_path / "model.dat"
    # the reference's 3-column file, rewritten with repeated spaces
    with open(path, "w") as f:
        for r, d, y in g["model_rows"]:
            f.write(f"{float(r)!r}  {float(d)!r}   {float(y)!r}\n")
    out = oracle.atmos_file(0, g["cosv"], g["ev"], str(path),
                            osc_for(NOM, 0), 22.0)
    assert np.abs(out - g["prob"]).max() <= TOL_GOLD


def test_oracle_matches_compiled_reference(oracle):
    from oracle.loader import load_ref
    ref = load_ref(deep=True)
    if ref is None:
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    e = W.earth_12()
    cz, ev = W.cos_grid(40), W.log_grid(40, 0.1, 100.0)
    for par in [NOM] + [[p.th12, p.th13, p.th23, p.dcp, p.dm12, p.dm23]
                        for p in W.draw_param_sets(2, 99)]:
        for nt in (0, 1):
            a = oracle.atmos(nt, cz, ev, e.radii, e.rhos, e.yps,
                             osc_for(par, nt), 22.0, threads=4)
            b = ref.atmos(nt, cz, ev, e.radii, e.rhos, e.yps,
                          osc_for(par, nt), 22.0, threads=4)
            assert np.array_equal(a, b)


# ---------------------------------------------------------------- math pins

def pdg_u(t12, t13, t23, d):
    s12, c12 = np.sin(t12), np.cos(t12)
    s13, c13 = np.sin(t13), np.cos(t13)
    s23, c23 = np.sin(t23), np.cos(t23)
    R23 = np.array([[1, 0, 0], [0, c23, s23], [0, -s23, c23]], complex)
    U13 = np.array([[c13, 0, s13 * np.exp(-1j * d)], [0, 1, 0],
                    [-s13 * np.exp(1j * d), 0, c13]])
    R12 = np.array([[c12, s12, 0], [-s12, c12, 0], [0, 0, 1]], complex)
    return R23 @ U13 @ R12


def test_pmns_closed_form(oracle):
    for knu in (1, -1):
        U = oracle.pmns(W.TH12, W.TH13, W.TH23, W.DCP, knu)
        assert np.abs(U @ U.conj().T - np.eye(3)).max() < 1e-15
        expect = pdg_u(W.TH12, W.TH13, W.TH23, W.DCP * knu)
        assert np.abs(U - expect).max() < 1e-15


def mass_matrix(osc, E, rho_ye, nutype):
    """2E*H in the mass basis, eV^2: diag(0, dm21, dm31) + A U^+ e e^+ U."""
    U = pdg_u(osc.th12, osc.th13, osc.th23, osc.dcp * osc.knutype)
    d = np.diag([0.0, osc.dm12, osc.dm12 + osc.dm23])
    A = 1.52588e-4 * E * rho_ye * (-1 if nutype else 1)
    return d + A * np.outer(U[0].conj(), U[0])


@pytest.mark.parametrize("nutype", [0, 1])
def test_matter_eigenvalues(oracle, nutype):
    osc = osc_for(NOM, nutype)
    for E in [0.1, 0.5, 2.0, 5.0, 20.0, 100.0]:
        for rye in [0.0, 1.5, 6.5]:
            lam, order = oracle.matter_eigs(osc, E, rye, nutype)
            ev = np.linalg.eigvalsh(mass_matrix(osc, E, rye, nutype))
            assert np.abs(np.sort(lam) - ev).max() < 1e-15 + 1e-12 * np.abs(ev).max()
            assert sorted(order) == [0, 1, 2]
    # continuity to vacuum: at zero density l_i = DM(i,0)
    lam, _ = oracle.matter_eigs(osc, 1.0, 0.0, nutype)
    assert np.abs(lam - [0.0, W.DM12, W.DM12 + W.DM23]).max() < 1e-15


@pytest.mark.parametrize("nutype", [0, 1])
def test_one_layer_is_exponential(oracle, nutype):
    """T = U expm(-i 2.534 M L / E) U^+ (flavour basis, [after][before])."""
    osc = osc_for(NOM, nutype)
    U = pdg_u(osc.th12, osc.th13, osc.th23, osc.dcp * osc.knutype)
    for E, rye, L in [(0.3, 1.6, 500.0), (3.0, 6.5, 2400.0),
                      (25.0, 4.0, 9000.0), (1.0, 0.0, 295.0)]:
        T, P = oracle.layer(osc, E, rye, L, nutype)
        M = mass_matrix(osc, E, rye, nutype)
        expect = U @ scipy.linalg.expm(-1j * 2.534 * M * L / E) @ U.conj().T
        assert np.abs(T - expect).max() < 1e-12
        # Sylvester projectors: idempotent, complete, eigen-projectors of M
        Pk = [P[:, :, k] for k in range(3)]
        assert np.abs(sum(Pk) - np.eye(3)).max() < 1e-11
        for p in Pk:
            assert np.abs(p @ p - p).max() < 1e-11


def vacuum_prob(osc, L, E):
    U = pdg_u(osc.th12, osc.th13, osc.th23, osc.dcp * osc.knutype)
    m = np.array([0.0, osc.dm12, osc.dm12 + osc.dm23])
    amp = U @ np.diag(np.exp(-1j * 2.534 * m * L / E)) @ U.conj().T
    return np.abs(amp) ** 2  # [after][before]


def test_vacuum_paths(oracle):
    """cos >= 0 crosses no shell: P(a->b) = |sum_k U_bk U*_ak e^{-i..}|^2
    over the path length from the production height."""
    e = W.prem4()
    osc = osc_for(NOM, 0)
    cz = np.array([0.0, 0.3, 0.9, 1.0])
    ev = np.array([0.2, 1.0, 7.0])
    out = oracle.atmos(0, cz, ev, e.radii, e.rhos, e.yps, osc, 15.0)
    R, h = 6371.0, 15.0
    for i, c in enumerate(cz):
        L = np.sqrt((R + h) ** 2 - R * R * (1 - c * c)) - R * c
        for j, E in enumerate(ev):
            P = vacuum_prob(osc, L, E)
            got = out[i, j].reshape(3, 3)  # [a][b] = P(a->b)
            assert np.abs(got - P.T).max() < 1e-12


def test_beam_vacuum_limit(oracle):
    osc = osc_for(NOM, 0)
    ev = np.array([0.3, 1.0, 2.5])
    out = oracle.beam(0, ev, 0.0, 810.0, osc)
    for j, E in enumerate(ev):
        assert np.abs(out[j].reshape(3, 3) - vacuum_prob(osc, 810.0, E).T).max() < 1e-12


def test_geometry(oracle):
    """Segments sum to the path length; shells are crossed symmetrically;
    each shell's chord is 2 sqrt(r^2 - b^2)."""
    e = W.earth_12()
    R, h = 6371.0, 22.0
    for c in [-1.0, -0.93, -0.5, -0.21, -0.04]:
        M, seg, rye = oracle.segments(c, e.radii, e.rhos, e.yps, h)
        L = np.sqrt((R + h) ** 2 - R * R * (1 - c * c)) - R * c
        total = seg[0] + 2 * seg[1:M].sum() + seg[M]
        assert abs(total - L) < 1e-9
        b = R * np.sqrt(1 - c * c)
        radii = e.radii[::-1]
        chords = 2 * np.sqrt(np.maximum(radii[:M] ** 2 - b * b, 0))
        # the innermost shell is crossed once along its full chord
        assert abs(seg[M] - chords[M - 1]) < 1e-9
        assert abs(seg[0] - (L - (-2 * c * R))) < 1e-9
        expect_M = int(np.sum(c < -np.sqrt(1 - (radii[1:] / R) ** 2))) + 1
        assert M == expect_M


def test_path_invariants(oracle):
    """Unitarity of P; T-symmetry P(a->b) = P(b->a) at delta = 0 on the
    symmetric Earth path (no atmosphere); nu-bar equals nu with
    (delta, A) -> (-delta, -A)."""
    e = W.earth_12()
    cz = np.array([-0.999, -0.8, -0.45, -0.1])
    ev = W.log_grid(15, 0.2, 80.0)
    p = osc_for(NOM, 0)
    out = oracle.atmos(0, cz, ev, e.radii, e.rhos, e.yps, p, 22.0)
    P = out.reshape(-1, 3, 3)
    assert np.abs(P.sum(-1) - 1).max() < 1e-12
    assert np.abs(P.sum(-2) - 1).max() < 1e-12
    # T-symmetry: h = 0 removes the atmosphere segment, so the path is
    # palindromic; with delta = 0 the amplitude matrix is symmetric.
    p0 = osc_for([W.TH12, W.TH13, W.TH23, 0.0, W.DM12, W.DM23], 0)
    out0 = oracle.atmos(0, cz, ev, e.radii, e.rhos, e.yps, p0, 0.0)
    P0 = out0.reshape(-1, 3, 3)
    assert np.abs(P0 - P0.transpose(0, 2, 1)).max() < 1e-12
    # antineutrino relation: calculate(Antineutrino) with kNuType = -1 equals
    # the neutrino code path run with the conjugated parameters, checked
    # through T-symmetry of the conjugate: P_nubar(a->b; d) = P_nu(b->a; -d)
    # only holds for symmetric paths; use h = 0 again.
    pb = osc_for(NOM, 1)
    outb = oracle.atmos(1, cz, ev, e.radii, e.rhos, e.yps, pb, 0.0)
    Pb = outb.reshape(-1, 3, 3)
    assert np.abs(Pb.sum(-1) - 1).max() < 1e-12
    # CPT-free check of the sign conventions: at zero density nu-bar with
    # kNuType = -1 is the complex conjugate amplitude, P_nubar(a->b) =
    # P_nu(b->a)
    z = np.zeros_like(e.rhos)
    pn = oracle.atmos(0, cz, ev, e.radii, z, e.yps, p, 0.0).reshape(-1, 3, 3)
    pa = oracle.atmos(1, cz, ev, e.radii, z, e.yps, pb, 0.0).reshape(-1, 3, 3)
    assert np.abs(pa - pn.transpose(0, 2, 1)).max() < 1e-12


def test_oracle_errors(oracle):
    e = W.prem4()
    with pytest.raises(RuntimeError, match="radii order"):
        oracle.atmos(0, [-0.5], [1.0], [0, 3000, 1000, 6371], [1, 2, 3, 4],
                     [0.5] * 4, osc_for(NOM, 0), 22.0)
    with pytest.raises(RuntimeError, match="must not be empty"):
        oracle.atmos(0, [-0.5], [1.0], [], [], [], osc_for(NOM, 0), 22.0)
    with pytest.raises(RuntimeError, match="could not open"):
        oracle.atmos_file(0, [-0.5], [1.0], "/nonexistent.dat",
                          osc_for(NOM, 0), 22.0)
    del e


# ------------------------------------------------------ polynomial densities

def poly_model():
    return W.prem4_poly()


@pytest.mark.parametrize("nutype", [0, 1])
def test_poly_golden(oracle, nutype):
    """setDensity(radii, a, b, c, yps) against the compiled reference's own
    polynomial path (tests/golden/g6_*), cos = -1 (straight through the
    centre) included."""
    g = golden(f"g6_prem4poly_{'nubar' if nutype else 'nu'}")
    out = oracle.atmos_poly(nutype, g["cosv"], g["ev"], g["radii"], g["a"],
                            g["b"], g["c"], g["yps"], osc_for(NOM, nutype),
                            22.0, threads=4)
    assert np.abs(out - g["prob"]).max() <= TOL_GOLD


def test_poly_file_and_reference(oracle, tmp_path):
    """The 5-column density file (radius, a, b, c, Ye) loads the same model;
    where the compiled reference is built, its file and array paths agree
    with the oracle's bitwise, also for drawn parameter sets."""
    ep = poly_model()
    path = tmp_path / "poly.dat"
    ep.write(str(path))
    cz, ev = W.cos_grid(24), W.log_grid(24, 0.1, 100.0)
    a = oracle.atmos_poly(0, cz, ev, ep.radii, ep.a, ep.b, ep.c, ep.yps,
                          osc_for(NOM, 0), 22.0)
    assert np.array_equal(oracle.atmos_file(0, cz, ev, str(path), osc_for(NOM, 0), 22.0), a)
    from oracle.loader import load_ref
    ref = load_ref()
    if ref is None:
        return
    assert np.array_equal(ref.atmos_file(0, cz, ev, str(path), osc_for(NOM, 0), 22.0), a)
    for p in W.draw_param_sets(2, 606):
        row = [p.th12, p.th13, p.th23, p.dcp, p.dm12, p.dm23]
        for nt in (0, 1):
            x = oracle.atmos_poly(nt, cz, ev, ep.radii, ep.a, ep.b, ep.c,
                                  ep.yps, osc_for(row, nt), 22.0)
            y = ref.atmos_poly(nt, cz, ev, ep.radii, ep.a, ep.b, ep.c,
                               ep.yps, osc_for(row, nt), 22.0)
            assert np.array_equal(x, y)


def test_poly_average_is_the_segment_mean(oracle):
    """Independently of the reference's closed forms: each (cosine, layer)
    density equals the mean of rho(R(t)) * Ye over the layer's segment by
    64-point Gauss-Legendre quadrature along the chord (R(t)^2 = t^2 - 2 d t
    + R_i^2 from the shell's outer boundary, d its half chord), and the
    radial mean over the shell at cos = -1."""
    ep = poly_model()
    desc = lambda v: np.asarray(v)[::-1]  # noqa: E731  rows as the model holds them
    radii, A, B, Cc, ye = map(desc, (ep.radii, ep.a, ep.b, ep.c, ep.yps))
    cz = np.array([-1.0, -0.999, -0.97, -0.9, -0.75, -0.6, -0.45, -0.3, -0.1, -0.02])
    got = oracle.poly_rye(cz, ep.radii, ep.a, ep.b, ep.c, ep.yps, 22.0)
    x, w = np.polynomial.legendre.leggauss(64)
    R = W.R_EARTH_KM
    for ic, c in enumerate(cz):
        M, segs, _ = oracle.segments(c, ep.radii, ep.c, ep.yps, 22.0)
        rmin = R * np.sqrt(1 - c * c)
        for l in range(1, M + 1):
            i = l - 1
            if c == -1.0:
                lo, hi = radii[i + 1] if i + 1 < len(radii) else 0.0, radii[i]
                r = 0.5 * (hi - lo) * x + 0.5 * (hi + lo)
                mean = 0.5 * np.sum(w * (A[i] * r * r + B[i] * r + Cc[i]))
            else:
                t2 = segs[l]
                d = np.sqrt(radii[i] ** 2 - rmin ** 2)
                t = 0.5 * t2 * (x + 1)
                r = np.sqrt(np.maximum(t * t - 2 * d * t + radii[i] ** 2, 0.0))
                mean = 0.5 * np.sum(w * (A[i] * r * r + B[i] * r + Cc[i]))
            assert abs(got[ic, i] - mean * ye[i]) <= 1e-9 * abs(mean * ye[i]), (c, l)
        assert np.all(got[ic, M:] == 0.0)


def test_poly_constant_profile_equals_constant_model(oracle):
    """a = b = 0 reduces the polynomial path to the constant-density one."""
    e = W.prem4()
    z = np.zeros_like(e.rhos)
    cz, ev = W.cos_grid(24), W.log_grid(24, 0.1, 100.0)
    for nt in (0, 1):
        x = oracle.atmos_poly(nt, cz, ev, e.radii, z, z, e.rhos, e.yps,
                              osc_for(NOM, nt), 22.0)
        y = oracle.atmos(nt, cz, ev, e.radii, e.rhos, e.yps, osc_for(NOM, nt), 22.0)
        assert np.abs(x - y).max() < 1e-13


# ------------------------------------------- production-height averaging

def _bin_prob(nC, nE, nb, seed):
    """Seeded bin probabilities [2][3][nE][nC][nb], normalised per
    (type, flavour, E, cos)."""
    p = np.random.default_rng(seed).random((2, 3, nE, nC, nb))
    return p / p.sum(axis=-1, keepdims=True)


def test_height_avg_golden(oracle):
    """The reference's averaging arithmetic (tests/golden/g7_*, made through
    the compiled reference with its averaging flag set) reproduced with the
    compiled reference's sinc, whose abs() truncates to an integer."""
    g = golden("g7_prem4_height_avg")
    for nt in (0, 1):
        out = oracle.atmos_pha(nt, g["cosv"], g["ev"], g["radii"], g["rhos"],
                               g["yps"], osc_for(NOM, nt), 22.0,
                               g["height_prob"], g["edges_km"],
                               int_abs_sinc=True)
        assert np.abs(out - g["prob"][nt]).max() <= TOL_GOLD
        # the intended sinc moves results by at most a few 1e-5 here
        fixed = oracle.atmos_pha(nt, g["cosv"], g["ev"], g["radii"], g["rhos"],
                                 g["yps"], osc_for(NOM, nt), 22.0,
                                 g["height_prob"], g["edges_km"])
        assert 0 < np.abs(fixed - g["prob"][nt]).max() < 1e-4


def test_height_avg_reference_drawn_sets(oracle):
    """Bitwise against the compiled reference for drawn parameter sets and
    bin layouts, where it is built."""
    from oracle.loader import load_ref
    ref = load_ref()
    if ref is None:
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    e = W.prem4()
    cz = np.linspace(-1.0, 1.0, 11)
    ev = W.log_grid(9, 0.1, 50.0)
    for k, p in enumerate(W.draw_param_sets(3, 997)):
        row = [p.th12, p.th13, p.th23, p.dcp, p.dm12, p.dm23]
        nb = 1 + 3 * k
        edges = np.sort(np.random.default_rng(k).uniform(1.0, 60.0, nb + 1))
        prob = _bin_prob(len(cz), len(ev), nb, 50 + k)
        for nt in (0, 1):
            a = ref.atmos_pha(nt, cz, ev, e.radii, e.rhos, e.yps, osc_for(row, nt),
                              15.0, prob, edges)
            b = oracle.atmos_pha(nt, cz, ev, e.radii, e.rhos, e.yps,
                                 osc_for(row, nt), 15.0, prob, edges,
                                 int_abs_sinc=True)
            assert np.array_equal(a, b)


def test_height_avg_zero_width_bin_is_that_height(oracle):
    """One bin of zero width at height h' with probability 1 moves the
    production point: the result is the plain propagation from h'."""
    e = W.prem4()
    cz = np.linspace(-1.0, 1.0, 13)
    ev = W.log_grid(11, 0.1, 30.0)
    for h2 in (3.0, 22.0, 61.0):
        prob = np.ones((2, 3, len(ev), len(cz), 1))
        for nt in (0, 1):
            avg = oracle.atmos_pha(nt, cz, ev, e.radii, e.rhos, e.yps,
                                   osc_for(NOM, nt), 22.0, prob, [h2, h2])
            plain = oracle.atmos(nt, cz, ev, e.radii, e.rhos, e.yps,
                                 osc_for(NOM, nt), h2)
            assert np.abs(avg - plain).max() < 1e-12


def _height_at_shift(cz, h_km, s_km):
    """Production height whose path length exceeds that of h_km by s_km
    (inverse of the path length of physics.hpp:897-898)."""
    R = 6371.0
    P0 = np.sqrt((R + h_km) ** 2 - R * R * (1 - cz * cz)) - R * cz
    P = P0 + s_km
    return np.sqrt((P + R * cz) ** 2 + R * R * (1 - cz * cz)) - R


def test_height_avg_bin_is_the_mean_over_path_shifts(oracle):
    """A bin averages the propagation uniformly over the path-length shifts
    between its edges: one bin with probability 1 equals the 96-point
    Gauss-Legendre mean of plain propagations from the heights of those
    shifts.  The grid has sinc arguments between 1e-6 and 1, where the
    compiled reference's truncating abs() takes the series instead."""
    e = W.prem4()
    cz = np.array([-0.9, -0.5, -0.1, 0.0, 0.2, 0.7])
    ev = np.array([0.15, 0.6, 2.0, 8.0, 40.0])
    h, edges = 22.0, np.array([12.0, 35.0])
    x, wq = np.polynomial.legendre.leggauss(96)
    prob = np.ones((2, 3, len(ev), len(cz), 1))
    for nt in (0, 1):
        osc = osc_for(NOM, nt)
        avg = oracle.atmos_pha(nt, cz, ev, e.radii, e.rhos, e.yps, osc, h,
                               prob, edges)
        quirk = oracle.atmos_pha(nt, cz, ev, e.radii, e.rhos, e.yps, osc, h,
                                 prob, edges, int_abs_sinc=True)
        R = 6371.0
        for i, c in enumerate(cz):
            path = lambda hk: np.sqrt((R + hk) ** 2 - R * R * (1 - c * c)) - R * c
            s0, s1 = path(edges[0]) - path(h), path(edges[1]) - path(h)
            mean = np.zeros((len(ev), 9))
            for xi, wi in zip(x, wq):
                s = 0.5 * (s0 + s1) + 0.5 * (s1 - s0) * xi
                hk = _height_at_shift(c, h, s)
                mean += 0.5 * wi * oracle.atmos(nt, [c], ev, e.radii, e.rhos,
                                                e.yps, osc, hk)[0]
            assert np.abs(avg[i] - mean).max() < 1e-9
        assert np.abs(quirk - avg).max() > 1e-6


def test_height_avg_no_bins_is_incoherent(oracle):
    """With averaging on and no bins the interference terms vanish: above
    the horizon (vacuum only) P(a->b) = sum_k |U_ak|^2 |U_bk|^2; everywhere
    each initial flavour's probabilities still sum to 1."""
    e = W.prem4()
    cz = np.linspace(-1.0, 1.0, 9)
    ev = W.log_grid(7, 0.1, 30.0)
    prob = np.zeros((2, 3, len(ev), len(cz), 0))
    for nt in (0, 1):
        osc = osc_for(NOM, nt)
        out = oracle.atmos_pha(nt, cz, ev, e.radii, e.rhos, e.yps, osc, 22.0,
                               prob, [22.0])
        U = pdg_u(osc.th12, osc.th13, osc.th23, osc.dcp * osc.knutype)
        inc = (np.abs(U) ** 2) @ (np.abs(U) ** 2).T  # [a][b]
        for i, c in enumerate(cz):
            if c >= 0:
                assert np.abs(out[i] - inc.ravel()).max() < 1e-12
        assert np.abs(out.reshape(len(cz), len(ev), 3, 3).sum(-1) - 1).max() < 1e-12


def test_height_avg_weights(oracle):
    """Results are affine in the bin probabilities (normalised mixtures mix
    the results), the probabilities of one initial flavour or neutrino type
    touch only that flavour's channels or that type's run, and each initial
    flavour's probabilities sum to 1."""
    e = W.prem4()
    cz = np.linspace(-1.0, 0.9, 8)
    ev = W.log_grid(6, 0.1, 20.0)
    edges = np.array([4.0, 9.0, 17.0, 26.0, 44.0])
    p1 = _bin_prob(len(cz), len(ev), 4, 1)
    p2 = _bin_prob(len(cz), len(ev), 4, 2)
    osc = osc_for(NOM, 0)
    run = lambda p, nt=0: oracle.atmos_pha(nt, cz, ev, e.radii, e.rhos, e.yps,
                                           osc_for(NOM, nt), 22.0, p, edges)
    a, b, m = run(p1), run(p2), run(0.3 * p1 + 0.7 * p2)
    assert np.abs(m - (0.3 * a + 0.7 * b)).max() < 1e-13
    assert np.abs(a.reshape(len(cz), len(ev), 3, 3).sum(-1) - 1).max() < 1e-12
    q = p1.copy()
    q[0, 1] = p2[0, 1]  # neutrino, initial mu
    c = run(q)
    mu = [3, 4, 5]
    others = [0, 1, 2, 6, 7, 8]
    assert np.array_equal(c[..., others], a[..., others])
    assert np.abs(c[..., mu] - a[..., mu]).max() > 1e-6
    q = p1.copy()
    q[1] = p2[1]  # the antineutrino slice
    assert np.array_equal(run(q), a)
    assert not np.array_equal(run(q, 1), run(p1, 1))
    del osc


def test_height_avg_poly_reference(oracle):
    """Averaging through a polynomial model, bitwise against the compiled
    reference where it is built (grazing cosines, where both return NaN,
    left out)."""
    from oracle.loader import load_ref
    ref = load_ref()
    if ref is None:
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    ep = W.prem4_poly()
    cz = np.concatenate([W.cos_grid(12), np.linspace(0.0, 0.9, 4)])
    ev = W.log_grid(10, 0.1, 50.0)
    edges = np.array([3.0, 11.0, 22.0, 47.0])
    prob = _bin_prob(len(cz), len(ev), 3, 71)
    for nt in (0, 1):
        a = ref.atmos_poly_pha(nt, cz, ev, ep.radii, ep.a, ep.b, ep.c, ep.yps,
                               osc_for(NOM, nt), 22.0, prob, edges)
        b = oracle.atmos_poly_pha(nt, cz, ev, ep.radii, ep.a, ep.b, ep.c, ep.yps,
                                  osc_for(NOM, nt), 22.0, prob, edges,
                                  int_abs_sinc=True)
        assert np.array_equal(a, b)
