"""The kernel arithmetic (cp3_math.hpp) run on the CPU through its host build
(libcp3model.so) against the oracle: the exact rewrites of the Barger method
keep FP64 within 1e-10 of the reference path, the mixed form within 1e-5; and
the operation counts used for the roofline are stable."""
import ctypes as C

import numpy as np
import pytest

from conftest import golden, osc_for
from cudaprob3_amd import workloads as W
from cudaprob3_amd import flops as FL

dp = C.POINTER(C.c_double)


@pytest.fixture(scope="module")
def model():
    from cudaprob3_amd._native import MODEL_SO
    m = C.CDLL(MODEL_SO)
    m.cp3m_atmos.argtypes = ([C.c_int, C.c_int, dp, C.c_int, dp, C.c_int, dp,
                              dp, dp, C.c_int] + [C.c_double] * 4 +
                             [C.c_int] + [C.c_double] * 3 +
                             [dp, C.c_char_p, C.c_int])
    m.cp3m_atmos_poly.argtypes = ([C.c_int, C.c_int, dp, C.c_int, dp, C.c_int] +
                                  [dp] * 5 + [C.c_int] + [C.c_double] * 4 +
                                  [C.c_int] + [C.c_double] * 3 +
                                  [dp, C.c_char_p, C.c_int])
    m.cp3m_poly_rye.argtypes = ([dp, C.c_int] + [dp] * 5 +
                                [C.c_int, C.c_double, dp])
    m.cp3m_atmos_pha.argtypes = ([C.c_int, C.c_int, dp, C.c_int, dp, C.c_int, dp,
                                  dp, dp, C.c_int] + [C.c_double] * 4 +
                                 [C.c_int] + [C.c_double] * 3 +
                                 [C.c_int, dp, dp, dp, C.c_char_p, C.c_int])
    m.cp3m_beam.argtypes = ([C.c_int, C.c_int, dp, C.c_int, C.c_double,
                             C.c_double] + [C.c_double] * 4 + [C.c_int] +
                            [C.c_double] * 2 + [dp, C.c_char_p, C.c_int])
    return m


def P(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(dp)


def m_atmos(m, prec, nt, cz, ev, earth, row, h=22.0):
    cz = np.ascontiguousarray(cz, float)
    ev = np.ascontiguousarray(ev, float)
    out = np.zeros((len(cz), len(ev), 9))
    err = C.create_string_buffer(256)
    t12, t13, t23, d, m12, m23 = [float(x) for x in row]
    rc = m.cp3m_atmos(prec, nt, P(cz), len(cz), P(ev), len(ev),
                      P(earth.radii), P(earth.rhos), P(earth.yps),
                      len(earth.radii), t12, t13, t23, d, -1 if nt else 1,
                      m12, m23, h, out.ctypes.data_as(dp), err, 256)
    assert rc == 0, err.value
    return out


NOM = [W.TH12, W.TH13, W.TH23, W.DCP, W.DM12, W.DM23]


@pytest.mark.parametrize("nutype", [0, 1])
def test_fp64_rewrite_vs_goldens(model, nutype):
    g = golden(f"g1_shells12_{'nubar' if nutype else 'nu'}")
    e = W.EarthModel(g["radii"], g["rhos"], g["yps"])
    out = m_atmos(model, 0, nutype, g["cosv"], g["ev"], e, NOM)
    assert np.abs(out - g["prob"]).max() < 1e-10


def test_fp64_rewrite_edges(model):
    g = golden("g4_edges")
    e = W.EarthModel(g["radii"], g["rhos"], g["yps"])
    for key in ["nominal", "dm12_zero", "dm23_zero", "dcp_zero", "inverted"]:
        for nt in (0, 1):
            out = m_atmos(model, 0, nt, g["cosv"], g["ev"], e,
                          g[f"params_{key}"])
            assert np.abs(out - g[f"prob_{key}_{nt}"]).max() < 1e-10, key


def test_mixed_rewrite_59_shells(model, oracle):
    e = W.earth_59()
    cz, ev = W.cos_grid(96), W.log_grid(96, 0.5, 50.0)
    for p in W.draw_param_sets(2, 77):
        row = [p.th12, p.th13, p.th23, p.dcp, p.dm12, p.dm23]
        for nt in (0, 1):
            ref = oracle.atmos(nt, cz, ev, e.radii, e.rhos, e.yps,
                               osc_for(row, nt), 22.0, threads=4)
            assert np.abs(m_atmos(model, 2, nt, cz, ev, e, row) - ref).max() < 1e-5
            assert np.abs(m_atmos(model, 0, nt, cz, ev, e, row) - ref).max() < 1e-10


def test_beam_rewrite(model, oracle):
    g = golden("g2_beam")
    for nt in (0, 1):
        for i, par in enumerate(g["params"]):
            out = np.zeros((len(g["ev"]), 9))
            err = C.create_string_buffer(256)
            t12, t13, t23, d, m12, m23 = [float(x) for x in par]
            rc = model.cp3m_beam(0, nt, P(g["ev"]), len(g["ev"]), 2.6, 1300.0,
                                 t12, t13, t23, d, -1 if nt else 1, m12, m23,
                                 out.ctypes.data_as(dp), err, 256)
            assert rc == 0
            assert np.abs(out - g["prob"][nt, i]).max() < 1e-10


def test_geometry_table_matches_oracle(oracle):
    e = W.earth_59()
    cz = W.cos_grid(200)
    M = FL.max_layers(cz, e)
    for i in range(0, 200, 17):
        Mo, _, _ = oracle.segments(cz[i], e.radii, e.rhos, e.yps, 22.0)
        assert M[i] == Mo


def test_stage_counts():
    f64 = FL.stage_ops("fp64")
    mixed = FL.stage_ops("mixed")
    # a middle layer: two rank-one updates of two complex 3-vectors each
    # (6 + 4 + 6 multiply-adds = 28 FLOPs per update) in F, and per phase
    # q = om L and q - rint(q) in D; the determinant phase: FP64 adds each
    # reduced phase (2 per layer), mixed advances by (oma + omb) L (3)
    assert mixed["middle"]["f32"] == 4 * 28
    assert mixed["middle"]["f64"] == 2 * 2 + 3
    assert f64["middle"]["f64"] == 4 * 28 + 2 * 3
    assert f64["middle"]["sincos64"] == 2 and mixed["middle"]["sincos32"] == 2
    # the cubic's third angle: an FP32 acos and sincos seed (one FP32
    # multiply), then FP64 Newton steps for the cube root of e^{i acos(x)}
    for form in (f64, mixed):
        assert form["solve"]["acos64"] == 0 and form["solve"]["sincos64"] == 0
        assert form["solve"]["acos32"] == 1 and form["solve"]["sincos32"] == 1
    assert mixed["solve"]["f32"] == 1  # the eigen-solve's arithmetic stays FP64
    # ... but for the beam's F eigenvectors, normalised in F: its FP64 work is
    # below one solve plus its two phase reductions (q = om L, q - rint(q))
    assert mixed["beam"]["f64"] < mixed["solve"]["f64"] + 2 * 2
    # the hoisting: fewer solves per trajectory with larger cosine blocks
    wl = W.atmos1024(n_sets=4)
    c1 = FL.batch_counts(wl, "mixed", 1)
    c8 = FL.batch_counts(wl, "mixed", 8)
    assert c8["acos32"] < c1["acos32"] / 6
    # the layer arithmetic does not depend on the blocking (one FP32 seed
    # multiply per solve aside)
    assert c8["f32"] - c8["acos32"] == c1["f32"] - c1["acos32"]


def test_third_angle():
    """cos and sin of acos(x)/3 (physics.hpp:310), which the FP64 forms take
    as the cube root of x + i sqrt(1 - x^2) by Newton's method from an FP32
    seed: against 40-digit values over [-1, 1], including both ends and
    points within 2^-52 of them, where acos is ill-conditioned; FP32 (acos and
    sincos) within its own rounding."""
    import mpmath
    from cudaprob3_amd._native import MODEL_SO
    m = C.CDLL(MODEL_SO)
    m.cp3m_third_angle.argtypes = [C.c_int, dp, C.c_int, dp, dp]
    near = np.array([2.0 ** -k for k in range(1, 53)])
    x = np.concatenate([np.linspace(-1.0, 1.0, 4001), [0.0, -1.0, 1.0],
                        1.0 - near, -1.0 + near])
    mpmath.mp.dps = 40

    def ref(xs):
        th = [mpmath.acos(mpmath.mpf(float(v))) / 3 for v in xs]
        return (np.array([float(mpmath.cos(t)) for t in th]),
                np.array([float(mpmath.sin(t)) for t in th]))

    # the FP32 form sees x rounded to float, so its reference is taken there
    for prec, xr, tol in ((0, x, 4e-16), (1, x.astype(np.float32).astype(np.float64), 2e-6)):
        c_ref, s_ref = ref(xr)
        c, s = np.zeros_like(x), np.zeros_like(x)
        m.cp3m_third_angle(prec, P(x), len(x), c.ctypes.data_as(dp), s.ctypes.data_as(dp))
        assert np.abs(c - c_ref).max() < tol, prec
        assert np.abs(s - s_ref).max() < tol, prec
    # the ends exactly: theta = 0 at x = 1 and pi/3 at x = -1
    c, s = np.zeros(2), np.zeros(2)
    m.cp3m_third_angle(0, P(np.array([1.0, -1.0])), 2, c.ctypes.data_as(dp),
                       s.ctypes.data_as(dp))
    assert c[0] == 1.0 and s[0] == 0.0
    assert abs(c[1] - 0.5) < 2e-16 and abs(s[1] - np.sqrt(3) / 2) < 2e-16


def test_phase_polynomial():
    """m_cispi_m1(f) = e^{2 pi i f} - 1 on |f| <= 1/2 against numpy's sin:
    a few ulp of |w| <= 2 everywhere, exact zero at f = 0, w = -2 at the half
    revolution, and relative accuracy for tiny f (no cancellation)."""
    from cudaprob3_amd._native import MODEL_SO
    m = C.CDLL(MODEL_SO)
    m.cp3m_cispi_m1.argtypes = [C.c_int, dp, C.c_int, dp, dp]
    grid = np.linspace(-0.5, 0.5, 100001)
    tiny = np.array([1e-15, -1e-9, 3e-7, -2e-12])  # squares stay normal
    f0 = np.concatenate([grid, [0.0], tiny])
    for prec, tol in ((0, 2e-15), (1, 8e-7)):
        f = f0 if prec == 0 else f0.astype(np.float32).astype(np.float64)
        wr, wi = np.zeros_like(f), np.zeros_like(f)
        m.cp3m_cispi_m1(prec, P(f), len(f), wr.ctypes.data_as(dp),
                        wi.ctypes.data_as(dp))
        assert np.abs(wr + 2 * np.sin(np.pi * f) ** 2).max() < tol
        assert np.abs(wi - np.sin(2 * np.pi * f)).max() < tol
        n = len(grid)
        assert wr[n] == 0.0 and wi[n] == 0.0
        assert abs(wr[0] + 2) < tol and abs(wi[0]) < tol
        t = f[n + 1:]
        assert np.all(np.abs(wi[n + 1:] / (2 * np.pi * t) - 1) < 1e-6)
        assert np.all(np.abs(wr[n + 1:] / (-2 * (np.pi * t) ** 2) - 1) < 1e-5)


def m_atmos_poly(m, prec, nt, cz, ev, ep, row, h=22.0):
    cz = np.ascontiguousarray(cz, float)
    ev = np.ascontiguousarray(ev, float)
    out = np.zeros((len(cz), len(ev), 9))
    err = C.create_string_buffer(256)
    t12, t13, t23, d, m12, m23 = [float(x) for x in row]
    rc = m.cp3m_atmos_poly(prec, nt, P(cz), len(cz), P(ev), len(ev),
                           P(ep.radii), P(ep.a), P(ep.b), P(ep.c), P(ep.yps),
                           len(ep.radii), t12, t13, t23, d, -1 if nt else 1,
                           m12, m23, h, out.ctypes.data_as(dp), err, 256)
    assert rc == 0, err.value
    return out


def test_poly_trajectory_densities_match_oracle(model, oracle):
    """The product's per-(cosine, layer) densities of a polynomial model are
    the oracle's (pinned to the compiled reference) bit for bit."""
    ep = W.prem4_poly()
    cz = np.concatenate([W.cos_grid(64), [-0.999999, -1e-6]])
    out = np.zeros((len(cz), len(ep.radii)))
    assert model.cp3m_poly_rye(P(cz), len(cz), P(ep.radii), P(ep.a), P(ep.b),
                               P(ep.c), P(ep.yps), len(ep.radii), 22.0,
                               out.ctypes.data_as(dp)) == 0
    ref = oracle.poly_rye(cz, ep.radii, ep.a, ep.b, ep.c, ep.yps, 22.0)
    assert np.array_equal(out, ref)


def test_poly_grazing_trajectories_are_finite(model, oracle):
    """Just below the horizon the reference's segment average of a
    polynomial model reaches 0/0 (its density, and every probability, is NaN;
    the oracle restates that).  The product takes the limit of the mean over
    a vanishing segment, the density at the shell's outer radius, and its
    probabilities stay unitary."""
    ep = W.prem4_poly()
    cz = np.array([-1.1102230246251565e-16, -1e-12, -1e-9])
    out = np.zeros((len(cz), len(ep.radii)))
    assert model.cp3m_poly_rye(P(cz), len(cz), P(ep.radii), P(ep.a), P(ep.b),
                               P(ep.c), P(ep.yps), len(ep.radii), 22.0,
                               out.ctypes.data_as(dp)) == 0
    ref = oracle.poly_rye(cz, ep.radii, ep.a, ep.b, ep.c, ep.yps, 22.0)
    assert np.isnan(ref[:, 0]).all()
    R = ep.radii[-1]
    rho_outer = (ep.a[-1] * R * R + ep.b[-1] * R + ep.c[-1]) * ep.yps[-1]
    assert np.allclose(out[:, 0], rho_outer, rtol=1e-14, atol=0)
    assert np.array_equal(out[:, 1:], ref[:, 1:])
    probs = m_atmos_poly(model, 0, 0, cz, W.log_grid(8, 0.1, 50.0), ep, NOM)
    assert np.isfinite(probs).all()
    assert np.abs(probs.reshape(-1, 3, 3).sum(-1) - 1).max() < 1e-12


@pytest.mark.parametrize("nutype", [0, 1])
def test_poly_rewrite_vs_golden_and_oracle(model, oracle, nutype):
    """The kernel arithmetic through a polynomial model: FP64 within 1e-10 of
    the compiled reference's golden and of the oracle, mixed within 1e-5."""
    g = golden(f"g6_prem4poly_{'nubar' if nutype else 'nu'}")
    ep = W.prem4_poly()
    assert np.array_equal(g["a"], ep.a) and np.array_equal(g["c"], ep.c)
    out = m_atmos_poly(model, 0, nutype, g["cosv"], g["ev"], ep, NOM)
    assert np.abs(out - g["prob"]).max() < 1e-10
    cz, ev = W.cos_grid(60), W.log_grid(60, 0.1, 100.0)
    ref = oracle.atmos_poly(nutype, cz, ev, ep.radii, ep.a, ep.b, ep.c, ep.yps,
                            osc_for(NOM, nutype), 22.0, threads=4)
    assert np.abs(m_atmos_poly(model, 0, nutype, cz, ev, ep, NOM) - ref).max() < 1e-10
    assert np.abs(m_atmos_poly(model, 2, nutype, cz, ev, ep, NOM) - ref).max() < 1e-5


def test_fp32_form_vs_reference_float_build(model, oracle):
    """The FP32 form is at least as accurate as the reference's own
    AtmosCpuPropagator<float> against FP64 (where oracle/_ref is built); the
    mixed form is far inside both."""
    from oracle.loader import load_ref
    ref = load_ref(deep=True)
    if ref is None:
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    e = W.earth_12()
    cz, ev = W.cos_grid(80), W.log_grid(80, 0.1, 100.0)
    for nt in (0, 1):
        exact = oracle.atmos(nt, cz, ev, e.radii, e.rhos, e.yps, osc_for(NOM, nt),
                             22.0, threads=4)
        ref32 = np.abs(ref.atmos(nt, cz, ev, e.radii, e.rhos, e.yps, osc_for(NOM, nt),
                                 22.0, threads=4, f32=True) - exact).max()
        ours32 = np.abs(m_atmos(model, 1, nt, cz, ev, e, NOM) - exact).max()
        mixed = np.abs(m_atmos(model, 2, nt, cz, ev, e, NOM) - exact).max()
        print(f"nutype {nt}: reference float {ref32:.2e}, FP32 form {ours32:.2e}, "
              f"mixed {mixed:.2e}")
        assert ours32 <= ref32
        assert mixed < 1e-5


def _load_file(model, path):
    model.cp3m_load_earth_file.argtypes = [C.c_char_p, dp, C.c_int,
                                           C.POINTER(C.c_int), C.POINTER(C.c_int),
                                           C.c_char_p, C.c_int]
    rows = np.zeros((64, 5))
    n, poly = C.c_int(), C.c_int()
    err = C.create_string_buffer(256)
    rc = model.cp3m_load_earth_file(str(path).encode(), rows.ctypes.data_as(dp), 64,
                                    C.byref(n), C.byref(poly), err, 256)
    return rc, rows[:n.value], poly.value, err.value.decode()


def test_density_file_formats(model, tmp_path):
    """setDensityFromFile's parsing (atmoscpupropagator.hpp:250-344) in the
    product's host code: comments and repeated spaces, the 3-column and the
    5-column (polynomial) formats, rows stored descending; 4 columns are
    refused with a message instead of ending the process."""
    ep = W.prem4_poly()
    p5 = tmp_path / "poly.dat"
    with open(p5, "w") as f:
        f.write("# radius a b c Ye\n")
        for row in zip(ep.radii, ep.a, ep.b, ep.c, ep.yps):
            f.write("   ".join(repr(float(x)) for x in row) + "\n")
    rc, rows, poly, _ = _load_file(model, p5)
    assert rc == 0 and poly == 1
    expect = np.stack([ep.radii, ep.a, ep.b, ep.c, ep.yps], axis=1)[::-1]
    assert np.array_equal(rows, expect)
    e = W.prem4()
    p3 = tmp_path / "const.dat"
    with open(p3, "w") as f:
        for r, d, y in zip(e.radii, e.rhos, e.yps):
            f.write(f"{float(r)!r}  {float(d)!r} {float(y)!r}\n")
    rc, rows, poly, _ = _load_file(model, p3)
    assert rc == 0 and poly == 0
    assert np.array_equal(rows[:, [0, 1, 4]], np.stack([e.radii, e.rhos, e.yps], 1)[::-1])
    p4 = tmp_path / "bad.dat"
    p4.write_text("0 1 2 3\n6371 1 2 3\n")
    rc, _, _, msg = _load_file(model, p4)
    assert rc == -1 and "4 entries per line" in msg


# ------------------------------------------- production-height averaging

def m_atmos_pha(m, prec, nt, cz, ev, earth, row, h, prob, edges):
    cz = np.ascontiguousarray(cz, float)
    ev = np.ascontiguousarray(ev, float)
    prob = np.ascontiguousarray(prob, float)
    edges = np.ascontiguousarray(edges, float)
    out = np.zeros((len(cz), len(ev), 9))
    err = C.create_string_buffer(256)
    t12, t13, t23, d, m12, m23 = [float(x) for x in row]
    rc = m.cp3m_atmos_pha(prec, nt, P(cz), len(cz), P(ev), len(ev),
                          P(earth.radii), P(earth.rhos), P(earth.yps),
                          len(earth.radii), t12, t13, t23, d, -1 if nt else 1,
                          m12, m23, h, len(edges) - 1, P(prob), P(edges),
                          out.ctypes.data_as(dp), err, 256)
    assert rc == 0, err.value
    return out


@pytest.mark.parametrize("nutype", [0, 1])
def test_height_avg_rewrite_vs_oracle(model, oracle, nutype):
    """Production-height averaging through the kernel arithmetic (the
    amplitude split by vacuum eigenstate in the rotated frame) against the
    oracle's restatement of the reference's expansion: FP64 within 1e-10,
    mixed within 1e-5, FP32 within 1e-3; on the golden's inputs too, where
    the difference from the golden is the reference CPU build's sinc."""
    e = W.earth_12()
    cz = np.linspace(-1.0, 1.0, 41)
    ev = W.log_grid(37, 0.1, 100.0)
    edges = np.array([0.5, 4.0, 9.0, 15.0, 22.0, 30.0, 41.0, 60.0])
    pr = np.random.default_rng(5).random((2, 3, len(ev), len(cz), len(edges) - 1))
    pr /= pr.sum(axis=-1, keepdims=True)
    for k, row in enumerate([NOM] + [[p.th12, p.th13, p.th23, p.dcp, p.dm12, p.dm23]
                                     for p in W.draw_param_sets(2, 31)]):
        ref = oracle.atmos_pha(nutype, cz, ev, e.radii, e.rhos, e.yps,
                               osc_for(row, nutype), 18.0, pr, edges, threads=4)
        out = m_atmos_pha(model, 0, nutype, cz, ev, e, row, 18.0, pr, edges)
        assert np.abs(out - ref).max() < 1e-10, k
        mixed = m_atmos_pha(model, 2, nutype, cz, ev, e, row, 18.0, pr, edges)
        assert np.abs(mixed - ref).max() < 1e-5, k
        f32 = m_atmos_pha(model, 1, nutype, cz, ev, e, row, 18.0, pr, edges)
        assert np.abs(f32 - ref).max() < 1e-3, k
    g = golden("g7_prem4_height_avg")
    e4 = W.EarthModel(g["radii"], g["rhos"], g["yps"])
    out = m_atmos_pha(model, 0, nutype, g["cosv"], g["ev"], e4, NOM, 22.0,
                      g["height_prob"], g["edges_km"])
    fixed = oracle.atmos_pha(nutype, g["cosv"], g["ev"], g["radii"], g["rhos"],
                             g["yps"], osc_for(NOM, nutype), 22.0,
                             g["height_prob"], g["edges_km"])
    assert np.abs(out - fixed).max() < 1e-10
    assert np.abs(out - g["prob"][nutype]).max() < 1e-4


def test_height_avg_limits_in_the_rewrite(model, oracle):
    """The split reassembles the plain amplitude: a zero-width bin at the
    production height itself gives the unaveraged result, no bins give the
    incoherent sum, and a zero-length atmosphere segment (cos = -1 at height
    0) turns no phase."""
    e = W.prem4()
    cz = np.linspace(-1.0, 1.0, 17)
    ev = W.log_grid(13, 0.1, 50.0)
    one = np.ones((2, 3, len(ev), len(cz), 1))
    for nt in (0, 1):
        plain = m_atmos(model, 0, nt, cz, ev, e, NOM, h=22.0)
        avg = m_atmos_pha(model, 0, nt, cz, ev, e, NOM, 22.0, one, [22.0, 22.0])
        assert np.abs(avg - plain).max() < 1e-12
        none = np.zeros((2, 3, len(ev), len(cz), 0))
        inc = m_atmos_pha(model, 0, nt, cz, ev, e, NOM, 22.0, none, [22.0])
        ref = oracle.atmos_pha(nt, cz, ev, e.radii, e.rhos, e.yps, osc_for(NOM, nt),
                               22.0, none, [22.0])
        assert np.abs(inc - ref).max() < 1e-12
        z = m_atmos_pha(model, 0, nt, [-1.0], ev, e, NOM, 0.0,
                        np.ones((2, 3, len(ev), 1, 1)), [0.0, 30.0])
        zr = oracle.atmos_pha(nt, [-1.0], ev, e.radii, e.rhos, e.yps,
                              osc_for(NOM, nt), 0.0, np.ones((2, 3, len(ev), 1, 1)),
                              [0.0, 30.0])
        assert np.abs(z - zr).max() < 1e-10


def test_height_avg_wide_bins(model, oracle):
    """Bins wide enough (down to 20 MeV, 80-300 km) that a pair's phase
    across a bin exceeds half a revolution, where the sinc divides the
    reduced sine instead of using the division-free quotient."""
    e = W.earth_12()
    cz = np.linspace(-1.0, 1.0, 41)
    ev = W.log_grid(37, 0.02, 100.0)
    edges = np.array([0.0, 80.0, 300.0])
    pr = np.random.default_rng(6).random((2, 3, len(ev), len(cz), 2))
    pr /= pr.sum(axis=-1, keepdims=True)
    for nt in (0, 1):
        ref = oracle.atmos_pha(nt, cz, ev, e.radii, e.rhos, e.yps, osc_for(NOM, nt),
                               18.0, pr, edges)
        assert np.abs(m_atmos_pha(model, 0, nt, cz, ev, e, NOM, 18.0, pr, edges)
                      - ref).max() < 1e-10
        assert np.abs(m_atmos_pha(model, 2, nt, cz, ev, e, NOM, 18.0, pr, edges)
                      - ref).max() < 1e-5


def test_extreme_energies_host_build(model, oracle):
    """The host build of the kernel arithmetic from 1 MeV to 10 TeV and for
    cosines above the horizon, against the oracle: FP64 within 1e-10, mixed
    within 1e-5 (layer phases of 1e4-1e5 radians at 1 MeV)."""
    e = W.earth_12()
    cz = np.concatenate([np.linspace(-1.0, -1e-4, 20), np.linspace(0.0, 0.99, 4)])
    ev = W.log_grid(30, 1e-3, 1e4)
    for nt in (0, 1):
        ref = oracle.atmos(nt, cz, ev, e.radii, e.rhos, e.yps, osc_for(NOM, nt), 22.0,
                           threads=4)
        assert np.abs(m_atmos(model, 0, nt, cz, ev, e, NOM) - ref).max() < 1e-10
        assert np.abs(m_atmos(model, 2, nt, cz, ev, e, NOM) - ref).max() < 1e-5
