"""Writes the golden vectors under tests/golden/ from the reference's own CPU
propagators (oracle/_ref/libcp3ref*.so, compiled from the reference headers by
oracle/Makefile).  Run in a container that has the reference sources:

    make -C oracle && python tests/golden/make_golden.py [--only-poly | --only-pha]

Each fixture is an .npz of plain float64/int arrays (inputs and expected
outputs, [cos][E][9] or [E][9], channel a*3+b = P(a->b)); manifest.json
records how each was made.
"""
from __future__ import annotations

import json
import os
import platform
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from cudaprob3_amd import workloads as W  # noqa: E402
from oracle.loader import Osc, load_ref  # noqa: E402


def osc_of(p: W.ParamSet, nutype: int) -> Osc:
    t12, t13, t23, d, k, m12, m23 = p.for_nutype(nutype)
    return Osc(t12, t13, t23, d, m12, m23, k)


def save(name, **arrays):
    path = os.path.join(HERE, name + ".npz")
    np.savez(path, **arrays)
    return os.path.getsize(path)


def poly_golden(ref, manifest):
    """G6: the quadratic-density 4-shell model through
    setDensity(radii, a, b, c, yps), 32 x 32 points (cos from -1, the
    straight-through branch included), nominal parameters, nu and nu-bar."""
    ep = W.prem4_poly()
    cz = W.cos_grid(32)
    ev = W.log_grid(32, 0.1, 100.0)
    for nt, tag in ((0, "nu"), (1, "nubar")):
        out = ref.atmos_poly(nt, cz, ev, ep.radii, ep.a, ep.b, ep.c, ep.yps,
                             osc_of(W.NOMINAL, nt), W.PROD_HEIGHT_KM)
        name = f"g6_prem4poly_{tag}"
        size = save(name, cosv=cz, ev=ev, radii=ep.radii, a=ep.a, b=ep.b,
                    c=ep.c, yps=ep.yps, prob=out, nutype=np.int64(nt))
        manifest["fixtures"][name] = {
            "workload": "4 quadratic-density shells, 32x32, cos -1+i/32, "
                        "E 0.1-100 GeV log, h 22 km",
            "nutype": nt, "lib": "libcp3ref.so", "bytes": size}


def pha_golden(ref, manifest):
    """G7: production-height averaging (physics.hpp:997-1067), which the
    reference class never enables; the harness's subclass sets its protected
    flag (oracle/ref_harness.cpp).  4-shell model, 16 x 16 points (cos from -1
    to 0.875, downgoing included), h 22 km, 5 bins with edges 5-50 km and
    seeded bin probabilities normalised per (type, flavour, E, cos)."""
    e = W.prem4()
    cz = -1.0 + np.arange(16, dtype=np.float64) / 8
    ev = W.log_grid(16, 0.1, 100.0)
    edges = np.array([5.0, 10.0, 15.0, 22.0, 30.0, 50.0])
    rng = np.random.default_rng(997)
    prob = rng.random((2, 3, len(ev), len(cz), len(edges) - 1))
    prob /= prob.sum(axis=-1, keepdims=True)
    outs = [ref.atmos_pha(nt, cz, ev, e.radii, e.rhos, e.yps,
                          osc_of(W.NOMINAL, nt), W.PROD_HEIGHT_KM, prob, edges)
            for nt in (0, 1)]
    name = "g7_prem4_height_avg"
    size = save(name, cosv=cz, ev=ev, radii=e.radii, rhos=e.rhos, yps=e.yps,
                edges_km=edges, height_prob=prob, prob=np.stack(outs))
    manifest["fixtures"][name] = {
        "workload": "4 shells, 16x16, cos -1+i/8, E 0.1-100 GeV log, h 22 km, "
                    "production-height averaging over 5 bins (edges 5-50 km), "
                    "bin probabilities seed 997",
        "prob_axes": "nutype, cos, E, channel",
        "height_prob_axes": "nutype, flavour, E, cos, bin",
        "lib": "libcp3ref.so", "bytes": size}


def only(fn):
    """Adds one fixture group to an existing set without rewriting the
    other fixtures."""
    ref = load_ref(deep=False)
    if ref is None:
        sys.exit("oracle/_ref is not built (needs the reference sources)")
    path = os.path.join(HERE, "manifest.json")
    with open(path) as f:
        manifest = json.load(f)
    fn(ref, manifest)
    with open(path, "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)


def main():
    ref = load_ref(deep=False)
    deep = load_ref(deep=True)
    if ref is None or deep is None:
        sys.exit("oracle/_ref is not built (needs the reference sources)")
    assert deep.lib.ref_max_layers() == 64 and ref.lib.ref_max_layers() == 10
    manifest = {
        "generator": "tests/golden/make_golden.py",
        "libraries": {"libcp3ref.so": "reference CPU propagators, MaxNLayers 10",
                      "libcp3ref_deep.so": "same, MaxNLayers 64"},
        "compiler": subprocess.run(["g++", "--version"], capture_output=True,
                                   text=True).stdout.splitlines()[0],
        "flags": "-O3 -std=c++14 -fopenmp -ffp-contract=off",
        "host": platform.machine(),
        "fixtures": {},
    }

    # G0: the 4-shell model, 100 x 100 grid (0.1-100 GeV), nominal parameters;
    # every 4th cosine and energy in full, plus per-channel sums over the grid.
    e4 = W.prem4()
    cz = W.cos_grid(100)
    ev = W.log_grid(100, 0.1, 100.0)
    for nt in (0, 1):
        osc = osc_of(W.NOMINAL, nt)
        full = ref.atmos(nt, cz, ev, e4.radii, e4.rhos, e4.yps, osc,
                         W.PROD_HEIGHT_KM, threads=8)
        full32 = ref.atmos(nt, cz, ev, e4.radii, e4.rhos, e4.yps, osc,
                           W.PROD_HEIGHT_KM, threads=8, f32=True)
        name = f"g0_prem4_{'nubar' if nt else 'nu'}"
        size = save(name, cosv=cz, ev=ev, radii=e4.radii, rhos=e4.rhos,
                    yps=e4.yps, stride=np.int64(4),
                    prob=full[::4, ::4], prob_f32=full32[::4, ::4],
                    chan_sum=full.sum(axis=(0, 1)),
                    chan_sumsq=(full * full).sum(axis=(0, 1)))
        manifest["fixtures"][name] = {
            "workload": "4 shells, 100x100, cos -1+i/100, E 0.1-100 GeV log, h 22 km",
            "nutype": nt, "lib": "libcp3ref.so", "bytes": size}

    # G1: the 12-shell model at 32 x 32 points of the 2000 x 2000 grid
    # (0.1-100 GeV), including the lowest energy.
    e12 = W.earth_12()
    ci = np.linspace(0, 1999, 32).round().astype(np.int64)
    cz = W.cos_grid(2000)[ci]
    ev = W.log_grid(2000, 0.1, 100.0)[ci]
    for nt in (0, 1):
        osc = osc_of(W.NOMINAL, nt)
        out = deep.atmos(nt, cz, ev, e12.radii, e12.rhos, e12.yps, osc,
                         W.PROD_HEIGHT_KM, threads=8)
        name = f"g1_shells12_{'nubar' if nt else 'nu'}"
        size = save(name, cosv=cz, ev=ev, radii=e12.radii, rhos=e12.rhos,
                    yps=e12.yps, grid_index=ci, prob=out)
        manifest["fixtures"][name] = {
            "workload": "12 seeded shells, 32x32 of the 2000x2000 grid",
            "nutype": nt, "lib": "libcp3ref_deep.so", "bytes": size}

    # G2: beam, 8 seeded parameter sets, 128 energies over 0.1-20 GeV,
    # 1300 km at 2.6 g/cm^3, neutrinos and antineutrinos.
    sets = W.draw_param_sets(8, seed=4096)
    ev = W.log_grid(128, 0.1, 20.0)
    par = np.array([[p.th12, p.th13, p.th23, p.dcp, p.dm12, p.dm23]
                    for p in sets])
    probs = np.zeros((2, len(sets), len(ev), 9))
    probs32 = np.zeros((2, len(sets), len(ev), 9), dtype=np.float32)
    for nt in (0, 1):
        for i, p in enumerate(sets):
            probs[nt, i] = ref.beam(nt, ev, 2.6, 1300.0, osc_of(p, nt))
            probs32[nt, i] = ref.beam(nt, ev, 2.6, 1300.0, osc_of(p, nt),
                                      f32=True)
    size = save("g2_beam", ev=ev, params=par, rho=np.float64(2.6),
                L_km=np.float64(1300.0), prob=probs, prob_f32=probs32)
    manifest["fixtures"]["g2_beam"] = {
        "workload": "beam 1300 km 2.6 g/cm3, 8 sets (seed 4096), 128 E",
        "params_columns": "th12 th13 th23 dcp dm12 dm23",
        "prob_axes": "nutype, set, E, channel", "lib": "libcp3ref.so",
        "bytes": size}

    # G3: the 59-shell model, 2 seeded sets, 24 x 24 points over 0.5-50 GeV.
    e59 = W.earth_59()
    sets = W.draw_param_sets(2, seed=1024)
    cz = W.cos_grid(24)
    ev = W.log_grid(24, 0.5, 50.0)
    par = np.array([[p.th12, p.th13, p.th23, p.dcp, p.dm12, p.dm23]
                    for p in sets])
    probs = np.zeros((2, len(sets), len(cz), len(ev), 9))
    for nt in (0, 1):
        for i, p in enumerate(sets):
            probs[nt, i] = deep.atmos(nt, cz, ev, e59.radii, e59.rhos,
                                      e59.yps, osc_of(p, nt),
                                      W.PROD_HEIGHT_KM, threads=8)
    size = save("g3_shells59", cosv=cz, ev=ev, radii=e59.radii,
                rhos=e59.rhos, yps=e59.yps, params=par, prob=probs)
    manifest["fixtures"]["g3_shells59"] = {
        "workload": "59 seeded shells, 2 sets (seed 1024), 24x24, 0.5-50 GeV",
        "prob_axes": "nutype, set, cos, E, channel",
        "lib": "libcp3ref_deep.so", "bytes": size}

    # G4: edge cases on the 4-shell model: straight down, cosines 1e-9 either
    # side of each shell's limit, grazing and above-horizon paths, the grid
    # ends in energy; plus degenerate splittings and delta = 0.
    lim = -np.sqrt(1.0 - (e4.radii[1:-1] / W.R_EARTH_KM) ** 2)
    cz = np.concatenate([[-1.0, -1e-6, 0.0, 0.5, 1.0], lim - 1e-9,
                         lim + 1e-9, lim])
    ev = np.array([0.1, 0.5, 2.5, 10.0, 100.0, 1000.0])
    cases = {
        "nominal": W.NOMINAL,
        "dm12_zero": W.ParamSet(dm12=0.0),
        "dm23_zero": W.ParamSet(dm23=0.0),
        "dcp_zero": W.ParamSet(dcp=0.0),
        "inverted": W.ParamSet(dm23=-2.494e-3 - 7.53e-5),
    }
    out = {}
    for key, p in cases.items():
        for nt in (0, 1):
            out[f"{key}_{nt}"] = ref.atmos(nt, cz, ev, e4.radii, e4.rhos,
                                           e4.yps, osc_of(p, nt),
                                           W.PROD_HEIGHT_KM, threads=1)
    par = {f"params_{k}": np.array([p.th12, p.th13, p.th23, p.dcp, p.dm12,
                                    p.dm23]) for k, p in cases.items()}
    size = save("g4_edges", cosv=cz, ev=ev, radii=e4.radii, rhos=e4.rhos,
                yps=e4.yps, **par, **{f"prob_{k}": v for k, v in out.items()})
    manifest["fixtures"]["g4_edges"] = {
        "workload": "4 shells, edge cosines/energies, degenerate splittings",
        "cases": list(cases), "lib": "libcp3ref.so", "bytes": size}

    # G5: the reference's own 3-column model file loaded with
    # setDensityFromFile, 20 x 20 points.
    cz = W.cos_grid(20)
    ev = W.log_grid(20, 1.0, 100.0)
    path = "/root/reference/models/PREM_4layer.dat"
    if os.path.exists(path):
        osc = osc_of(W.NOMINAL, 0)
        out = ref.atmos_file(0, cz, ev, path, osc, W.PROD_HEIGHT_KM)
        rows = np.loadtxt(path)
        size = save("g5_prem4_file", cosv=cz, ev=ev, model_rows=rows,
                    prob=out)
        manifest["fixtures"]["g5_prem4_file"] = {
            "workload": "models/PREM_4layer.dat via setDensityFromFile, 20x20",
            "lib": "libcp3ref.so", "bytes": size}

    poly_golden(ref, manifest)
    pha_golden(ref, manifest)

    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)
    total = sum(v["bytes"] for v in manifest["fixtures"].values())
    print(f"wrote {len(manifest['fixtures'])} fixtures, {total/1024:.0f} KiB")


if __name__ == "__main__":
    if sys.argv[1:] == ["--only-poly"]:
        only(poly_golden)
    elif sys.argv[1:] == ["--only-pha"]:
        only(pha_golden)
    else:
        main()
