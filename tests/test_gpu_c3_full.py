"""The headline batch at its full size: 1024 seeded parameter sets on the
59-shell seeded Earth model, 512 on 400 x 400 and 512 on 200 x 200 (cos x E)
grids over 0.5-50 GeV, each for neutrinos and antineutrinos (2048 jobs,
2.048e8 evaluations), built as the benchmark builds it (autotuned, replayed
from a hipGraph).

* mixed form: a spread of sampled jobs against the oracle (the reference's
  calculate_atmos, physics.hpp:789-1150) within 1e-5 on every point;
* FP64 form of the same batch: sampled jobs within 1e-10 of the oracle;
* graph replays equal the eager run bitwise, and sampled jobs equal the
  single-grid propagator bitwise;
* every point of every job, in both forms: finite, in [0, 1], and each row
  and column of its 3 x 3 probability matrix summing to 1 (1e-5 mixed,
  1e-12 FP64).

Every job's last nE mod 64 energies (16 on the 400-energy grid, 8 on the
200-energy grid) run in packed tail tiles of 4 and 8 jobs; the sampled jobs
take every slot position of those tiles.
"""
import os

import numpy as np
import pytest

from conftest import conservation_all_jobs, osc_for
from cudaprob3_amd import workloads as W
from cudaprob3_amd.batch import ATMOS, Batch, make_job

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-10
TOL_MIXED = 1e-5


def oracle_threads():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def build(precision, wl, graph=True, tune=True):
    b = Batch(precision)
    grids = {}
    for j in wl.jobs:
        if j.grid not in grids:
            grids[j.grid] = b.add_atmos_grid(j.cosv, j.ev, wl.earth, wl.h_km)
    b.set_jobs([make_job(ATMOS, grids[j.grid], j.params, j.nutype)
                for j in wl.jobs])
    if tune:
        b.autotune(reps=1)
    b.use_graph(graph)
    return b


def sampled_jobs(n_jobs, per_grid):
    """per_grid jobs of each grid half, both neutrino types, at every slot
    position of the 4- and 8-job tail tiles (a stride of 2 * 31 over the
    (set, nutype) pairs: 31 is odd, so the slot residues mod 8 cycle), plus
    the first and last job of each half."""
    half = n_jobs // 2
    picks = set()
    for base in (0, half):
        for k in range(per_grid):
            picks.add(base + (k * 31) % half)
        picks.update((base, base + half - 1))
    return sorted(picks)


def oracle_job(oracle, wl, j):
    job = wl.jobs[j]
    ref = oracle.atmos(job.nutype, job.cosv, job.ev, wl.earth.radii,
                       wl.earth.rhos, wl.earth.yps,
                       osc_for([job.params.th12, job.params.th13,
                                job.params.th23, job.params.dcp,
                                job.params.dm12, job.params.dm23], job.nutype),
                       wl.h_km, threads=oracle_threads())
    return np.transpose(ref, (2, 0, 1))  # [9][nC][nE]


def test_c3_full_size_mixed_and_fp64(oracle):
    wl = W.atmos1024()
    assert len(wl.jobs) == 2048 and wl.n_evals == 204_800_000
    print("c3: 2048 jobs, 59 shells, 400^2 / 200^2", flush=True)
    mixed = build("mixed", wl)
    tun = mixed.tuning()
    print("mixed tuning", tun, flush=True)
    # first run captures the graph, the next replay it; eager work and host
    # synchronisation between replays
    picks = sampled_jobs(len(wl.jobs), 32)
    replays = []
    for _ in range(3):
        mixed.run()
        mixed.wait()
        replays.append({j: mixed.result(j) for j in picks[::9]})
    mixed.use_graph(False)
    mixed.run()
    mixed.wait()
    for j in picks[::9]:
        eager = mixed.result(j)
        for r in replays:
            assert np.array_equal(r[j], eager), f"replay differs on job {j}"
    # mixed against the oracle
    worst, sq, npts = 0.0, 0.0, 0
    kinds = set()
    for j in picks:
        job = wl.jobs[j]
        kinds.add((len(job.ev), job.nutype))
        got = mixed.result(j)
        ref = oracle_job(oracle, wl, j)
        d = np.abs(got - ref)
        worst = max(worst, float(d.max()))
        sq += float((d ** 2).sum())
        npts += d.size
        assert d.max() < TOL_MIXED, f"job {j}: {d.max():.3e}"
        # the packed tail energies are among the compared points
        tail = len(job.ev) % 64
        assert d[:, :, -tail:].max() < TOL_MIXED
        assert np.abs(got.reshape(3, 3, *got.shape[1:]).sum(1) - 1).max() < 1e-5
    assert kinds == {(400, 0), (400, 1), (200, 0), (200, 1)}
    print(f"mixed vs oracle over {len(picks)} jobs: max {worst:.3e} "
          f"rms {np.sqrt(sq / npts):.3e}", flush=True)

    # sampled jobs equal the single-grid propagator (in-kernel solve) bitwise
    from cudaprob3_amd import AtmosPropagator
    for j in (picks[0], picks[1], picks[-2], picks[-1]):
        job = wl.jobs[j]
        t12, t13, t23, d, k, m12, m23 = job.params.for_nutype(job.nutype)
        with AtmosPropagator(len(job.cosv), len(job.ev), precision="mixed") as p:
            p.setSolveMode("shared")
            p.setEnergyList(job.ev)
            p.setCosineList(job.cosv)
            p.setMNSMatrix(t12, t13, t23, d, k)
            p.setNeutrinoMasses(m12, m23)
            p.setDensity(wl.earth.radii, wl.earth.rhos, wl.earth.yps)
            p.setProductionHeight(wl.h_km)
            p.calculateProbabilities(job.nutype)
            single = np.transpose(p.probabilities(), (2, 0, 1))
        assert np.array_equal(single, mixed.result(j)), f"job {j}"
    keep = {j: mixed.result(j) for j in picks[::4]}
    # all 2.048e8 points: probability conserved
    dev = conservation_all_jobs(mixed, len(wl.jobs), TOL_MIXED)
    print(f"mixed, all {len(wl.jobs)} jobs: sums within {dev:.3e} of 1", flush=True)
    mixed.close()

    # the FP64 form of the same batch
    f64 = build("fp64", wl)
    print("fp64 tuning", f64.tuning(), flush=True)
    f64.run()
    f64.run()
    f64.wait()
    fp64_picks = picks[::4][:12]
    worst64, worst_mix64 = 0.0, 0.0
    for j in fp64_picks:
        got = f64.result(j)
        ref = oracle_job(oracle, wl, j)
        worst64 = max(worst64, float(np.abs(got - ref).max()))
        worst_mix64 = max(worst_mix64, float(np.abs(keep[j] - got).max()))
        P = got.reshape(3, 3, *got.shape[1:])
        assert max(np.abs(P.sum(0) - 1).max(), np.abs(P.sum(1) - 1).max()) < 1e-12
    dev64 = conservation_all_jobs(f64, len(wl.jobs), 1e-12)
    print(f"fp64, all {len(wl.jobs)} jobs: sums within {dev64:.3e} of 1", flush=True)
    f64.close()
    print(f"fp64 vs oracle over {len(fp64_picks)} jobs: max {worst64:.3e}; "
          f"mixed vs fp64 (GPU) max {worst_mix64:.3e}", flush=True)
    assert worst64 < TOL_FP64
    assert worst_mix64 < TOL_MIXED
