"""Batched engine: every job's result equals the single-propagator result of
the same inputs bitwise (a point's arithmetic does not depend on the batch),
graph replays equal eager runs, parameter updates take effect, and the
batched FP64 form keeps parity with the oracle."""
import numpy as np
import pytest

from conftest import osc_for
from cudaprob3_amd import workloads as W
from cudaprob3_amd.batch import ATMOS, BEAM, Batch, make_job

pytestmark = pytest.mark.gpu


def single_atmos(job, earth, precision):
    """The single propagator solving inside its kernel (shared mode), an
    independent path from the batch's slab table."""
    from cudaprob3_amd import AtmosPropagator
    t12, t13, t23, d, k, m12, m23 = job.params.for_nutype(job.nutype)
    with AtmosPropagator(len(job.cosv), len(job.ev), precision=precision) as p:
        p.setSolveMode("shared")
        p.setEnergyList(job.ev)
        p.setCosineList(job.cosv)
        p.setMNSMatrix(t12, t13, t23, d, k)
        p.setNeutrinoMasses(m12, m23)
        p.setDensity(earth.radii, earth.rhos, earth.yps)
        p.setProductionHeight(22.0)
        p.calculateProbabilities(job.nutype)
        return np.transpose(p.probabilities(), (2, 0, 1))  # [9][nC][nE]


def single_beam(params, nutype, ev, precision):
    from cudaprob3_amd import BeamPropagator
    t12, t13, t23, d, k, m12, m23 = params.for_nutype(nutype)
    with BeamPropagator(len(ev), precision=precision) as p:
        p.setEnergyList(ev)
        p.setMNSMatrix(t12, t13, t23, d, k)
        p.setNeutrinoMasses(m12, m23)
        p.setDensity(2.6)
        p.setPathLength(1300.0)
        p.calculateProbabilities(nutype)
        return p.probabilities().T  # [9][nE]


def small_mix():
    """Uneven grids on the 59-shell model, both neutrino types, plus beam."""
    earth = W.earth_59()
    sets = W.draw_param_sets(3, 321)
    grids = {"a": (W.cos_grid(37), W.log_grid(70, 0.5, 50.0)),
             "b": (W.cos_grid(19), W.log_grid(130, 0.5, 50.0))}
    jobs = []
    for i, p in enumerate(sets):
        g = "a" if i % 2 == 0 else "b"
        for nt in (0, 1):
            jobs.append(W.AtmosJob(p, grids[g][0], grids[g][1], nt, grid=g))
    beam_ev = W.log_grid(300, 0.1, 20.0)
    return earth, grids, jobs, sets, beam_ev


def build(precision, nc=0, group=0, waves=0, mode="auto"):
    earth, grids, jobs, sets, beam_ev = small_mix()
    b = Batch(precision)
    b.set_solve_mode(mode)
    gid = {k: b.add_atmos_grid(c, e, earth, 22.0) for k, (c, e) in grids.items()}
    bg = b.add_beam_grid(beam_ev, 2.6, 1300.0)
    cj = [make_job(ATMOS, gid[j.grid], j.params, j.nutype) for j in jobs]
    cj += [make_job(BEAM, bg, sets[0], 0), make_job(BEAM, bg, sets[1], 1)]
    b.set_jobs(cj)
    if nc or group or waves:
        b.set_tuning(nc, group, waves)
    return b, earth, jobs, sets, beam_ev


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("tuning", [(1, 1, 0, "shared"), (4, 1, -1, "shared"),
                                    (8, 2, 2, "shared"), (2, 4, 0, "shared"),
                                    (4, 4, -1, "shared"), (1, 0, 0, "table"),
                                    (2, 0, -1, "table"), (4, 0, 2, "table"),
                                    (8, 0, -1, "table")])
def test_batch_equals_single(precision, tuning):
    b, earth, jobs, sets, beam_ev = build(precision, *tuning)
    assert b.tuning()[3] == tuning[3]
    assert b.tuning()[0] == tuning[0]
    b.run()
    b.wait()
    for i, job in enumerate(jobs):
        ref = single_atmos(job, earth, precision)
        assert np.array_equal(b.result(i), ref), (i, job.grid, job.nutype)
    n = len(jobs)
    assert np.array_equal(b.result(n), single_beam(sets[0], 0, beam_ev, precision))
    assert np.array_equal(b.result(n + 1), single_beam(sets[1], 1, beam_ev, precision))
    b.close()


@pytest.mark.parametrize("mode", ["shared", "table"])
def test_graph_replays_equal_eager(mode):
    b, earth, jobs, sets, beam_ev = build("mixed", mode=mode)
    b.run()
    b.wait()
    eager = [b.result(i) for i in range(len(jobs) + 2)]
    b.use_graph(True)
    for rep in range(3):
        b.run()
        b.wait()
        for i, ref in enumerate(eager):
            assert np.array_equal(b.result(i), ref), (rep, i)
    b.close()


def test_update_jobs_and_fp64_parity(oracle):
    b, earth, jobs, sets, beam_ev = build("fp64")
    b.use_graph(True)
    b.run()
    b.wait()
    new = W.draw_param_sets(3, 999)
    # same kinds and grids, new parameter values
    cj, nj = [], []
    for i, job in enumerate(jobs):
        p = new[i // 2]
        nj.append(W.AtmosJob(p, job.cosv, job.ev, job.nutype, job.grid))
        cj.append(make_job(ATMOS, 0 if job.grid == "a" else 1, p, job.nutype))
    cj += [make_job(BEAM, 0, new[0], 0), make_job(BEAM, 0, new[1], 1)]
    b.update_jobs(cj)
    b.run()
    b.wait()
    for i, job in enumerate(nj):
        t12, t13, t23, d, k, m12, m23 = job.params.for_nutype(job.nutype)
        from oracle.loader import Osc
        ref = oracle.atmos(job.nutype, job.cosv, job.ev, earth.radii,
                           earth.rhos, earth.yps,
                           Osc(t12, t13, t23, d, m12, m23, k), 22.0, threads=4)
        got = np.transpose(b.result(i), (1, 2, 0))
        assert np.abs(got - ref).max() < 1e-10
    row = [new[1].th12, new[1].th13, new[1].th23, new[1].dcp, new[1].dm12,
           new[1].dm23]
    ref = oracle.beam(1, beam_ev, 2.6, 1300.0, osc_for(row, 1))
    assert np.abs(b.result(len(nj) + 1).T - ref).max() < 1e-10
    b.close()


def test_autotune_picks_valid():
    b, *_ = build("mixed")
    nc, group, waves = b.autotune(reps=1)
    assert nc in (1, 2, 4, 8) and group in (1, 2, 4)
    assert b.tuning()[:3] == (nc, group, waves)
    assert b.tuning()[3] in ("shared", "table")
    b.run()
    b.wait()
    assert b.last_ms() > 0
    b.close()


def test_batch_errors():
    from cudaprob3_amd import Cp3Error
    b = Batch("mixed")
    with pytest.raises(Cp3Error, match="unknown atmospheric grid"):
        b.set_jobs([make_job(ATMOS, 3, W.NOMINAL, 0)])
    # a rejected job list leaves the batch empty
    with pytest.raises(Cp3Error, match="no jobs"):
        b.run()
    g = b.add_beam_grid(W.log_grid(10, 0.1, 20.0), 2.6, 1300.0)
    g2 = b.add_beam_grid(W.log_grid(12, 0.1, 20.0), 2.6, 810.0)
    b.set_jobs([make_job(BEAM, g, W.NOMINAL, 0)])
    with pytest.raises(Cp3Error, match="kinds and grids"):
        b.update_jobs([make_job(BEAM, g2, W.NOMINAL, 0)])
    with pytest.raises(Cp3Error, match="out of range"):
        b.result_device(5)
    b.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("tuning", [(1, 1, 0, "shared"), (4, 4, -1, "shared"),
                                    (8, 2, 0, "shared"), (1, 0, 0, "table"),
                                    (4, 0, -1, "table"), (8, 0, 0, "table")])
def test_packed_tails_edge_cases(precision, tuning):
    """Jobs sharing a grid run their last nE % 64 energies in packed tail
    tiles: a grid with fewer than 64 energies (every lane in tails, several
    tiles, the last one partial), a tail above 32 (not packed), an exact
    multiple of 64 (no tail) and a lone job on a grid (not packed); every
    result equals the single propagator's bitwise."""
    earth = W.earth_59()
    sets = W.draw_param_sets(4, 55)
    grids = {"c": (W.cos_grid(11), W.log_grid(20, 0.5, 50.0)),
             "d": (W.cos_grid(9), W.log_grid(100, 0.5, 50.0)),
             "e": (W.cos_grid(5), W.log_grid(64, 0.5, 50.0)),
             "f": (W.cos_grid(7), W.log_grid(90, 0.5, 50.0))}
    plan = [("c", s, nt) for s in sets for nt in (0, 1)][:7]
    plan += [("d", sets[0], 0), ("d", sets[1], 1), ("e", sets[2], 0),
             ("e", sets[3], 1), ("f", sets[0], 1)]
    b = Batch(precision)
    b.set_solve_mode(tuning[3])
    gid = {k: b.add_atmos_grid(c, e, earth, 22.0) for k, (c, e) in grids.items()}
    jobs = [W.AtmosJob(p, grids[g][0], grids[g][1], nt, grid=g) for g, p, nt in plan]
    b.set_jobs([make_job(ATMOS, gid[j.grid], j.params, j.nutype) for j in jobs])
    b.set_tuning(*tuning[:3])
    assert b.tuning()[3] == tuning[3]
    b.run()
    b.wait()
    for i, job in enumerate(jobs):
        ref = single_atmos(job, earth, precision)
        assert np.array_equal(b.result(i), ref), (i, job.grid, job.nutype)
    b.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("mode", ["shared", "table"])
def test_tiny_grids(precision, mode):
    """Grids of 1, 3 and 5 energies and 1 or 2 cosines: a pre-pass wave
    spans many table rows (every lane walks its own row), every lane of a
    work unit but a few is masked; results equal the single propagator's
    bitwise."""
    earth = W.earth_59()
    sets = W.draw_param_sets(2, 77)
    grids = {"g1": (W.cos_grid(1), W.log_grid(1, 0.5, 50.0)),
             "g3": (W.cos_grid(2), W.log_grid(3, 0.5, 50.0)),
             "g5": (np.array([-0.99, -0.05]), W.log_grid(5, 0.1, 100.0))}
    plan = [(g, s, nt) for g in grids for s in sets for nt in (0, 1)]
    b = Batch(precision)
    b.set_solve_mode(mode)
    gid = {k: b.add_atmos_grid(c, e, earth, 22.0) for k, (c, e) in grids.items()}
    jobs = [W.AtmosJob(p, grids[g][0], grids[g][1], nt, grid=g) for g, p, nt in plan]
    b.set_jobs([make_job(ATMOS, gid[j.grid], j.params, j.nutype) for j in jobs])
    b.run()
    b.wait()
    assert b.tuning()[3] == mode
    for i, job in enumerate(jobs):
        ref = single_atmos(job, earth, precision)
        assert np.array_equal(b.result(i), ref), (i, job.grid, job.nutype)
    b.close()


@pytest.mark.parametrize("mode", ["shared", "table"])
def test_phase_timings(mode):
    """Events between the launches split a run into its phases, which lie
    within the run's time; only table mode has a pre-pass."""
    b, *_ = build("mixed", mode=mode)
    for graph in (False, True):
        b.use_graph(graph)
        b.run()
        b.wait()
        ph = b.phase_ms()
        assert all(v >= 0 for v in ph.values())
        # the phases lie within the run (the rest is launch overhead)
        assert sum(ph.values()) <= b.last_ms() + 0.02
        if graph:
            continue
        assert ph["atmos"] > 0 and ph["beam"] > 0
        if mode == "table":
            assert ph["prepass"] > 0
        else:  # nothing launched: only the events' own spacing
            assert ph["prepass"] < 0.05
    b.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("mode", ["shared", "table"])
def test_random_grid_shapes(precision, mode):
    """Seeded random grids: 1-70 cosines drawn over [-1, 1) (down-going
    trajectories without shells among them), 1-200 energies, several jobs per
    grid so that packed tails form; every result equals the single
    propagator's bitwise."""
    rng = np.random.default_rng(777)
    earth = W.earth_59()
    sets = W.draw_param_sets(3, 778)
    grids = {}
    for k in range(6):
        nc, ne = int(rng.integers(1, 71)), int(rng.integers(1, 201))
        cz = np.sort(rng.uniform(-1.0, 1.0, nc))
        ev = np.sort(rng.uniform(0.2, 60.0, ne))
        grids[f"r{k}"] = (cz, ev)
    plan = [(g, s, nt) for g in grids for s in sets for nt in (0, 1)]
    b = Batch(precision)
    b.set_solve_mode(mode)
    gid = {k: b.add_atmos_grid(c, e, earth, 22.0) for k, (c, e) in grids.items()}
    jobs = [W.AtmosJob(p, grids[g][0], grids[g][1], nt, grid=g) for g, p, nt in plan]
    b.set_jobs([make_job(ATMOS, gid[j.grid], j.params, j.nutype) for j in jobs])
    b.run()
    b.wait()
    for i, job in enumerate(jobs):
        ref = single_atmos(job, earth, precision)
        assert np.array_equal(b.result(i), ref), (i, job.grid, job.nutype)
    b.close()


def test_beam_batch_over_65535_jobs():
    """70000 beam jobs (100 energies each, two grids of different lengths):
    the beam launch runs in slices of 65535 jobs; sampled jobs from every
    slice, on both grids, equal the single propagator bitwise."""
    sets = W.draw_param_sets(7, 70000)
    ev_a, ev_b = W.log_grid(100, 0.1, 20.0), W.log_grid(37, 0.3, 9.0)
    with Batch("mixed") as b:
        ga = b.add_beam_grid(ev_a, 2.6, 1300.0)
        gb = b.add_beam_grid(ev_b, 2.6, 1300.0)
        n = 70000
        jobs = [make_job(BEAM, ga if j % 3 else gb, sets[j % 7], j % 2) for j in range(n)]
        b.set_jobs(jobs)
        b.run()
        b.wait()
        for j in (0, 1, 2, 65534, 65535, 65536, 69998, 69999):
            ev = ev_a if j % 3 else ev_b
            assert np.array_equal(b.result(j), single_beam(sets[j % 7], j % 2, ev, "mixed")), j
