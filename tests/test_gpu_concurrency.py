"""Propagator objects hold no process-global state (the reference keeps its
parameters in __constant__ memory and static arrays, physics.hpp:115-125, so
its instances race): several objects driven from their own host threads at
once give exactly the results each gives alone."""
import threading

import numpy as np
import pytest

from cudaprob3_amd import workloads as W

pytestmark = pytest.mark.gpu


def atmos_once(params, nutype, precision, cz, ev, earth, out, key):
    from cudaprob3_amd import AtmosPropagator
    t12, t13, t23, d, k, m12, m23 = params.for_nutype(nutype)
    with AtmosPropagator(len(cz), len(ev), precision=precision) as p:
        p.setEnergyList(ev)
        p.setCosineList(cz)
        p.setMNSMatrix(t12, t13, t23, d, k)
        p.setNeutrinoMasses(m12, m23)
        p.setDensity(earth.radii, earth.rhos, earth.yps)
        p.setProductionHeight(22.0)
        for _ in range(3):  # repeated launches overlap the other threads' work
            p.calculateProbabilitiesAsync(nutype)
            p.waitForCompletion()
        out[key] = p.probabilities()


def beam_once(params, nutype, precision, ev, out, key):
    from cudaprob3_amd import BeamPropagator
    t12, t13, t23, d, k, m12, m23 = params.for_nutype(nutype)
    with BeamPropagator(len(ev), precision=precision) as p:
        p.setEnergyList(ev)
        p.setMNSMatrix(t12, t13, t23, d, k)
        p.setNeutrinoMasses(m12, m23)
        p.setDensity(2.6)
        p.setPathLength(1300.0)
        for _ in range(3):
            p.calculateProbabilitiesAsync(nutype)
            p.waitForCompletion()
        out[key] = p.probabilities()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_threads_equal_alone(precision):
    earth = W.earth_12()
    cz, ev = W.cos_grid(120), W.log_grid(150, 0.1, 100.0)
    bev = W.log_grid(20000, 0.1, 20.0)
    sets = W.draw_param_sets(6, 4242)
    work = [("a", i, s, i % 2) for i, s in enumerate(sets[:4])]
    work += [("b", i, s, i % 2) for i, s in enumerate(sets[4:])]

    def run(kind, i, s, nt, out):
        if kind == "a":
            atmos_once(s, nt, precision, cz, ev, earth, out, (kind, i))
        else:
            beam_once(s, nt, precision, bev, out, (kind, i))

    alone = {}
    for w in work:
        run(*w, alone)
    together = {}
    errors = []

    def guarded(w):
        try:
            run(*w, together)
        except Exception as exc:  # reported below, in the test's own thread
            errors.append((w[:2], exc))

    threads = [threading.Thread(target=guarded, args=(w,)) for w in work]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not errors, errors
    assert sorted(together) == sorted(alone)
    for key in alone:
        assert np.array_equal(together[key], alone[key]), key


def test_setters_after_an_asynchronous_launch():
    """Setters called while an asynchronous calculation is still running (no
    wait in between) change the tables only after it: a new Earth model with
    more shells, new cosines and new energies, then a calculation, give the
    bits of a fresh object set up that way."""
    from cudaprob3_amd import AtmosPropagator, BeamPropagator
    n = 512
    e4, e12 = W.prem4(), W.earth_12()
    cz, ev = W.cos_grid(n), W.log_grid(n, 0.1, 100.0)
    cz2, ev2 = W.cos_grid(n)[::-1].copy(), W.log_grid(n, 0.2, 50.0)

    def setup(p, cosv, en, earth):
        p.setEnergyList(en)
        p.setCosineList(cosv)
        p.setMNSMatrix(W.TH12, W.TH13, W.TH23, W.DCP, 1)
        p.setNeutrinoMasses(W.DM12, W.DM23)
        p.setDensity(earth.radii, earth.rhos, earth.yps)
        p.setProductionHeight(22.0)

    for mode in ("table", "shared"):
        with AtmosPropagator(n, n, precision="mixed") as p:
            p.setSolveMode(mode)
            setup(p, cz, ev, e4)
            p.calculateProbabilitiesAsync(0)
            setup(p, cz2, ev2, e12)
            p.calculateProbabilities(0)
            got = p.probabilities()
        with AtmosPropagator(n, n, precision="mixed") as q:
            q.setSolveMode(mode)
            setup(q, cz2, ev2, e12)
            q.calculateProbabilities(0)
            assert np.array_equal(got, q.probabilities())
    bev = W.log_grid(100000, 0.1, 20.0)
    with BeamPropagator(len(bev)) as b:
        b.setEnergyList(bev)
        b.setMNSMatrix(W.TH12, W.TH13, W.TH23, W.DCP, 1)
        b.setNeutrinoMasses(W.DM12, W.DM23)
        b.setDensity(2.6)
        b.setPathLength(1300.0)
        b.calculateProbabilitiesAsync(0)
        b.setEnergyList(bev[::-1].copy())
        b.calculateProbabilities(0)
        got = b.probabilities()
    with BeamPropagator(len(bev)) as c:
        c.setEnergyList(bev[::-1].copy())
        c.setMNSMatrix(W.TH12, W.TH13, W.TH23, W.DCP, 1)
        c.setNeutrinoMasses(W.DM12, W.DM23)
        c.setDensity(2.6)
        c.setPathLength(1300.0)
        c.calculateProbabilities(0)
        assert np.array_equal(got, c.probabilities())
