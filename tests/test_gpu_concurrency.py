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
