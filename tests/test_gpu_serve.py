"""Request server: continuous batching.  A request's result is bitwise the
same alone and in a mix of other requests (and equal to the single
propagator's), admission respects the device-memory cap, oversized requests
are refused, and batches overlap on the device."""
import numpy as np
import pytest

from cudaprob3_amd import Cp3Error
from cudaprob3_amd import workloads as W
from cudaprob3_amd.batch import ATMOS, BEAM, make_job
from cudaprob3_amd.serving import Server, draw_requests, serving_setup

pytestmark = pytest.mark.gpu


def single_atmos(n, params, nutype, precision):
    from cudaprob3_amd import AtmosPropagator
    e = W.earth_12()
    t12, t13, t23, d, k, m12, m23 = params.for_nutype(nutype)
    with AtmosPropagator(n, n, precision=precision) as p:
        p.setEnergyList(W.log_grid(n, 0.1, 100.0))
        p.setCosineList(W.cos_grid(n))
        p.setMNSMatrix(t12, t13, t23, d, k)
        p.setNeutrinoMasses(m12, m23)
        p.setDensity(e.radii, e.rhos, e.yps)
        p.setProductionHeight(W.PROD_HEIGHT_KM)
        p.calculateProbabilities(nutype)
        return np.transpose(p.probabilities(), (2, 0, 1))


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
@pytest.mark.parametrize("mode", ["table", "shared"])
def test_alone_equals_in_mix(precision, mode):
    with Server(precision, mem_cap_bytes=2 << 30, slots=3) as srv:
        srv.set_solve_mode(mode)
        agrid, bgrid = serving_setup(srv)
        p = W.draw_param_sets(1, 4242)[0]
        probe_a = make_job(ATMOS, agrid[250], p, 1)
        probe_b = make_job(BEAM, bgrid[10000], p, 0)
        # alone: each probe is a batch of its own (its job passed by value),
        # while in the mix below it shares a batch's job table
        ra = srv.submit(probe_a)
        srv.wait(ra)
        rb = srv.submit(probe_b)
        srv.wait(rb)
        assert srv.stats()["batches"] == 2
        alone_a, alone_b = srv.fetch(ra), srv.fetch(rb)
        mix = draw_requests(40, 77, agrid, bgrid)
        ids = [srv.submit(j) for j in mix[:20]]
        ra2 = srv.submit(probe_a)
        ids += [srv.submit(j) for j in mix[20:]]
        rb2 = srv.submit(probe_b)
        srv.wait(ra2)
        srv.wait(rb2)
        assert np.array_equal(srv.fetch(ra2), alone_a)
        assert np.array_equal(srv.fetch(rb2), alone_b)
        for i in ids:
            srv.wait(i)
            srv.release(i)
        assert srv.stats()["mem_in_use"] == 0
    assert np.array_equal(alone_a, single_atmos(250, p, 1, precision))


def test_memory_cap_respected():
    one = 250 * 250 * 9 * 4  # a 250^2 mixed-precision output
    cap = 2 * one + 4096
    with Server("mixed", mem_cap_bytes=cap, slots=4) as srv:
        agrid, _ = serving_setup(srv)
        sets = W.draw_param_sets(10, 5)
        ids = [srv.submit(make_job(ATMOS, agrid[250], p, 0)) for p in sets]
        for i in ids:
            srv.wait(i, timeout_ms=20000)
            out = srv.fetch(i)  # frees its memory, admitting the next
            assert out.shape == (9, 250, 250)
            assert np.isfinite(out).all()
        st = srv.stats()
        assert st["completed"] == 10
        assert 0 < st["mem_peak"] <= cap
        with pytest.raises(Cp3Error, match="the server's memory cap"):
            srv.submit(make_job(ATMOS, agrid[1000], sets[0], 0))


def test_batches_overlap_and_throughput():
    with Server("mixed", mem_cap_bytes=8 << 30, slots=4) as srv:
        agrid, bgrid = serving_setup(srv)
        jobs = draw_requests(300, 11, agrid, bgrid)
        ids = [srv.submit(j) for j in jobs]
        for i in ids:
            srv.wait(i, timeout_ms=60000)
            srv.release(i)
        st = srv.stats()
        print(st)
        assert st["completed"] == 300
        assert st["batches"] >= 2
        assert st["peak_inflight_batches"] >= 2


def test_server_errors():
    with Server("mixed", mem_cap_bytes=1 << 20, slots=1) as srv:
        with pytest.raises(Cp3Error, match="unknown atmospheric grid"):
            srv.submit(make_job(ATMOS, 0, W.NOMINAL, 0))
        with pytest.raises(Cp3Error, match="unknown request"):
            srv.release(12345)


def test_stream_keeps_up_at_ten_times_the_rate():
    """The request stream of the serving workload at 20,000 requests/s (ten
    times its nominal rate) for half a second: every request completes, the
    results are collected as fast as they arrive (the arena stays far below
    its cap instead of filling up and stalling admission) and the median
    latency stays under a millisecond (about 30 us on MI355X)."""
    from cudaprob3_amd.serving import run_stream
    r = run_stream("mixed", 20000.0, duration_s=0.5, warmup_s=0.1)
    print("20,000 req/s:", {k: r[k] for k in ("requests", "p50_us", "p99_us",
                                               "mem_peak_bytes")}, flush=True)
    assert r["requests"] > 0.8 * 20000 * 0.5
    assert r["mem_peak_bytes"] < r["mem_cap_bytes"] // 4
    assert r["p50_us"] < 1000.0
