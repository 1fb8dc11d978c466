"""Request server (the ``cp3_server_*`` part of the C ABI) and the Poisson
serving workload.

The server batches single-evaluation requests continuously on one device:
requests are admitted in arrival order into free launch slots while their
outputs fit in a device-memory arena, batches on different slots overlap,
and outputs stay resident until fetched or released.

``python -m cudaprob3_amd.serving --precision mixed`` runs the serving
workload: Poisson arrivals at 2000 requests/s, each with one seeded parameter
set; 70% atmospheric on the 12-shell model with a grid drawn from
{100^2, 250^2, 500^2, 1000^2} (cos in [-1, 0), E 0.1-100 GeV), 30% beam with
1e3-1e5 log-spaced energies (0.1-20 GeV, 1300 km, 2.6 g/cm^3), under a 32 GB
cap.  Latency is submit -> result resident in HBM.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import time

import numpy as np

from . import _native as N
from . import workloads as W
from .batch import ATMOS, BEAM, make_job


class Server:
    def __init__(self, precision="mixed", mem_cap_bytes=32 << 30, slots=4,
                 max_batch=256, device=0):
        self._lib = N.load()
        self._h = C.c_void_p()
        self.precision = precision
        N.check(self._lib.cp3_server_create(int(device),
                                            N.PRECISIONS[precision],
                                            int(mem_cap_bytes), int(slots),
                                            int(max_batch), C.byref(self._h)))
        self.grid_shapes = {ATMOS: [], BEAM: []}
        self.req_shape = {}

    def close(self):
        if self._h:
            N.check(self._lib.cp3_server_destroy(self._h))
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_solve_mode(self, mode="table"):
        """"table" (slabs pre-solved per request, the default) or "shared"
        (solved inside the propagation kernel); results do not depend on it."""
        N.check(self._lib.cp3_server_set_solve_mode(self._h, N.SOLVE_MODES[mode]))

    def set_tuning(self, cosine_block, group_waves):
        N.check(self._lib.cp3_server_set_tuning(self._h, int(cosine_block),
                                                int(group_waves)))

    def add_atmos_grid(self, cosv, ev, earth, h_km) -> int:
        f = lambda x: np.ascontiguousarray(x, dtype=np.float64)
        c, e, r, d, y = f(cosv), f(ev), f(earth.radii), f(earth.rhos), f(earth.yps)
        dp = C.POINTER(C.c_double)
        gid = C.c_int()
        N.check(self._lib.cp3_server_add_atmos_grid(
            self._h, c.ctypes.data_as(dp), len(c), e.ctypes.data_as(dp), len(e),
            r.ctypes.data_as(dp), d.ctypes.data_as(dp), y.ctypes.data_as(dp),
            len(r), float(h_km), C.byref(gid)))
        self.grid_shapes[ATMOS].append((9, len(c), len(e)))
        return gid.value

    def add_beam_grid(self, ev, rho, L_km) -> int:
        e = np.ascontiguousarray(ev, dtype=np.float64)
        gid = C.c_int()
        N.check(self._lib.cp3_server_add_beam_grid(
            self._h, e.ctypes.data_as(C.POINTER(C.c_double)), len(e),
            float(rho), float(L_km), C.byref(gid)))
        self.grid_shapes[BEAM].append((9, len(e)))
        return gid.value

    def submit(self, job) -> int:
        rid = C.c_int64()
        N.check(self._lib.cp3_server_submit(self._h, C.byref(job), C.byref(rid)))
        self.req_shape[rid.value] = self.grid_shapes[job.kind][job.grid]
        return rid.value

    def wait(self, rid, timeout_ms=-1):
        N.check(self._lib.cp3_server_wait(self._h, C.c_int64(rid), int(timeout_ms)))

    def fetch(self, rid) -> np.ndarray:
        out = np.zeros(self.req_shape.pop(rid))
        N.check(self._lib.cp3_server_fetch(self._h, C.c_int64(rid),
                                           out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def release(self, rid):
        self.req_shape.pop(rid, None)
        N.check(self._lib.cp3_server_release(self._h, C.c_int64(rid)))

    def latency(self, rid):
        q, t = C.c_double(), C.c_double()
        N.check(self._lib.cp3_server_latency(self._h, C.c_int64(rid),
                                             C.byref(q), C.byref(t)))
        return q.value, t.value

    def stats(self) -> dict:
        v = [C.c_int64() for _ in range(6)]
        N.check(self._lib.cp3_server_stats(self._h, *[C.byref(x) for x in v]))
        keys = ["completed", "evals", "batches", "mem_in_use", "mem_peak",
                "peak_inflight_batches"]
        return dict(zip(keys, [x.value for x in v]))


ATMOS_SIZES = (100, 250, 500, 1000)
BEAM_SIZES = tuple(int(round(x)) for x in np.logspace(3, 5, 9))


def serving_setup(server):
    earth = W.earth_12()
    agrid = {n: server.add_atmos_grid(W.cos_grid(n), W.log_grid(n, 0.1, 100.0),
                                      earth, W.PROD_HEIGHT_KM)
             for n in ATMOS_SIZES}
    bgrid = {n: server.add_beam_grid(W.log_grid(n, 0.1, 20.0), 2.6, 1300.0)
             for n in BEAM_SIZES}
    return agrid, bgrid


def draw_requests(n, seed, agrid, bgrid):
    """The request mix: 70% atmospheric, 30% beam; one seeded parameter set
    and a random neutrino type per request."""
    rng = np.random.default_rng(seed)
    sets = W.draw_param_sets(n, seed + 1)
    kinds = rng.random(n) < 0.7
    asz = rng.integers(0, len(ATMOS_SIZES), n)
    bsz = rng.integers(0, len(BEAM_SIZES), n)
    nts = rng.integers(0, 2, n)
    jobs = []
    for i in range(n):
        if kinds[i]:
            jobs.append(make_job(ATMOS, agrid[ATMOS_SIZES[asz[i]]], sets[i], nts[i]))
        else:
            jobs.append(make_job(BEAM, bgrid[BEAM_SIZES[bsz[i]]], sets[i], nts[i]))
    return jobs


def run_stream(precision="mixed", rate=2000.0, duration_s=3.0, warmup_s=0.5,
               seed=2000, slots=4, mem_cap_bytes=32 << 30):
    """Poisson arrivals; returns latency percentiles (microseconds, submit ->
    result resident in HBM, timed inside the server) and throughput.

    One thread submits each request at its arrival time and, between
    arrivals, polls the oldest outstanding request without blocking and
    releases it once done.  (A separate collector thread would have to take
    the GIL from the spinning submitter for every call; above about 5000
    requests/s it then released results more slowly than they arrived, and
    the unreleased ones filled the memory cap and stalled admission.)"""
    from collections import deque
    with Server(precision, mem_cap_bytes, slots) as srv:
        agrid, bgrid = serving_setup(srv)
        n = int(rate * (duration_s + warmup_s)) + 1
        jobs = draw_requests(n, seed, agrid, bgrid)
        rng = np.random.default_rng(seed + 7)
        arrivals = np.cumsum(rng.exponential(1.0 / rate, n))
        ids = [None] * n
        lib, h = srv._lib, srv._h
        lat, queue, evals = [], [], 0
        outstanding = deque()
        q, t = C.c_double(), C.c_double()
        t0 = time.perf_counter()
        i = 0
        while i < n or outstanding:
            if i < n and time.perf_counter() - t0 >= arrivals[i]:
                ids[i] = srv.submit(jobs[i])
                outstanding.append(i)
                i += 1
                continue
            if not outstanding:
                continue
            k = outstanding[0]
            rc = lib.cp3_server_wait(h, C.c_int64(ids[k]), 0)
            if rc == N.CP3_ERR_STATE:
                # not finished yet: each poll takes the server's lock, which
                # its scheduler needs too, so the next one comes 5 us later
                # (or at the next arrival)
                t_next = time.perf_counter() + 5e-6
                while time.perf_counter() < t_next and not (
                        i < n and time.perf_counter() - t0 >= arrivals[i]):
                    pass
                continue
            N.check(rc)
            N.check(lib.cp3_server_latency(h, C.c_int64(ids[k]), C.byref(q), C.byref(t)))
            if arrivals[k] >= warmup_s:
                lat.append(t.value)
                queue.append(q.value)
                shape = srv.req_shape[ids[k]]
                evals += int(np.prod(shape[1:]))
            srv.release(ids[k])
            outstanding.popleft()
        st = srv.stats()
        measured_span = arrivals[-1] - warmup_s
    lat, queue = np.array(lat), np.array(queue)
    return {
        "precision": precision, "rate_req_s": rate, "requests": int(len(lat)),
        "p50_us": float(np.percentile(lat, 50)),
        "p99_us": float(np.percentile(lat, 99)),
        "mean_us": float(lat.mean()),
        "queue_p50_us": float(np.percentile(queue, 50)),
        "queue_p99_us": float(np.percentile(queue, 99)),
        "evals_per_s": evals / measured_span,
        "req_per_s": len(lat) / measured_span,
        "batches": st["batches"], "peak_inflight_batches": st["peak_inflight_batches"],
        "mem_peak_bytes": st["mem_peak"], "mem_cap_bytes": mem_cap_bytes,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--precision", default="mixed")
    ap.add_argument("--rate", type=float, default=2000.0)
    ap.add_argument("--duration", type=float, default=3.0)
    args = ap.parse_args()
    print(json.dumps(run_stream(args.precision, args.rate, args.duration)))


if __name__ == "__main__":
    main()
