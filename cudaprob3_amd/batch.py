"""Batched evaluation of many parameter sets on resident grids (the
``cp3_batch_*`` part of the C ABI): one launch per kind and phase over all
jobs (persistent or one workgroup per work unit, as the autotuner finds
faster), optional hipGraph replay, autotuned cosine block and grid size."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from ._native import Cp3Job

ATMOS, BEAM = 0, 1


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _f64(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64).ravel())


def make_job(kind, grid, params, nutype) -> Cp3Job:
    """params: a workloads.ParamSet (or anything with th12 ... dm23)."""
    return Cp3Job(kind, grid, int(nutype), -1 if nutype else 1,
                  params.th12, params.th13, params.th23, params.dcp,
                  params.dm12, params.dm23)


class Batch:
    def __init__(self, precision: str = "mixed", device: int = 0):
        self._lib = N.load()
        self._h = C.c_void_p()
        self.precision = precision
        N.check(self._lib.cp3_batch_create(int(device),
                                           N.PRECISIONS[precision],
                                           C.byref(self._h)))
        self.shapes = []  # per job: output shape
        self.grid_shapes = {ATMOS: [], BEAM: []}

    def close(self):
        if self._h:
            N.check(self._lib.cp3_batch_destroy(self._h))
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def add_atmos_grid(self, cosv, ev, earth, h_km) -> int:
        c, e = _f64(cosv), _f64(ev)
        r, d, y = _f64(earth.radii), _f64(earth.rhos), _f64(earth.yps)
        gid = C.c_int()
        N.check(self._lib.cp3_batch_add_atmos_grid(
            self._h, _dptr(c), len(c), _dptr(e), len(e), _dptr(r), _dptr(d),
            _dptr(y), len(r), float(h_km), C.byref(gid)))
        self.grid_shapes[ATMOS].append((9, len(c), len(e)))
        return gid.value

    def add_beam_grid(self, ev, rho, L_km) -> int:
        e = _f64(ev)
        gid = C.c_int()
        N.check(self._lib.cp3_batch_add_beam_grid(self._h, _dptr(e), len(e),
                                                  float(rho), float(L_km),
                                                  C.byref(gid)))
        self.grid_shapes[BEAM].append((9, len(e)))
        return gid.value

    def set_jobs(self, jobs):
        arr = (Cp3Job * len(jobs))(*jobs)
        N.check(self._lib.cp3_batch_set_jobs(self._h, arr, len(jobs)))
        self.shapes = [self.grid_shapes[j.kind][j.grid] for j in jobs]

    def update_jobs(self, jobs):
        arr = (Cp3Job * len(jobs))(*jobs)
        N.check(self._lib.cp3_batch_update_jobs(self._h, arr, len(jobs)))

    def set_tuning(self, cosine_block=0, group_waves=0, workgroups_per_cu=0):
        """(cosine block, waves sharing each solve, resident workgroups per
        CU or -1 for one workgroup per work unit); 0 = automatic."""
        N.check(self._lib.cp3_batch_set_tuning(self._h, int(cosine_block),
                                               int(group_waves),
                                               int(workgroups_per_cu)))

    def set_solve_mode(self, mode="auto"):
        """Where the eigen-solutions come from: "shared" (solved in the
        propagation kernel, shared through LDS by a work unit), "table"
        (pre-solved once per job, shell and energy into device memory) or
        "auto" (the table when it fits)."""
        N.check(self._lib.cp3_batch_set_solve_mode(self._h,
                                                   N.SOLVE_MODES[mode]))

    def tuning(self):
        """(cosine block, group waves, workgroups per CU, solve mode) in
        effect."""
        v = [C.c_int() for _ in range(4)]
        N.check(self._lib.cp3_batch_get_tuning(self._h, *[C.byref(x) for x in v]))
        names = {m: k for k, m in N.SOLVE_MODES.items()}
        return v[0].value, v[1].value, v[2].value, names[v[3].value]

    def autotune(self, reps=2):
        nc, g, w = C.c_int(), C.c_int(), C.c_int()
        N.check(self._lib.cp3_batch_autotune(self._h, int(reps), C.byref(nc),
                                             C.byref(g), C.byref(w)))
        return nc.value, g.value, w.value

    def use_graph(self, enable=True):
        N.check(self._lib.cp3_batch_use_graph(self._h, int(bool(enable))))

    def run(self):
        N.check(self._lib.cp3_batch_run(self._h))

    def wait(self):
        N.check(self._lib.cp3_batch_wait(self._h))

    def last_ms(self) -> float:
        v = C.c_float()
        N.check(self._lib.cp3_batch_last_ms(self._h, C.byref(v)))
        return v.value

    def phase_ms(self) -> dict:
        """Milliseconds of the last run's phases: slab-table pre-pass,
        atmospheric propagation, packed tails, beam."""
        v = (C.c_float * 4)()
        N.check(self._lib.cp3_batch_phase_ms(self._h, v))
        return dict(zip(("prepass", "atmos", "tails", "beam"), list(v)))

    def stream(self) -> int:
        v = C.c_void_p()
        N.check(self._lib.cp3_batch_get_stream(self._h, C.byref(v)))
        return v.value or 0

    def stats(self):
        u, e, b = C.c_int64(), C.c_int64(), C.c_int64()
        N.check(self._lib.cp3_batch_stats(self._h, C.byref(u), C.byref(e),
                                          C.byref(b)))
        return u.value, e.value, b.value

    def result(self, job: int) -> np.ndarray:
        """One job's probabilities, [9][nC][nE] or [9][nE], as float64."""
        out = np.zeros(self.shapes[job])
        N.check(self._lib.cp3_batch_copy_result(self._h, int(job), _dptr(out)))
        return out

    def result_device(self, job: int):
        p, el = C.c_void_p(), C.c_int()
        N.check(self._lib.cp3_batch_result_device(self._h, int(job),
                                                  C.byref(p), C.byref(el)))
        return p.value, el.value
