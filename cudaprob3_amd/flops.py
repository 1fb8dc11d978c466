"""Algorithmic work of a launch: FLOPs (split FP64 / FP32) and transcendental
calls of the exact decomposition the kernels execute.

Per-stage counts come from the host build of the kernel arithmetic run with an
operation-counting scalar type (``cp3m_stage_ops`` in libcp3model.so):
add/sub/mul/div/sqrt count 1, a fused multiply-add 2; sincos/acos are counted
apart.  A launch's total follows the kernel's work list: per work unit of
``w`` waves x ``nc`` trajectories and per valid energy lane, one eigen-solve
per shell down to the unit's deepest trajectory plus one atmosphere solve per
wave; per trajectory, M-1 middle shells, one innermost shell (M >= 1) and the
finish.
"""
from __future__ import annotations

import ctypes as C
import json
import os

import numpy as np

from ._native import MODEL_SO

STAGES = ["solve", "middle", "inner", "finish", "beam"]
KEYS = ["f64", "f32", "sincos64", "sincos32", "acos64", "acos32", "sqrt64",
        "sqrt32", "rint64"]
PREC = {"fp64": 0, "fp32": 1, "mixed": 2}
PROFILES = os.path.join(os.path.dirname(os.path.dirname(
    os.path.abspath(__file__))), "profiles")

_model = None


def _lib():
    global _model
    if _model is None:
        if not os.path.exists(MODEL_SO):
            raise ImportError(f"{MODEL_SO} is not built")
        m = C.CDLL(MODEL_SO)
        m.cp3m_stage_ops.argtypes = [C.c_int, C.POINTER(C.c_double)]
        m.cp3m_stage_ops.restype = None
        m.cp3m_max_layers.argtypes = [C.POINTER(C.c_double), C.c_int,
                                      C.POINTER(C.c_double),
                                      C.POINTER(C.c_double),
                                      C.POINTER(C.c_double), C.c_int,
                                      C.POINTER(C.c_int32)]
        m.cp3m_max_layers.restype = C.c_int
        _model = m
    return _model


def stage_ops(precision: str) -> dict:
    out = np.zeros(len(STAGES) * len(KEYS))
    _lib().cp3m_stage_ops(PREC[precision],
                          out.ctypes.data_as(C.POINTER(C.c_double)))
    out = out.reshape(len(STAGES), len(KEYS))
    return {s: dict(zip(KEYS, out[i])) for i, s in enumerate(STAGES)}


def max_layers(cosv, earth) -> np.ndarray:
    c = np.ascontiguousarray(cosv, dtype=np.float64)
    r = np.ascontiguousarray(earth.radii, dtype=np.float64)
    d = np.ascontiguousarray(earth.rhos, dtype=np.float64)
    y = np.ascontiguousarray(earth.yps, dtype=np.float64)
    M = np.zeros(len(c), dtype=np.int32)
    dp = C.POINTER(C.c_double)
    rc = _lib().cp3m_max_layers(c.ctypes.data_as(dp), len(c),
                                r.ctypes.data_as(dp), d.ctypes.data_as(dp),
                                y.ctypes.data_as(dp), len(r),
                                M.ctypes.data_as(C.POINTER(C.c_int32)))
    if rc != 0:
        raise ValueError("bad Earth model")
    return M


def grid_stage_calls(M: np.ndarray, n_e: int, nc: int, w: int = 1) -> dict:
    """Stage calls of one atmospheric job: cosines ordered by decreasing M
    (the kernel's block order), work units of w waves x nc trajectories,
    every valid energy lane: one solve per shell down to the unit's deepest
    trajectory plus one atmosphere (vacuum) solve per unit."""
    Ms = np.sort(M)[::-1]
    per = nc * w
    nunits = (len(Ms) + per - 1) // per
    solves = sum(int(Ms[u * per]) + 1 for u in range(nunits))
    return {
        "solve": solves * n_e,
        "middle": int(np.maximum(Ms - 1, 0).sum()) * n_e,
        "inner": int((Ms >= 1).sum()) * n_e,
        "finish": len(Ms) * n_e,
    }


def table_stage_calls(M: np.ndarray, n_e: int, n_shells: int) -> dict:
    """Stage calls of one atmospheric job in slab-table mode: the pre-pass
    solves every shell of the Earth model and the vacuum once per energy."""
    calls = grid_stage_calls(M, n_e, 1, 1)
    calls["solve"] = (n_shells + 1) * n_e
    return calls


def batch_counts(wl, precision: str, nc: int, w: int = 1,
                 mode: str = "shared") -> dict:
    """Total counted work of one launch over an AtmosBatch."""
    ops = stage_ops(precision)
    tot = {k: 0.0 for k in KEYS}
    cache = {}
    for j in wl.jobs:
        key = (j.grid, len(j.cosv), len(j.ev))
        if key not in cache:
            M = max_layers(j.cosv, wl.earth)
            if mode == "table":
                cache[key] = table_stage_calls(M, len(j.ev), len(wl.earth.radii))
            else:
                cache[key] = grid_stage_calls(M, len(j.ev), nc, w)
        calls = cache[key]
        for st, n in calls.items():
            for k in KEYS:
                tot[k] += n * ops[st][k]
    return tot


def beam_counts(n_points: int, precision: str) -> dict:
    ops = stage_ops(precision)
    return {k: n_points * ops["beam"][k] for k in KEYS}


def measured_traffic(name: str):
    """HBM bytes per launch measured with rocprofv3 counters (FETCH_SIZE,
    WRITE_SIZE), as recorded in profiles/traffic.json, or None."""
    path = os.path.join(PROFILES, "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        data = json.load(f)
    return data.get(name)
