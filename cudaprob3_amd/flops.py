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


def _zero():
    return {k: 0.0 for k in KEYS}


def batch_phase_counts(wl, precision: str, nc: int, w: int = 1,
                       mode: str = "shared") -> dict:
    """Counted work of each phase of a batch launch over an AtmosBatch:
    "prepass" (slab-table solves), "atmos" (the jobs' own energy tiles) and
    "tails" (the last nE % 64 energies of jobs sharing a grid, when two or
    more such tails fit in a wave), following the batch engine's plan."""
    ops = stage_ops(precision)
    res = {"prepass": _zero(), "atmos": _zero(), "tails": _zero()}
    per_grid = {}
    for j in wl.jobs:
        per_grid[j.grid] = per_grid.get(j.grid, 0) + 1
    n_shells = len(wl.earth.radii)
    cache = {}

    def add(phase, calls):
        for st, n in calls.items():
            for k in KEYS:
                res[phase][k] += n * ops[st][k]

    for j in wl.jobs:
        n_e = len(j.ev)
        tail = n_e % 64
        packed = tail and 64 // tail >= 2 and per_grid[j.grid] >= 2
        t = tail if packed else 0
        key = (j.grid, len(j.cosv), n_e)
        if key not in cache:
            M = max_layers(j.cosv, wl.earth)
            if mode == "table":
                main = grid_stage_calls(M, n_e - t, 1, 1)
                tl = grid_stage_calls(M, t, 1, 1)
                main.pop("solve")
                tl.pop("solve")
                cache[key] = (main, tl, {"solve": (n_shells + 1) * n_e})
            else:
                cache[key] = (grid_stage_calls(M, n_e - t, nc, w),
                              grid_stage_calls(M, t, nc, w), {})
        main, tl, pre = cache[key]
        add("atmos", main)
        add("tails", tl)
        add("prepass", pre)
    return res


def batch_counts(wl, precision: str, nc: int, w: int = 1,
                 mode: str = "shared") -> dict:
    """Total counted work of one launch over an AtmosBatch."""
    ph = batch_phase_counts(wl, precision, nc, w, mode)
    return {k: sum(ph[p][k] for p in ph) for k in KEYS}


def beam_counts(n_points: int, precision: str) -> dict:
    ops = stage_ops(precision)
    return {k: n_points * ops["beam"][k] for k in KEYS}


def measured_valu(name: str):
    """The VALU issue record of a kernel, or None: its occupancy (compiler
    resource usage) and the per-form issue rates of a probe of independent
    instruction chains, from profiles/valu_counters.json, with the kernel's
    instruction counts per launch (rocprofv3 SQ_INSTS_VALU and the per-type
    counters) from profiles/valu_flops.json."""
    path = os.path.join(PROFILES, "valu_counters.json")
    flops = measured_flops(name)
    if not os.path.exists(path) or flops is None:
        return None
    with open(path) as f:
        data = json.load(f)
    occ = data["kernels"].get(name)
    if occ is None:
        return None
    return {"occupancy": occ, "probe": data["probe"],
            "probe_per_clock": data.get("probe_per_clock"),
            "counters": flops["counters_per_launch"]}


def _rate_at(table: dict, waves: int) -> float:
    """The probe rate measured at the largest wave count not above `waves`."""
    keys = sorted(int(k) for k in table)
    return table[str(max([k for k in keys if k <= waves] or keys[:1]))]


def issue_ceiling(rec, per_clock: bool = False) -> dict:
    """Wave-instructions per CU-cycle the kernel's instruction mix allows at
    its occupancy, from the probe's per-form rates: FP32 arithmetic at the
    rate of two-operand FP32 forms with VGPR operands, FP64 arithmetic and
    conversions at the FP64 rate, everything else at the FP32 rate.
    per_clock: the probe's rates per cycle of the clock it ran at (measured
    with SQ_BUSY_CYCLES), for a kernel rate taken at its own measured clock;
    otherwise per cycle of the nominal 2.4 GHz."""
    w = rec["occupancy"]["waves_per_simd"]
    if per_clock:
        p = rec["probe_per_clock"]
        r32 = _rate_at(p["fp32_fmac_16_chains"], w)
        r64 = _rate_at(p["fp64_fmac_8_chains"], w)
    else:
        p = rec["probe"]
        r32 = p["fp32_vop2_vgpr_operands"].get(str(w)) or p["fp32_vop2_vgpr_operands"]["4"]
        r64 = p["fp64_fmac"].get(str(w)) or p["fp64_fmac"]["4"]
    c = rec["counters"]
    f32 = c["ADD_F32"] + c["MUL_F32"] + c["FMA_F32"]
    f64 = c["ADD_F64"] + c["MUL_F64"] + c["FMA_F64"] + c["TRANS_F64"] + c["CVT"]
    rest = max(c["VALU"] - f32 - f64, 0.0)
    cycles = (f32 + rest) / r32 + f64 / r64
    return {"ceiling": c["VALU"] / cycles, "fp32_rate": r32, "fp64_rate": r64,
            "waves_per_simd": rec["occupancy"]["waves_per_simd"]}


def measured_traffic(name: str):
    """HBM bytes per launch measured with rocprofv3 counters (FETCH_SIZE,
    WRITE_SIZE), as recorded in profiles/traffic.json, or None."""
    path = os.path.join(PROFILES, "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        data = json.load(f)
    return data.get(name)


def measured_flops(name: str):
    """FLOPs per launch from rocprofv3's VALU instruction counters
    (64 x (ADD + MUL + 2 FMA) of SQ_INSTS_VALU_{ADD,MUL,FMA}_{F32,F64}, plus
    the hardware's own SQ_INSTS_VALU_FLOPS_FP32/FP64 and the TRANS counts),
    as recorded in profiles/valu_flops.json, or None.  Unlike the counted
    FLOPs above, these include the phase polynomials' multiply-adds (counted
    above as sincos calls) and exclude work the compiler shares between a
    lane's trajectories."""
    path = os.path.join(PROFILES, "valu_flops.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        data = json.load(f)
    return data.get(name)


# FLOPs of one layer-phase evaluation e^{2 pi i f} - 1 (cp3_math.hpp
# m_cispi_m1), which the counts above record as one sincos call: z = f*f,
# Horner chains of degree 9 and 10 in f (4 + 5 fused multiply-adds) in FP32 or
# 15 and 16 (7 + 8) in FP64, then a = pa f, wr = -(a a), wi = a b.
PHASE_POLY_FLOP = {"f32": 1 + 2 * 4 + 2 * 5 + 3, "f64": 1 + 2 * 7 + 2 * 8 + 3}


def with_phase_polynomials(counts: dict) -> dict:
    """Counted FLOPs of a propagation kernel with no eigen-solve inside (the
    slab-table kernels), its phase polynomials included: every sincos call
    there is a layer-phase evaluation.  This is the figure the VALU FLOP
    counters measure."""
    return {"f64": counts["f64"] + PHASE_POLY_FLOP["f64"] * counts["sincos64"],
            "f32": counts["f32"] + PHASE_POLY_FLOP["f32"] * counts["sincos32"]}
