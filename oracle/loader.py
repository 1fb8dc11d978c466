"""ctypes loader for the oracle (test infrastructure only).

Loads ``oracle/liboracle.so`` (the CPU restatement of the reference propagator)
and, where it was built, ``oracle/_ref/libcp3ref.so`` /
``oracle/_ref/libcp3ref_deep.so`` (the reference's own CPU propagators compiled
from its headers).  Only tests, the smoke check and the benchmark's CPU-baseline
leg use this module; it is the checker, never the thing measured or shipped.

All atmospheric results are returned in the reference CPU layout
``[cos][E][9]`` (physics.hpp:1142-1144), beam results as ``[E][9]``; channel
``a*3+b`` is P(a->b) with e, mu, tau = 0, 1, 2 (types.hpp:23-27).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# CP3_ORACLE_SO selects another build of the oracle (the sanitizer build,
# liboracle_asan.so)
ORACLE_SO = os.environ.get("CP3_ORACLE_SO") or os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libcp3ref.so")
REF_DEEP_SO = os.path.join(HERE, "_ref", "libcp3ref_deep.so")

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)


def _ptr(a, t=_dp):
    return a.ctypes.data_as(t)


def _atmos_argtypes(outp):
    return [C.c_int, _dp, C.c_int, _dp, C.c_int, _dp, _dp, _dp, C.c_int,
            C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
            C.c_double, C.c_double, C.c_double, C.c_int, outp, C.c_char_p,
            C.c_int]


def _atmos_file_argtypes(outp):
    return [C.c_int, _dp, C.c_int, _dp, C.c_int, C.c_char_p, C.c_double,
            C.c_double, C.c_double, C.c_double, C.c_int, C.c_double,
            C.c_double, C.c_double, C.c_int, outp, C.c_char_p, C.c_int]


def _atmos_poly_argtypes(outp):
    return [C.c_int, _dp, C.c_int, _dp, C.c_int, _dp, _dp, _dp, _dp, _dp,
            C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
            C.c_double, C.c_double, C.c_double, C.c_int, outp, C.c_char_p,
            C.c_int]


def _atmos_pha_argtypes(oracle: bool):
    return ([C.c_int, _dp, C.c_int, _dp, C.c_int, _dp, _dp, _dp, C.c_int,
             C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
             C.c_double, C.c_double, C.c_double, C.c_int, _dp, _dp] +
            ([C.c_int] if oracle else []) + [C.c_int, _dp, C.c_char_p, C.c_int])


def _atmos_poly_pha_argtypes(oracle: bool):
    return ([C.c_int, _dp, C.c_int, _dp, C.c_int, _dp, _dp, _dp, _dp, _dp, C.c_int,
             C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
             C.c_double, C.c_double, C.c_double, C.c_int, _dp, _dp] +
            ([C.c_int] if oracle else []) + [C.c_int, _dp, C.c_char_p, C.c_int])


def _beam_argtypes(outp):
    return [C.c_int, _dp, C.c_int, C.c_double, C.c_double, C.c_double,
            C.c_double, C.c_double, C.c_double, C.c_int, C.c_double,
            C.c_double, C.c_int, outp, C.c_char_p, C.c_int]


@dataclass
class Osc:
    """Oscillation parameters in the reference's setter units: radians and eV^2.

    ``dm23`` is the reference's ``dm23sq`` argument, i.e. Delta m^2_32
    (propagator.hpp:188-193).  ``knutype`` is setMNSMatrix's kNuType
    (delta is negated when it is negative, propagator.hpp:150-153).
    """

    th12: float
    th13: float
    th23: float
    dcp: float
    dm12: float
    dm23: float
    knutype: int = 1

    def args(self):
        return (self.th12, self.th13, self.th23, self.dcp, int(self.knutype),
                self.dm12, self.dm23)


class PropLib:
    """One of the CPU libraries (prefix ``orc_`` or ``ref_``)."""

    def __init__(self, path: str, prefix: str):
        self.path = path
        self.prefix = prefix
        self.lib = C.CDLL(path)
        L = self.lib
        self._atmos = getattr(L, prefix + "atmos_f64")
        self._atmos.argtypes = _atmos_argtypes(_dp)
        self._atmos.restype = C.c_int
        self._atmos_file = getattr(L, prefix + "atmos_file_f64")
        self._atmos_file.argtypes = _atmos_file_argtypes(_dp)
        self._atmos_file.restype = C.c_int
        self._atmos_poly = getattr(L, prefix + "atmos_poly_f64")
        self._atmos_poly.argtypes = _atmos_poly_argtypes(_dp)
        self._atmos_poly.restype = C.c_int
        self._atmos_pha = getattr(L, prefix + "atmos_pha_f64")
        self._atmos_pha.argtypes = _atmos_pha_argtypes(prefix == "orc_")
        self._atmos_pha.restype = C.c_int
        self._atmos_poly_pha = getattr(L, prefix + "atmos_poly_pha_f64")
        self._atmos_poly_pha.argtypes = _atmos_poly_pha_argtypes(prefix == "orc_")
        self._atmos_poly_pha.restype = C.c_int
        self._beam = getattr(L, prefix + "beam_f64")
        self._beam.argtypes = _beam_argtypes(_dp)
        self._beam.restype = C.c_int
        self._atmos32 = getattr(L, prefix + "atmos_f32", None)
        if self._atmos32 is not None:
            self._atmos32.argtypes = _atmos_argtypes(_fp)
            self._atmos32.restype = C.c_int
        self._beam32 = getattr(L, prefix + "beam_f32", None)
        if self._beam32 is not None:
            self._beam32.argtypes = _beam_argtypes(_fp)
            self._beam32.restype = C.c_int

    def _check(self, rc, err):
        if rc != 0:
            raise RuntimeError(err.value.decode(errors="replace"))

    def atmos(self, nutype: int, cosv, ev, radii, rhos, yps, osc: Osc,
              h_km: float, threads: int = 1, f32: bool = False):
        cosv = np.ascontiguousarray(cosv, dtype=np.float64)
        ev = np.ascontiguousarray(ev, dtype=np.float64)
        radii = np.ascontiguousarray(radii, dtype=np.float64)
        rhos = np.ascontiguousarray(rhos, dtype=np.float64)
        yps = np.ascontiguousarray(yps, dtype=np.float64)
        err = C.create_string_buffer(512)
        if f32:
            out = np.zeros((len(cosv), len(ev), 9), dtype=np.float32)
            fn, op = self._atmos32, _fp
        else:
            out = np.zeros((len(cosv), len(ev), 9), dtype=np.float64)
            fn, op = self._atmos, _dp
        rc = fn(int(nutype), _ptr(cosv), len(cosv), _ptr(ev), len(ev),
                _ptr(radii), _ptr(rhos), _ptr(yps), len(radii), *osc.args(),
                float(h_km), int(threads), _ptr(out, op), err, 512)
        self._check(rc, err)
        return out

    def atmos_poly(self, nutype: int, cosv, ev, radii, a, b, c, yps, osc: Osc,
                   h_km: float, threads: int = 1):
        """setDensity(radii, a, b, c, yps): row densities a R^2 + b R + c
        (R in km) averaged per trajectory segment."""
        arrs = [np.ascontiguousarray(x, dtype=np.float64)
                for x in (cosv, ev, radii, a, b, c, yps)]
        cosv, ev, radii, a, b, c, yps = arrs
        err = C.create_string_buffer(512)
        out = np.zeros((len(cosv), len(ev), 9), dtype=np.float64)
        rc = self._atmos_poly(int(nutype), _ptr(cosv), len(cosv), _ptr(ev),
                              len(ev), _ptr(radii), _ptr(a), _ptr(b), _ptr(c),
                              _ptr(yps), len(radii), *osc.args(), float(h_km),
                              int(threads), _ptr(out), err, 512)
        self._check(rc, err)
        return out

    def atmos_pha(self, nutype: int, cosv, ev, radii, rhos, yps, osc: Osc,
                  h_km: float, prob, edges_km, threads: int = 1,
                  int_abs_sinc: bool = False):
        """Production-height averaging (physics.hpp:997-1067): prob is
        [2][3][nE][nC][nb] (neutrino type, initial flavour, energy, cosine,
        bin), edges_km the nb + 1 bin edges.  int_abs_sinc (oracle only)
        reproduces the compiled reference's sinc, whose abs() truncates to an
        integer (math.hpp:50)."""
        arrs = [np.ascontiguousarray(x, dtype=np.float64)
                for x in (cosv, ev, radii, rhos, yps, prob, edges_km)]
        cosv, ev, radii, rhos, yps, prob, edges_km = arrs
        nb = len(edges_km) - 1
        if prob.size != nb * 2 * 3 * len(ev) * len(cosv):
            raise ValueError("prob must hold 2 * 3 * nE * nC * nb values")
        err = C.create_string_buffer(512)
        out = np.zeros((len(cosv), len(ev), 9), dtype=np.float64)
        extra = [int(bool(int_abs_sinc))] if self.prefix == "orc_" else []
        rc = self._atmos_pha(int(nutype), _ptr(cosv), len(cosv), _ptr(ev), len(ev),
                             _ptr(radii), _ptr(rhos), _ptr(yps), len(radii),
                             *osc.args(), float(h_km), nb, _ptr(prob),
                             _ptr(edges_km), *extra, int(threads), _ptr(out),
                             err, 512)
        self._check(rc, err)
        return out

    def atmos_poly_pha(self, nutype: int, cosv, ev, radii, a, b, c, yps, osc: Osc,
                       h_km: float, prob, edges_km, threads: int = 1,
                       int_abs_sinc: bool = False):
        """Production-height averaging through a polynomial model."""
        arrs = [np.ascontiguousarray(x, dtype=np.float64)
                for x in (cosv, ev, radii, a, b, c, yps, prob, edges_km)]
        cosv, ev, radii, a, b, c, yps, prob, edges_km = arrs
        nb = len(edges_km) - 1
        if prob.size != nb * 2 * 3 * len(ev) * len(cosv):
            raise ValueError("prob must hold 2 * 3 * nE * nC * nb values")
        err = C.create_string_buffer(512)
        out = np.zeros((len(cosv), len(ev), 9), dtype=np.float64)
        extra = [int(bool(int_abs_sinc))] if self.prefix == "orc_" else []
        rc = self._atmos_poly_pha(int(nutype), _ptr(cosv), len(cosv), _ptr(ev), len(ev),
                                  _ptr(radii), _ptr(a), _ptr(b), _ptr(c), _ptr(yps),
                                  len(radii), *osc.args(), float(h_km), nb, _ptr(prob),
                                  _ptr(edges_km), *extra, int(threads), _ptr(out),
                                  err, 512)
        self._check(rc, err)
        return out

    def atmos_file(self, nutype: int, cosv, ev, path: str, osc: Osc,
                   h_km: float, threads: int = 1):
        cosv = np.ascontiguousarray(cosv, dtype=np.float64)
        ev = np.ascontiguousarray(ev, dtype=np.float64)
        err = C.create_string_buffer(512)
        out = np.zeros((len(cosv), len(ev), 9), dtype=np.float64)
        rc = self._atmos_file(int(nutype), _ptr(cosv), len(cosv), _ptr(ev),
                              len(ev), path.encode(), *osc.args(),
                              float(h_km), int(threads), _ptr(out), err, 512)
        self._check(rc, err)
        return out

    def beam(self, nutype: int, ev, rho: float, L_km: float, osc: Osc,
             threads: int = 1, f32: bool = False):
        ev = np.ascontiguousarray(ev, dtype=np.float64)
        err = C.create_string_buffer(512)
        if f32:
            out = np.zeros((len(ev), 9), dtype=np.float32)
            fn, op = self._beam32, _fp
        else:
            out = np.zeros((len(ev), 9), dtype=np.float64)
            fn, op = self._beam, _dp
        rc = fn(int(nutype), _ptr(ev), len(ev), float(rho), float(L_km),
                *osc.args(), int(threads), _ptr(out, op), err, 512)
        self._check(rc, err)
        return out


class Oracle(PropLib):
    """The restatement, plus its single-function probes."""

    def __init__(self, path: str = ORACLE_SO):
        super().__init__(path, "orc_")
        L = self.lib
        L.orc_pmns.argtypes = [C.c_double] * 4 + [C.c_int, _dp]
        L.orc_poly_rye.argtypes = ([_dp, C.c_int] + [_dp] * 5 +
                                   [C.c_int, C.c_double, _dp, C.c_char_p, C.c_int])
        L.orc_poly_rye.restype = C.c_int
        L.orc_pmns.restype = None
        L.orc_matter_eigs.argtypes = ([C.c_double] * 4 + [C.c_int] +
                                      [C.c_double] * 4 + [C.c_int, _dp, _ip])
        L.orc_matter_eigs.restype = None
        L.orc_layer.argtypes = ([C.c_double] * 4 + [C.c_int] +
                                [C.c_double] * 5 + [C.c_int, _dp, _dp])
        L.orc_layer.restype = None
        L.orc_segments.argtypes = [C.c_double, _dp, _dp, _dp, C.c_int,
                                   C.c_double, _dp, _dp, C.c_int]
        L.orc_segments.restype = C.c_int

    def poly_rye(self, cosv, radii, a, b, c, yps, h_km):
        """rho*Ye of each (cosine, layer 1..rows) of a polynomial model as
        the propagation uses it ([nC][rows], 0 past the innermost layer)."""
        arrs = [np.ascontiguousarray(x, dtype=np.float64)
                for x in (cosv, radii, a, b, c, yps)]
        cosv, radii, a, b, c, yps = arrs
        out = np.zeros((len(cosv), len(radii)))
        err = C.create_string_buffer(512)
        rc = self.lib.orc_poly_rye(_ptr(cosv), len(cosv), _ptr(radii), _ptr(a),
                                   _ptr(b), _ptr(c), _ptr(yps), len(radii),
                                   float(h_km), _ptr(out), err, 512)
        self._check(rc, err)
        return out

    def pmns(self, th12, th13, th23, dcp, knutype=1):
        out = np.zeros(18)
        self.lib.orc_pmns(th12, th13, th23, dcp, knutype, _ptr(out))
        return (out[0::2] + 1j * out[1::2]).reshape(3, 3)

    def matter_eigs(self, osc: Osc, E, rho_ye, nutype):
        lam = np.zeros(3)
        order = np.zeros(3, dtype=np.int32)
        self.lib.orc_matter_eigs(osc.th12, osc.th13, osc.th23, osc.dcp,
                                 osc.knutype, osc.dm12, osc.dm23, E, rho_ye,
                                 nutype, _ptr(lam), _ptr(order, _ip))
        return lam, order

    def layer(self, osc: Osc, E, rho_ye, L_km, nutype):
        T = np.zeros(18)
        P = np.zeros(54)
        self.lib.orc_layer(osc.th12, osc.th13, osc.th23, osc.dcp, osc.knutype,
                           osc.dm12, osc.dm23, E, rho_ye, L_km, nutype,
                           _ptr(T), _ptr(P))
        Tc = (T[0::2] + 1j * T[1::2]).reshape(3, 3)
        Pc = (P[0::2] + 1j * P[1::2]).reshape(3, 3, 3)
        return Tc, Pc

    def segments(self, cz, radii, rhos, yps, h_km, cap=256):
        radii = np.ascontiguousarray(radii, dtype=np.float64)
        rhos = np.ascontiguousarray(rhos, dtype=np.float64)
        yps = np.ascontiguousarray(yps, dtype=np.float64)
        seg = np.zeros(cap)
        rye = np.zeros(cap)
        M = self.lib.orc_segments(cz, _ptr(radii), _ptr(rhos), _ptr(yps),
                                  len(radii), h_km, _ptr(seg), _ptr(rye), cap)
        if M < 0:
            raise RuntimeError("bad geometry input")
        return M, seg[:M + 1].copy(), rye[:M + 1].copy()


def load_oracle() -> Oracle:
    if not os.path.exists(ORACLE_SO):
        raise FileNotFoundError(
            f"{ORACLE_SO} is not built; run `make -C oracle` first")
    return Oracle()


def load_ref(deep: bool = False):
    """The compiled reference, or None when it was not built here."""
    path = REF_DEEP_SO if deep else REF_SO
    if not os.path.exists(path):
        return None
    return PropLib(path, "ref_")
