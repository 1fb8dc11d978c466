"""Python view of the propagators, with the reference's method names.

``AtmosPropagator`` stands for ``AtmosCudaPropagatorSingle`` /
``AtmosCpuPropagator<T>`` (atmoscudapropagator.cuh:42-145,
atmoscpupropagator.hpp:35-547) and ``BeamPropagator`` for
``BeamCudaPropagatorSingle`` (beamcudapropagator.cuh:41-121).  Every call goes
through the C ABI of ``libcp3hip.so``; misuse raises :class:`Cp3Error` where
the reference throws ``std::runtime_error``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from ._native import Cp3Error

# types.hpp:23-29
e_e, e_m, e_t, m_e, m_m, m_t, t_e, t_m, t_t = range(9)
Neutrino, Antineutrino = 0, 1


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _f64(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64).ravel())


class AtmosPropagator:
    """One atmospheric grid on one GPU (precision "fp64", "fp32" or "mixed")."""

    def __init__(self, n_cosines: int, n_energies: int, device: int = 0,
                 precision: str = "fp64"):
        self._lib = N.load()
        self._h = C.c_void_p()
        self.n_cosines = int(n_cosines)
        self.n_energies = int(n_energies)
        self.precision = precision
        N.check(self._lib.cp3_atmos_create(int(device), self.n_cosines,
                                           self.n_energies,
                                           N.PRECISIONS[precision],
                                           C.byref(self._h)))

    def close(self):
        if self._h:
            N.check(self._lib.cp3_atmos_destroy(self._h))
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

    # setters (propagator.hpp:143-218, atmoscpupropagator.hpp:116-393)
    def setEnergyList(self, energies):
        a = _f64(energies)
        N.check(self._lib.cp3_atmos_set_energy_list(self._h, _dptr(a), len(a)))

    def setCosineList(self, cosines):
        a = _f64(cosines)
        N.check(self._lib.cp3_atmos_set_cosine_list(self._h, _dptr(a), len(a)))

    def setMNSMatrix(self, theta12, theta13, theta23, dcp, kNuType=1):
        N.check(self._lib.cp3_atmos_set_mns_matrix(
            self._h, float(theta12), float(theta13), float(theta23),
            float(dcp), int(kNuType)))

    def setNeutrinoMasses(self, dm12sq, dm23sq):
        N.check(self._lib.cp3_atmos_set_neutrino_masses(
            self._h, float(dm12sq), float(dm23sq)))

    def setDensity(self, radii, *cols):
        """setDensity(radii, rhos, yps) or, for a polynomial model,
        setDensity(radii, a, b, c, yps) (density a R^2 + b R + c per shell,
        R in km, averaged per trajectory segment)."""
        if len(cols) == 4:
            r, a, b, c, y = [_f64(x) for x in (radii,) + cols]
            if len(a) != len(r) or len(a) != len(y):
                msg = ("setDensity : a.size() != radii.size()"
                       if len(a) != len(r) else
                       "setDensity : a.size() != yps.size()")
                raise Cp3Error(N.CP3_ERR_INVALID, msg)
            if len(b) != len(r) or len(c) != len(r):
                raise Cp3Error(N.CP3_ERR_INVALID, "setDensity : b.size() and "
                               "c.size() must equal radii.size()")
            N.check(self._lib.cp3_atmos_set_density_poly(
                self._h, _dptr(r), _dptr(a), _dptr(b), _dptr(c), _dptr(y), len(r)))
            return
        if len(cols) != 2:
            raise TypeError("setDensity(radii, rhos, yps) or "
                            "setDensity(radii, a, b, c, yps)")
        rhos, yps = cols
        r, d, y = _f64(radii), _f64(rhos), _f64(yps)
        if not (len(r) == len(d) == len(y)):
            # the C ABI takes one length; mirror the reference's size checks
            msg = ("setDensity : rhos.size() != radii.size()"
                   if len(r) != len(d) else
                   "setDensity : rhos.size() != yps.size()")
            raise Cp3Error(N.CP3_ERR_INVALID, msg)
        N.check(self._lib.cp3_atmos_set_density(self._h, _dptr(r), _dptr(d),
                                                _dptr(y), len(r)))

    def setDensityFromFile(self, path: str):
        N.check(self._lib.cp3_atmos_set_density_from_file(self._h,
                                                          path.encode()))

    def setChemicalComposition(self, yps):
        y = _f64(yps)
        N.check(self._lib.cp3_atmos_set_chemical_composition(self._h,
                                                             _dptr(y), len(y)))

    def setProductionHeight(self, height_km):
        N.check(self._lib.cp3_atmos_set_production_height(self._h,
                                                          float(height_km)))

    def SetNumberOfProductionHeightBinsForAveraging(self, n_bins: int):
        """At most 20 bins; like the reference (atmoscpupropagator.hpp:88-108)
        any n >= 1 turns averaging off."""
        N.check(self._lib.cp3_atmos_set_production_height_bins(self._h, int(n_bins)))

    def setProductionHeightAveraging(self, on: bool = True):
        """No counterpart in the reference class, which keeps this flag
        protected (atmoscpupropagator.hpp:532) and never sets it: turns the
        averaging of physics.hpp:997-1067 on or off."""
        N.check(self._lib.cp3_atmos_set_production_height_averaging(self._h,
                                                                    int(bool(on))))

    def setProductionHeightList(self, list_prob, list_bins):
        """Bin probabilities [nutype][flavour][energy][cosine][bin] and the
        n_bins + 1 bin edges in km (atmoscpupropagator.hpp:395-427)."""
        p, b = _f64(list_prob), _f64(list_bins)
        N.check(self._lib.cp3_atmos_set_production_height_list(
            self._h, _dptr(p), len(p), _dptr(b), len(b)))

    def getNlayerBoundaries(self) -> int:
        n = C.c_int()
        N.check(self._lib.cp3_atmos_get_nlayer_boundaries(self._h, C.byref(n)))
        return n.value

    def setTuning(self, cosine_block: int = 0, group_waves: int = 0):
        """Trajectories per lane and waves sharing each eigen-solve (0 =
        automatic); results do not depend on them."""
        N.check(self._lib.cp3_atmos_set_tuning(self._h, int(cosine_block),
                                               int(group_waves)))

    def setSolveMode(self, mode: str = "auto"):
        """Where the eigen-solutions come from: "table" (solved once per
        shell and energy into device memory by a pre-pass), "shared" (solved
        in the propagation kernel, shared through LDS) or "auto" (the
        table); results do not depend on it."""
        N.check(self._lib.cp3_atmos_set_solve_mode(self._h,
                                                   N.SOLVE_MODES[mode]))

    # calculation and results
    def calculateProbabilities(self, nutype=Neutrino):
        N.check(self._lib.cp3_atmos_calculate(self._h, int(nutype)))

    def calculateProbabilitiesAsync(self, nutype=Neutrino):
        N.check(self._lib.cp3_atmos_calculate_async(self._h, int(nutype)))

    def waitForCompletion(self):
        N.check(self._lib.cp3_atmos_wait(self._h))

    def getProbability(self, index_cosine, index_energy, t) -> float:
        v = C.c_double()
        N.check(self._lib.cp3_atmos_get_probability(
            self._h, int(index_cosine), int(index_energy), int(t), C.byref(v)))
        return v.value

    def getProbabilityArr(self, t) -> np.ndarray:
        out = np.zeros(self.n_cosines * self.n_energies)
        N.check(self._lib.cp3_atmos_get_probability_arr(self._h, _dptr(out),
                                                        int(t)))
        return out

    def lastKernelMs(self) -> float:
        v = C.c_float()
        N.check(self._lib.cp3_atmos_last_kernel_ms(self._h, C.byref(v)))
        return v.value

    def probabilities(self) -> np.ndarray:
        """All channels as [cos][E][9] (the reference CPU layout)."""
        chans = [self.getProbabilityArr(t).reshape(self.n_energies,
                                                   self.n_cosines).T
                 for t in range(9)]
        return np.stack(chans, axis=-1)


class BeamPropagator:
    """One beam energy list on one GPU."""

    def __init__(self, n_energies: int, device: int = 0,
                 precision: str = "fp64"):
        self._lib = N.load()
        self._h = C.c_void_p()
        self.n_energies = int(n_energies)
        self.precision = precision
        N.check(self._lib.cp3_beam_create(int(device), self.n_energies,
                                          N.PRECISIONS[precision],
                                          C.byref(self._h)))

    def close(self):
        if self._h:
            N.check(self._lib.cp3_beam_destroy(self._h))
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

    def setEnergyList(self, energies):
        a = _f64(energies)
        N.check(self._lib.cp3_beam_set_energy_list(self._h, _dptr(a), len(a)))

    def setMNSMatrix(self, theta12, theta13, theta23, dcp, kNuType=1):
        N.check(self._lib.cp3_beam_set_mns_matrix(
            self._h, float(theta12), float(theta13), float(theta23),
            float(dcp), int(kNuType)))

    def setNeutrinoMasses(self, dm12sq, dm23sq):
        N.check(self._lib.cp3_beam_set_neutrino_masses(
            self._h, float(dm12sq), float(dm23sq)))

    def setDensity(self, rho):
        N.check(self._lib.cp3_beam_set_density(self._h, float(rho)))

    def setPathLength(self, length_km):
        N.check(self._lib.cp3_beam_set_path_length(self._h, float(length_km)))

    def calculateProbabilities(self, nutype=Neutrino):
        N.check(self._lib.cp3_beam_calculate(self._h, int(nutype)))

    def calculateProbabilitiesAsync(self, nutype=Neutrino):
        N.check(self._lib.cp3_beam_calculate_async(self._h, int(nutype)))

    def waitForCompletion(self):
        N.check(self._lib.cp3_beam_wait(self._h))

    def getProbability(self, index_energy, t) -> float:
        v = C.c_double()
        N.check(self._lib.cp3_beam_get_probability(
            self._h, int(index_energy), int(t), C.byref(v)))
        return v.value

    def getProbabilityArr(self, t) -> np.ndarray:
        out = np.zeros(self.n_energies)
        N.check(self._lib.cp3_beam_get_probability_arr(self._h, _dptr(out),
                                                       int(t)))
        return out

    def lastKernelMs(self) -> float:
        v = C.c_float()
        N.check(self._lib.cp3_beam_last_kernel_ms(self._h, C.byref(v)))
        return v.value

    def probabilities(self) -> np.ndarray:
        """All channels as [E][9]."""
        return np.stack([self.getProbabilityArr(t) for t in range(9)], axis=-1)
