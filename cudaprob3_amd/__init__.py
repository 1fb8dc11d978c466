"""MI355X-native three-flavour neutrino oscillation propagator (Barger
method) with the CUDAProb3++ propagator interface.

The compute path is ``libcp3hip.so`` (HIP kernels for gfx950 behind the C ABI
in ``include/cp3_hip.h``); the Python classes here are a thin view of it.
"""
from .propagator import (AtmosPropagator, BeamPropagator, Cp3Error,  # noqa: F401
                         Neutrino, Antineutrino)

__all__ = ["AtmosPropagator", "BeamPropagator", "Cp3Error", "Neutrino",
           "Antineutrino"]
