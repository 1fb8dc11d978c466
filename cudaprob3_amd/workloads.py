"""Synthetic inputs for the propagation workloads (Earth models, grids and
oscillation-parameter sets), shared by the tests, the golden generator and the
benchmark.

Units follow the reference setters: radii in km, densities in g/cm^3, energies
in GeV, angles in radians, mass splittings in eV^2 (propagator.hpp:143-218,
atmoscpupropagator.hpp:109-170).  ``dm23`` is the reference's ``dm23sq``
argument, i.e. Delta m^2_32.

Workloads:
  * ``prem4``      - 4 shells (outer radii 1220/3480/5701/6371 km, densities
                     13.0/11.3/5.0/3.3, Ye 0.5), 100 x 100 grid, cos in [-1, 0),
                     E log-spaced over 0.1-100 GeV, production height 22 km.
  * ``shells12``   - 12 seeded shells, 2000 x 2000 grid (precision study).
  * ``beam4096``   - one 2.6 g/cm^3 layer, 1300 km, 1e5 energies over
                     0.1-20 GeV, 4096 seeded parameter sets.
  * ``atmos1024``  - 59 seeded shells, 1024 seeded parameter sets, half on
                     400 x 400 and half on 200 x 200 grids over 0.5-50 GeV,
                     neutrino and antineutrino.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np

R_EARTH_KM = 6371.0

# Nominal parameters: sin^2 th12 = 0.307, sin^2 th13 = 0.0220,
# sin^2 th23 = 0.546, dcp = -1.6, dm21^2 = 7.53e-5, dm32^2 = 2.494e-3.
TH12 = float(np.arcsin(np.sqrt(0.307)))
TH13 = float(np.arcsin(np.sqrt(0.0220)))
TH23 = float(np.arcsin(np.sqrt(0.546)))
DCP = -1.6
DM12 = 7.53e-5
DM23 = 2.494e-3
PROD_HEIGHT_KM = 22.0

DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


@dataclass(frozen=True)
class ParamSet:
    """One oscillation-parameter set (an MCMC walker position)."""

    th12: float = TH12
    th13: float = TH13
    th23: float = TH23
    dcp: float = DCP
    dm12: float = DM12
    dm23: float = DM23

    def for_nutype(self, nutype: int):
        """(th12, th13, th23, dcp, kNuType, dm12, dm23) as a caller passes
        them to setMNSMatrix/setNeutrinoMasses for neutrinos (kNuType = +1)
        or antineutrinos (kNuType = -1, which negates delta,
        propagator.hpp:150-153)."""
        return (self.th12, self.th13, self.th23, self.dcp,
                -1 if nutype else 1, self.dm12, self.dm23)


NOMINAL = ParamSet()


@dataclass
class EarthModel:
    """Rows ascending by radius with the 0-km row first, as in
    models/PREM_4layer.dat; the density on a row belongs to the shell whose
    outer radius is that row (atmoscpupropagator.hpp:109-112)."""

    radii: np.ndarray
    rhos: np.ndarray
    yps: np.ndarray

    @property
    def n_shells(self) -> int:
        return len(self.radii) - 1

    def write(self, path: str):
        with open(path, "w") as f:
            for r, d, y in zip(self.radii, self.rhos, self.yps):
                f.write(f"{float(r)!r} {float(d)!r} {float(y)!r}\n")

    @staticmethod
    def read(path: str) -> "EarthModel":
        rows = []
        with open(path) as f:
            for line in f:
                line = line.strip()
                if not line or line.startswith("#"):
                    continue
                rows.append([float(x) for x in line.split()])
        a = np.asarray(rows, dtype=np.float64)
        return EarthModel(a[:, 0].copy(), a[:, 1].copy(), a[:, 2].copy())


def prem4() -> EarthModel:
    return EarthModel(np.array([0.0, 1220.0, 3480.0, 5701.0, 6371.0]),
                      np.array([13.0, 13.0, 11.3, 5.0, 3.3]),
                      np.full(5, 0.5))


@dataclass
class PolyEarthModel:
    """Rows ascending by radius with the 0-km row first; the shell whose outer
    radius is row i has density a[i] R^2 + b[i] R + c[i] g/cm^3 at radius R km,
    averaged along each trajectory segment (setDensity(radii, a, b, c, yps),
    atmoscpupropagator.hpp:181-245; physics.hpp:653-739)."""

    radii: np.ndarray
    a: np.ndarray
    b: np.ndarray
    c: np.ndarray
    yps: np.ndarray

    def write(self, path: str):
        """The 5-column density-file format (radius, a, b, c, Ye)."""
        with open(path, "w") as f:
            for row in zip(self.radii, self.a, self.b, self.c, self.yps):
                f.write(" ".join(repr(float(x)) for x in row) + "\n")

    def rho(self, shell: int, r_km):
        """Density of row `shell` at radius r_km."""
        return self.a[shell] * r_km ** 2 + self.b[shell] * r_km + self.c[shell]


def prem4_poly() -> PolyEarthModel:
    """Four shells at the PREM radii with quadratic density profiles in
    x = R / 6371 km: inner core 13.0885 - 8.8381 x^2, outer core
    12.5815 - 1.2638 x - 3.6426 x^2, lower mantle 7.9565 - 6.4761 x +
    5.5283 x^2, upper mantle and crust 5.3197 - 1.4836 x; Ye 0.466 in the
    core, 0.494 outside.  The 0-km row repeats the inner core's row."""
    R = R_EARTH_KM
    shells = [(-8.8381, 0.0, 13.0885), (-3.6426, -1.2638, 12.5815),
              (5.5283, -6.4761, 7.9565), (0.0, -1.4836, 5.3197)]
    rows = [shells[0]] + shells
    a = np.array([s[0] / R ** 2 for s in rows])
    b = np.array([s[1] / R for s in rows])
    c = np.array([s[2] for s in rows])
    yps = np.array([0.466, 0.466, 0.466, 0.494, 0.494])
    return PolyEarthModel(np.array([0.0, 1220.0, 3480.0, 5701.0, 6371.0]),
                          a, b, c, yps)


def seeded_earth(n_shells: int, seed: int) -> EarthModel:
    """n_shells shells: n_shells-1 outer radii sorted uniform in (0, 6371) km
    plus 6371 km; densities linear in shell index from 13.0 (innermost) to 2.6
    (outermost); Ye 0.467 for shells with outer radius <= 3480 km, else 0.494.
    """
    rng = np.random.default_rng(seed)
    inner = np.sort(rng.uniform(0.0, R_EARTH_KM, size=n_shells - 1))
    outer = np.concatenate([inner, [R_EARTH_KM]])
    idx = np.arange(n_shells)
    rho = 13.0 + (2.6 - 13.0) * idx / max(n_shells - 1, 1)
    ye = np.where(outer <= 3480.0, 0.467, 0.494)
    radii = np.concatenate([[0.0], outer])
    rhos = np.concatenate([[rho[0]], rho])
    yps = np.concatenate([[ye[0]], ye])
    return EarthModel(radii, rhos, yps)


def earth_12() -> EarthModel:
    """The 12-shell model (seed 12), read from its committed data file."""
    return EarthModel.read(os.path.join(DATA_DIR, "earth_12shell.dat"))


def earth_59() -> EarthModel:
    """The 59-shell model (seed 59), read from its committed data file."""
    return EarthModel.read(os.path.join(DATA_DIR, "earth_59shell.dat"))


def cos_grid(n: int) -> np.ndarray:
    """n values uniform in [-1, 0): -1 + i/n."""
    return -1.0 + np.arange(n, dtype=np.float64) / n


def log_grid(n: int, lo: float, hi: float) -> np.ndarray:
    """n energies log-spaced from lo to hi inclusive."""
    if n == 1:
        return np.array([lo], dtype=np.float64)
    return lo * (hi / lo) ** (np.arange(n, dtype=np.float64) / (n - 1))


def draw_param_sets(n: int, seed: int) -> list:
    """th23 ~ N(nominal, 0.05 rad), dm32^2 ~ N(nominal, 3%),
    th13 ~ N(nominal, 2%), dcp ~ U[-pi, pi]; th12 and dm21^2 nominal."""
    rng = np.random.default_rng(seed)
    th23 = rng.normal(TH23, 0.05, size=n)
    dm23 = rng.normal(DM23, 0.03 * DM23, size=n)
    th13 = rng.normal(TH13, 0.02 * TH13, size=n)
    dcp = rng.uniform(-np.pi, np.pi, size=n)
    return [ParamSet(TH12, float(th13[i]), float(th23[i]), float(dcp[i]),
                     DM12, float(dm23[i])) for i in range(n)]


@dataclass
class AtmosJob:
    """One atmospheric evaluation: a parameter set on a grid for one
    neutrino type."""

    params: ParamSet
    cosv: np.ndarray
    ev: np.ndarray
    nutype: int
    grid: str = ""


@dataclass
class BeamJob:
    params: ParamSet
    ev: np.ndarray
    nutype: int
    rho: float = 2.6
    L_km: float = 1300.0


@dataclass
class AtmosBatch:
    earth: EarthModel
    h_km: float
    jobs: list = field(default_factory=list)

    @property
    def n_evals(self) -> int:
        return int(sum(len(j.cosv) * len(j.ev) for j in self.jobs))


def atmos1024(n_sets: int = 1024, seed: int = 1024, big: int = 400,
              small: int = 200) -> AtmosBatch:
    """Half the sets on big x big grids and half on small x small, each for
    neutrinos and antineutrinos; cos uniform in [-1, 0), E over 0.5-50 GeV."""
    sets = draw_param_sets(n_sets, seed)
    grids = {
        big: (cos_grid(big), log_grid(big, 0.5, 50.0)),
        small: (cos_grid(small), log_grid(small, 0.5, 50.0)),
    }
    jobs = []
    for i, p in enumerate(sets):
        n = big if i < n_sets // 2 else small
        c, e = grids[n]
        for nt in (0, 1):
            jobs.append(AtmosJob(p, c, e, nt, grid=f"{n}x{n}"))
    return AtmosBatch(earth_59(), PROD_HEIGHT_KM, jobs)


def beam4096(n_sets: int = 4096, n_e: int = 100000, seed: int = 4096):
    sets = draw_param_sets(n_sets, seed)
    ev = log_grid(n_e, 0.1, 20.0)
    return [BeamJob(p, ev, 0) for p in sets]
