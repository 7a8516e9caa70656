"""Seeded synthetic atom configurations.

FCC lattices (what setup_system's ``lattice: {structure: FCC}`` produces,
src/setup_system/bulk_lattice.cpp:17-31), Gaussian position jitter (as
``gaussian_noise_r``), Maxwell-Boltzmann velocities with the centre-of-mass
momentum removed and an exact rescale to the target temperature
(init_temperature_new.cpp:153-177,335-430, 3N degrees of freedom).  Every draw
comes from numpy's counter-based Philox generator with an explicit seed, so the
same configuration is produced on any machine and thread count.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from .units import KB

FCC_BASIS = np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.0], [0.5, 0.0, 0.5], [0.0, 0.5, 0.5]])


@dataclasses.dataclass
class Config:
    """Atoms of one or more periodic boxes (replicas)."""
    r: np.ndarray          # (n, 3) positions, A
    v: np.ndarray          # (n, 3) velocities, A/ps
    type: np.ndarray       # (n,) uint8
    lo: np.ndarray         # (nrep, 3)
    hi: np.ndarray         # (nrep, 3)
    periodic: tuple = (1, 1, 1)
    rep: np.ndarray | None = None  # (n,) uint16 when nrep > 1

    @property
    def n(self) -> int:
        return self.r.shape[0]

    @property
    def nrep(self) -> int:
        return self.lo.shape[0]


def rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.Philox(seed))


def fcc(nx: int, ny: int, nz: int, a: float) -> np.ndarray:
    """FCC sites of nx*ny*nz cubic cells; order (iz, iy, ix, basis)."""
    iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    cells = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=1).astype(np.float64)
    r = (cells[:, None, :] + FCC_BASIS[None, :, :]).reshape(-1, 3) * a
    return r


def maxwell_boltzmann(g: np.random.Generator, masses: np.ndarray, T: float) -> np.ndarray:
    """Velocities at exactly T with zero net momentum (init_temperature_new semantics)."""
    n = masses.shape[0]
    if T <= 0 or n == 0:
        return np.zeros((n, 3))
    v = g.normal(size=(n, 3)) * np.sqrt(KB * T / masses)[:, None]
    M = masses.sum()
    v -= (masses[:, None] * v).sum(axis=0) / M
    ke = 0.5 * float((masses[:, None] * v * v).sum())
    Tcur = 2.0 * ke / (3.0 * n * KB)
    return v * np.sqrt(T / Tcur)


def lattice_config(cells, a, species_masses, T=0.0, jitter=0.0, seed=0, types=None) -> Config:
    """A periodic FCC crystal with optional jitter and MB velocities."""
    nx, ny, nz = cells
    r = fcc(nx, ny, nz, a)
    g = rng(seed)
    if jitter > 0:
        r = r + g.normal(scale=jitter, size=r.shape)
    n = r.shape[0]
    t = np.zeros(n, dtype=np.uint8) if types is None else np.asarray(types, dtype=np.uint8)
    m = np.asarray(species_masses, dtype=np.float64)[t]
    v = maxwell_boltzmann(g, m, T)
    lo = np.zeros((1, 3))
    hi = np.array([[nx * a, ny * a, nz * a]])
    return Config(r, v, t, lo, hi)


def alloy_slab_config(cells, a, species_masses, vacancy_frac=0.01, vacuum=10.0, T=0.0,
                      seed=0) -> Config:
    """Binary FCC alloy: 50/50 random substitution, random vacancies, and a free
    surface along z (vacuum on both sides; periodic in x and y only)."""
    nx, ny, nz = cells
    r = fcc(nx, ny, nz, a)
    g = rng(seed)
    n0 = r.shape[0]
    types = (g.random(n0) < 0.5).astype(np.uint8)
    keep = g.random(n0) >= vacancy_frac
    r, types = r[keep], types[keep]
    r[:, 2] += vacuum
    m = np.asarray(species_masses, dtype=np.float64)[types]
    v = maxwell_boltzmann(g, m, T)
    lo = np.zeros((1, 3))
    hi = np.array([[nx * a, ny * a, nz * a + 2 * vacuum]])
    return Config(r, v, types, lo, hi, periodic=(1, 1, 0))


def strained_batch(nconf, cells, a, species_masses, T_range=(300.0, 1500.0), strain=0.03,
                   jitter=0.05, seed=0) -> Config:
    """nconf independent periodic FCC boxes, each with its own isotropic strain
    (uniform in +-strain) and temperature T (uniform in T_range), as one replica set.
    Each box gets MB velocities at T and thermal-like Gaussian displacements of
    width jitter * sqrt(T / 300 K)."""
    g = rng(seed)
    rs, vs, ts, los, his, reps = [], [], [], [], [], []
    nx, ny, nz = cells
    for k in range(nconf):
        s = 1.0 + g.uniform(-strain, strain)
        T = g.uniform(*T_range)
        ak = a * s
        r = fcc(nx, ny, nz, ak)
        if jitter > 0:
            r = r + g.normal(scale=jitter * np.sqrt(T / 300.0), size=r.shape)
        n = r.shape[0]
        t = np.zeros(n, dtype=np.uint8)
        v = maxwell_boltzmann(g, np.asarray(species_masses)[t], T)
        rs.append(r)
        vs.append(v)
        ts.append(t)
        los.append([0.0, 0.0, 0.0])
        his.append([nx * ak, ny * ak, nz * ak])
        reps.append(np.full(n, k, dtype=np.uint16))
    return Config(np.concatenate(rs), np.concatenate(vs), np.concatenate(ts), np.array(los),
                  np.array(his), (1, 1, 1), np.concatenate(reps))


def split_replicas(c: Config) -> list[Config]:
    """One Config per replica (for per-box references)."""
    if c.rep is None:
        return [c]
    out = []
    for k in range(c.nrep):
        sel = c.rep == k
        out.append(Config(c.r[sel], c.v[sel], c.type[sel], c.lo[k:k + 1], c.hi[k:k + 1],
                          c.periodic))
    return out
