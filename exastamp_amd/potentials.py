"""EAM table sets in setfl layout: synthetic Johnson-type Cu and Cu/Ni, and setfl I/O.

A setfl (.eam.alloy) file holds, per element, F(rho) at rho_k = k*drho and
rho(r) at r_k = k*dr (k = 0..n-1), then r*phi(r) (eV.A) for each element pair of
the packed lower triangle 11, 21, 22, 31, ... -- the layout read by
eam_alloy.cpp:86-201.  ``write_setfl`` writes that layout with full FP64
precision; ``read_setfl`` reads it back.

The synthetic tables follow the functional forms named for the workloads:
rho(r) = fe exp(-beta (r/re - 1)) (the density term of johnson.h:97-115 without
its high-order denominator), a Morse pair term phi(r) = D (exp(-2 alpha (r-r0)) -
2 exp(-alpha (r-r0))), and F(rho) = -E0 sqrt(rho/rho_e) + B (rho/rho_e - 1)^2
(the sqrt term of sutton_chen.h:65-70 plus a quadratic).  rho and phi are
multiplied by a quintic taper that brings value, slope and curvature to zero
between rs and rc, so energy is conserved across the cutoff.  r0 is solved so the
perfect FCC crystal at the given lattice constant is at zero pressure.
"""
from __future__ import annotations

import dataclasses
import math
from pathlib import Path

import numpy as np


@dataclasses.dataclass
class SetflTables:
    names: list[str]
    z: list[int]
    masses: list[float]
    a0: list[float]
    lattice: list[str]
    nrho: int
    drho: float
    nr: int
    dr: float
    rc: float
    frho: np.ndarray   # (nel, nrho) eV
    rhor: np.ndarray   # (nel, nr)
    z2r: np.ndarray    # (nel(nel+1)/2, nr) eV.A

    @property
    def nel(self) -> int:
        return len(self.names)

    def theta(self) -> np.ndarray:
        """All knot values as one parameter vector (F tables, rho tables, z2r tables)."""
        return np.concatenate([self.frho.ravel(), self.rhor.ravel(), self.z2r.ravel()])

    def with_theta(self, theta: np.ndarray) -> "SetflTables":
        nf, nrr = self.frho.size, self.rhor.size
        t = np.asarray(theta, dtype=np.float64)
        return dataclasses.replace(
            self,
            frho=t[:nf].reshape(self.frho.shape).copy(),
            rhor=t[nf:nf + nrr].reshape(self.rhor.shape).copy(),
            z2r=t[nf + nrr:].reshape(self.z2r.shape).copy(),
        )


def write_setfl(path, t: SetflTables, comments=("synthetic", "", "")) -> None:
    lines = [str(c) for c in comments[:3]]
    lines += [""] * (3 - len(lines))
    lines.append(f"{t.nel} " + " ".join(t.names))
    lines.append(f"{t.nrho} {t.drho:.17e} {t.nr} {t.dr:.17e} {t.rc:.17e}")

    def block(vals):
        v = np.asarray(vals, dtype=np.float64)
        return [" ".join(f"{x:.17e}" for x in v[i:i + 5]) for i in range(0, v.size, 5)]

    for e in range(t.nel):
        lines.append(f"{t.z[e]} {t.masses[e]:.17g} {t.a0[e]:.17g} {t.lattice[e]}")
        lines += block(t.frho[e])
        lines += block(t.rhor[e])
    for k in range(t.z2r.shape[0]):
        lines += block(t.z2r[k])
    Path(path).write_text("\n".join(lines) + "\n")


def read_setfl(path) -> SetflTables:
    """Token-order reader of the setfl layout (eam_alloy.cpp:86-201)."""
    text = Path(path).read_text().splitlines()
    toks = " ".join(text[3:]).split()
    pos = 0

    def take():
        nonlocal pos
        pos += 1
        return toks[pos - 1]

    nel = int(take())
    names = [take() for _ in range(nel)]
    nrho, drho, nr, dr, rc = int(take()), float(take()), int(take()), float(take()), float(take())
    z, masses, a0, lat = [], [], [], []
    frho = np.zeros((nel, nrho))
    rhor = np.zeros((nel, nr))
    for e in range(nel):
        z.append(int(take()))
        masses.append(float(take()))
        a0.append(float(take()))
        lat.append(take())
        frho[e] = [float(take()) for _ in range(nrho)]
        rhor[e] = [float(take()) for _ in range(nr)]
    nz = nel * (nel + 1) // 2
    z2r = np.array([[float(take()) for _ in range(nr)] for _ in range(nz)])
    return SetflTables(names, z, masses, a0, lat, nrho, drho, nr, dr, rc, frho, rhor, z2r)


# ---------------------------------------------------------------------------
# synthetic Johnson-type tables
# ---------------------------------------------------------------------------

def taper(r, rs, rc):
    """1 below rs, 0 above rc, quintic smoothstep (C2) in between."""
    r = np.asarray(r, dtype=np.float64)
    x = np.clip((r - rs) / (rc - rs), 0.0, 1.0)
    return 1.0 - x ** 3 * (10.0 - 15.0 * x + 6.0 * x * x)


@dataclasses.dataclass
class JohnsonElement:
    name: str
    z: int
    mass: float
    a: float            # FCC lattice constant (A)
    beta: float = 6.0   # density decay
    fe: float = 1.0
    E0: float = 2.6     # eV, sqrt embedding depth
    B: float = 0.6      # eV, quadratic embedding curvature
    D: float = 0.18     # eV, Morse depth
    alpha: float = 1.55  # 1/A, Morse range
    r0: float = 0.0     # solved for zero pressure when 0

    @property
    def re(self) -> float:
        return self.a / math.sqrt(2.0)


def _fcc_shells(a, rmax):
    """Distances and multiplicities of FCC neighbour shells up to rmax."""
    n = int(math.ceil(rmax / a)) + 1
    g = np.arange(-2 * n, 2 * n + 1)
    i, j, k = np.meshgrid(g, g, g, indexing="ij")
    sel = ((i + j + k) % 2 == 0) & ~((i == 0) & (j == 0) & (k == 0))
    s = (i[sel] ** 2 + j[sel] ** 2 + k[sel] ** 2).astype(np.int64)  # exact shell index
    u, c = np.unique(s, return_counts=True)
    d = 0.5 * a * np.sqrt(u.astype(np.float64))
    keep = d < rmax
    return d[keep], c[keep]


def _rho_fn(el, rs, rc):
    return lambda r: el.fe * np.exp(-el.beta * (np.asarray(r) / el.re - 1.0)) * taper(r, rs, rc)


def _phi_fn(D, alpha, r0, rs, rc):
    def phi(r):
        r = np.asarray(r, dtype=np.float64)
        x = np.exp(-alpha * (r - r0))
        return D * (x * x - 2.0 * x) * taper(r, rs, rc)
    return phi


def _solve_r0(el, rs, rc):
    """Morse r0 giving zero pressure for the perfect FCC crystal at el.a."""
    rho = _rho_fn(el, rs, rc)
    d, c = _fcc_shells(el.a, rc)

    def energy(a_, r0):
        s = a_ / el.a
        rho_e_ref = float(np.sum(c * rho(d)))
        rh = float(np.sum(c * rho(d * s)))
        F = -el.E0 * math.sqrt(rh / rho_e_ref) + el.B * (rh / rho_e_ref - 1.0) ** 2
        return F + 0.5 * float(np.sum(c * _phi_fn(el.D, el.alpha, r0, rs, rc)(d * s)))

    def dEda(r0, h=1e-5):
        return (energy(el.a + h, r0) - energy(el.a - h, r0)) / (2 * h)

    lo, hi = 0.6 * el.re, 1.6 * el.re
    flo, fhi = dEda(lo), dEda(hi)
    if flo * fhi > 0:
        raise RuntimeError("cannot bracket Morse r0 for zero pressure")
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        fm = dEda(mid)
        if (fm > 0) == (flo > 0):
            lo, flo = mid, fm
        else:
            hi = mid
    return 0.5 * (lo + hi)


CU = JohnsonElement("Cu", 29, 63.546, 3.615)
NI = JohnsonElement("Ni", 28, 58.693, 3.52, beta=6.3, E0=2.9, B=0.7, D=0.2, alpha=1.6)


def johnson_tables(elements=(CU,), nr=5000, nrho=5000, rc=4.95, taper_width=1.0,
                   rhomax_factor=2.0) -> SetflTables:
    """Synthetic setfl tables for one or more Johnson-type elements."""
    rs = rc - taper_width
    els = [dataclasses.replace(e) for e in elements]
    for e in els:
        if e.r0 == 0.0:
            e.r0 = _solve_r0(e, rs, rc)
    dr = rc / (nr - 1)
    r = np.arange(nr) * dr
    rho_e = []
    for e in els:
        d, c = _fcc_shells(e.a, rc)
        rho_e.append(float(np.sum(c * _rho_fn(e, rs, rc)(d))))
    rhomax = rhomax_factor * max(rho_e)
    drho = rhomax / (nrho - 1)
    rho_grid = np.arange(nrho) * drho
    frho = np.zeros((len(els), nrho))
    rhor = np.zeros((len(els), nr))
    for k, e in enumerate(els):
        x = rho_grid / rho_e[k]
        frho[k] = -e.E0 * np.sqrt(x) + e.B * (x - 1.0) ** 2
        rhor[k] = _rho_fn(e, rs, rc)(r)
    nz = len(els) * (len(els) + 1) // 2
    z2r = np.zeros((nz, nr))
    idx = 0
    for i in range(len(els)):
        for j in range(i + 1):
            ei, ej = els[i], els[j]
            phi_i = _phi_fn(ei.D, ei.alpha, ei.r0, rs, rc)(r)
            phi_j = _phi_fn(ej.D, ej.alpha, ej.r0, rs, rc)(r)
            z2r[idx] = r * 0.5 * (phi_i + phi_j)
            idx += 1
    return SetflTables([e.name for e in els], [e.z for e in els], [e.mass for e in els],
                       [e.a for e in els], ["fcc"] * len(els), nrho, drho, nr, dr, rc,
                       frho, rhor, z2r)
