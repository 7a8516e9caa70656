"""exaStamp's internal unit system and a small quantity parser.

Internal units are Angstrom, Dalton, picosecond, elementary charge, kelvin
(include/exaStamp/unit_system.h:26-36), so energy is Da.A^2/ps^2.  The elementary
charge and the Boltzmann constant are exact SI values; the atomic mass constant is
the CODATA-2022 value, the one onika uses: the density column of the run the
reference recorded for its AlCu B2 deck
(data/regression_new/potentials/eam/eam_alloy/thermodynamic_state.csv, 724 224 Da in
72^3 A^3 at 3.221992473234 g/cm^3) gives 1.6605390689e-27 kg to its 12 digits.

Quantities are written in decks as "<value> <unit>" strings, e.g. "0.0104 eV",
"3.4 ang", "39.948 Da", "1.0e-3 ps" (lennard_jones.h:66-67 reads them through
onika's Quantity); ``parse_quantity`` converts such a string to internal units.
"""
from __future__ import annotations

import re

AMU_KG = 1.66053906892e-27
E_CHARGE_C = 1.602176634e-19
KB_J_PER_K = 1.380649e-23

ENERGY_UNIT_J = AMU_KG * 1.0e4          # 1 Da.A^2/ps^2 in joule
EV = E_CHARGE_C / ENERGY_UNIT_J         # 1 eV in Da.A^2/ps^2 (~9648.53)
KB = KB_J_PER_K / ENERGY_UNIT_J         # Boltzmann constant in Da.A^2/ps^2/K
EV_PER_ANG_TO_INTERNAL_FORCE = EV       # 1 eV/A in Da.A/ps^2
PRESSURE_UNIT_PA = ENERGY_UNIT_J / 1e-30  # internal energy / A^3 in Pa

# unit name -> factor to internal units (length A, mass Da, time ps, energy Da.A^2/ps^2)
_UNITS = {
    "ang": 1.0, "angstrom": 1.0, "A": 1.0, "nm": 10.0, "m": 1e10, "um": 1e4,
    "Da": 1.0, "u": 1.0, "amu": 1.0, "Dalton": 1.0, "g/mol": 1.0, "kg": 1.0 / AMU_KG,
    "ps": 1.0, "fs": 1e-3, "ns": 1e3, "s": 1e12,
    "eV": EV, "meV": 1e-3 * EV, "J": 1.0 / ENERGY_UNIT_J,
    "K": 1.0,
    "e-": 1.0,
    "m/s": 1e10 / 1e12,
    "eV/ang": EV, "eV*ang": EV, "eV.ang": EV,
}

_QRE = re.compile(r"^\s*([-+0-9.eE]+)\s*([A-Za-z/*.\-^0-9]*)\s*$")


def parse_quantity(q) -> float:
    """Value of a deck quantity in internal units; bare numbers pass through."""
    if isinstance(q, (int, float)):
        return float(q)
    m = _QRE.match(str(q))
    if not m:
        raise ValueError(f"cannot parse quantity {q!r}")
    val, unit = float(m.group(1)), m.group(2)
    if unit == "":
        return val
    if unit not in _UNITS:
        raise ValueError(f"unknown unit {unit!r} in {q!r}")
    return val * _UNITS[unit]


def temperature_from_ke(ke_internal: float, n_atoms: int) -> float:
    """T = 2 KE / (3 N kB) in kelvin, KE in internal units (init_temperature_new.cpp:338-345)."""
    return 2.0 * ke_internal / (3.0 * n_atoms * KB)
