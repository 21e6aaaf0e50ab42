"""Species and gas constants of the reference (Constants.hpp:8-138,
ParticlePropertiesManager.hpp:73-217), with the overrides the workloads use.

omega (viscosity index) and alpha (VSS deflection) are `float` constants in the
reference (Constants.hpp:59-73), so they are carried as the float-rounded values.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

E_CHARGE = 1.602176565e-19      # Constants.hpp:14 (charge)
EV_J = 1.60217733000001e-19     # Constants.hpp:15 (energy)
KT_REFERENCE = 297.0 * 1.380658e-23
R_GAS = 8.314
N_AV = 6.02214076e23
MIN_VELOCITY = 10.0             # Constants.hpp:50
BBOX_EXPANSION = 1.01           # Constants.hpp:55

MASS = {"O2": 53.1e-27, "Ar": 6.6335209e-26, "Ne": 3.3509177e-26, "He": 6.6464731e-27,
        "Ti": 7.9485017e-26, "Al": 4.4803831e-26, "Sn": 1.9712258e-25, "W": 8.4590343e-26,
        "Au": 3.2707137e-25, "Cu": 1.0552061e-25, "Ni": 9.7462675e-26, "Ag": 1.7911901e-25}
RADIUS = {"O2": 2.035e-10, "Ar": 71e-12, "Ne": 160e-12, "He": 120e-12, "Ti": 140e-12,
          "Al": 125e-12, "Sn": 145e-12, "W": 135e-12, "Au": 135e-12, "Cu": 135e-12,
          "Ni": 135e-12, "Ag": 160e-12}
_f32 = lambda v: float(np.float32(v))
VTI = {"O2": _f32(0.77), "Ar": _f32(0.81), "Ne": _f32(0.66), "He": _f32(0.66)}
VSS_ALPHA = {"O2": _f32(1.40), "Ar": _f32(1.40), "Ne": _f32(1.31), "He": _f32(1.26)}
# ParticlePropertiesManager::getChargeFromType (:191-217); O2 and the noble gases are 0 there
CHARGE_UNITS = {"Ti": 2, "Al": 3, "Sn": 2, "W": 6, "Au": 3, "Cu": 1, "Ni": 2, "Ag": 1}


@dataclass
class Species:
    name: str
    mass: float
    charge: float
    diameter: float  # collision diameter r_p + r_t [m]


@dataclass
class Gas:
    name: str
    mass: float
    omega: float
    alpha: float
    concentration: float
    model: str  # "NONE" | "HS" | "VHS" | "VSS"


def species(name: str, gas: str, charge_units: float | None = None, diameter: float | None = None,
            mass: float | None = None) -> Species:
    q = CHARGE_UNITS.get(name, 0) if charge_units is None else charge_units
    d = RADIUS[name] + RADIUS[gas] if diameter is None else diameter
    return Species(name, MASS[name] if mass is None else mass, q * E_CHARGE, d)


def concentration(pressure: float, temperature: float) -> float:
    """util::calculateConcentration (Utilities.cpp:132-149)."""
    return (pressure / (R_GAS * temperature)) * N_AV


def gas(name: str, pressure: float, temperature: float, model: str) -> Gas:
    return Gas(name, MASS[name], VTI.get(name, 0.0), VSS_ALPHA.get(name, 0.0),
               concentration(pressure, temperature), model.upper())
