"""Physics parameters and the ``gs_params`` C struct.

The element table, masses and defaults restate the reference crate:
``ELEMENTS`` (src/physics/forces.rs:43-54), ``AtomType::mass``
(src/components/atom_type.rs:80-93) and ``Config::default_config``
(src/config/mod.rs:233-260).  ``GsParams`` is the ctypes mirror of
``gs_params`` in include/genesis_gpu.h.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field
from typing import Dict, Optional

NSPECIES = 10
SYMBOLS = ("H", "He", "C", "N", "O", "Na", "Si", "P", "S", "Fe")
SPECIES_INDEX = {s: i for i, s in enumerate(SYMBOLS)}

# src/physics/forces.rs:43-54
DEFAULT_SIGMA = (1.6, 1.4, 1.9, 1.7, 1.65, 2.0, 2.3, 2.2, 2.1, 1.9)
DEFAULT_EPSILON_K = (12.0, 8.0, 80.0, 40.0, 55.0, 110.0, 200.0, 150.0, 130.0, 350.0)
DEFAULT_CHARGE = (0.0, 0.0, 0.0, 0.0, -1.0, 1.0, 0.5, 0.0, 0.0, 1.0)
# src/components/atom_type.rs:80-93
DEFAULT_MASS = (1.008, 4.0026, 12.011, 14.007, 15.999, 22.99, 28.085, 30.974, 32.06, 55.845)

LJ_CUTOFF_FACTOR = 2.5  # src/physics/forces.rs:169
LJ_NUCLEUS = 0.5  # src/physics/forces.rs:173
BOND_FACTOR = 1.5  # src/analysis/mod.rs:20
DEBOUNCE = 2  # src/systems/bonds.rs:27

GRID_REFERENCE = 0
GRID_MINIMAGE = 1

# Unit map used to express the workloads' femtoseconds in simulation time:
# length = angstrom, mass = u, energy = thermal_constant * kelvin with the
# reference's thermal_constant = 0.01, i.e. one energy unit = 100 k_B K.
_KB = 1.380649e-23
_U = 1.66053906660e-27
TIME_UNIT_FS = 1e15 * 1e-10 * math.sqrt(_U / (100.0 * _KB))  # ~109.67 fs


def fs_to_time(fs: float) -> float:
    """Femtoseconds -> simulation time units (thermal_constant = 0.01)."""
    return fs / TIME_UNIT_FS


class GsParams(ctypes.Structure):
    """ctypes mirror of ``gs_params`` (include/genesis_gpu.h)."""

    _fields_ = [
        ("box", ctypes.c_double * 3),
        ("dt", ctypes.c_double),
        ("sigma", ctypes.c_double * NSPECIES),
        ("epsilon_k", ctypes.c_double * NSPECIES),
        ("charge", ctypes.c_double * NSPECIES),
        ("mass", ctypes.c_double * NSPECIES),
        ("thermal_constant", ctypes.c_double),
        ("lj_cutoff_factor", ctypes.c_double),
        ("coulomb_constant", ctypes.c_double),
        ("coulomb_cutoff", ctypes.c_double),
        ("gravity_constant", ctypes.c_double),
        ("gravity_cutoff", ctypes.c_double),
        ("thermostat_temperature", ctypes.c_double),
        ("thermostat_tau", ctypes.c_double),
        ("bond_k_bind", ctypes.c_double),
        ("bond_min_periods", ctypes.c_double),
        ("enable_boundaries", ctypes.c_int32),
        ("enable_electrostatics", ctypes.c_int32),
        ("enable_gravity", ctypes.c_int32),
        ("enable_thermostat", ctypes.c_int32),
        ("enable_bond_interaction", ctypes.c_int32),
        ("enable_bond_observation", ctypes.c_int32),
        ("grid_semantics", ctypes.c_int32),
        ("reserved0", ctypes.c_int32),
    ]


@dataclass
class Physics:
    """The subset of the reference Config the hot path reads.

    Defaults are ``Config::default_config`` (src/config/mod.rs:233-260) with
    forces and boundaries on, as ``config/universe.toml`` ships.
    """

    box: tuple = (128.0, 128.0, 128.0)
    dt: float = 1.0 / 60.0
    sigma: tuple = DEFAULT_SIGMA
    epsilon_k: tuple = DEFAULT_EPSILON_K
    charge: tuple = DEFAULT_CHARGE
    mass: tuple = DEFAULT_MASS
    thermal_constant: float = 0.01
    lj_cutoff_factor: float = LJ_CUTOFF_FACTOR
    coulomb_constant: float = 1.0
    coulomb_cutoff: float = 2.5
    gravity_constant: float = 6.674e-11
    gravity_cutoff: float = 3.0
    thermostat_temperature: float = 300.0
    thermostat_tau: float = 20.0
    bond_k_bind: float = 1.5
    bond_min_periods: float = 10.0
    enable_boundaries: bool = True
    enable_electrostatics: bool = False
    enable_gravity: bool = False
    enable_thermostat: bool = False
    enable_bond_interaction: bool = False
    enable_bond_observation: bool = False
    grid_semantics: int = GRID_REFERENCE

    def with_overrides(self, overrides: Dict[str, Dict[str, float]]) -> "Physics":
        """ElementTable::apply_overrides (src/physics/forces.rs:90-109).

        Unknown symbols raise ``ValueError`` before anything is changed, as
        the reference fails fast on an unknown affinity-table key.
        """
        for sym in overrides:
            if sym not in SPECIES_INDEX:
                raise ValueError(f"unknown element symbol in affinity table: {sym}")
        sig, eps, q, m = list(self.sigma), list(self.epsilon_k), list(self.charge), list(self.mass)
        for sym, o in overrides.items():
            i = SPECIES_INDEX[sym]
            if "sigma" in o:
                sig[i] = float(o["sigma"])
            if "epsilon_k" in o:
                eps[i] = float(o["epsilon_k"])
            if "charge" in o:
                q[i] = float(o["charge"])
            if "mass" in o:
                m[i] = float(o["mass"])
        out = Physics(**self.__dict__)
        out.sigma, out.epsilon_k, out.charge, out.mass = tuple(sig), tuple(eps), tuple(q), tuple(m)
        return out

    def to_c(self) -> GsParams:
        p = GsParams()
        p.box[:] = [float(v) for v in self.box]
        p.dt = self.dt
        p.sigma[:] = list(self.sigma)
        p.epsilon_k[:] = list(self.epsilon_k)
        p.charge[:] = list(self.charge)
        p.mass[:] = list(self.mass)
        for name in ("thermal_constant", "lj_cutoff_factor", "coulomb_constant", "coulomb_cutoff",
                     "gravity_constant", "gravity_cutoff", "thermostat_temperature", "thermostat_tau",
                     "bond_k_bind", "bond_min_periods"):
            setattr(p, name, float(getattr(self, name)))
        for name in ("enable_boundaries", "enable_electrostatics", "enable_gravity", "enable_thermostat",
                     "enable_bond_interaction", "enable_bond_observation"):
            setattr(p, name, int(bool(getattr(self, name))))
        p.grid_semantics = int(self.grid_semantics)
        return p

    # --- derived quantities the reference computes in ForceSystem::new -------
    def max_sigma(self) -> float:
        return max(self.sigma)

    def lj_rc(self) -> float:
        """LjTable::new: rc = max(cutoff_factor, 1) * max mixed sigma."""
        return max(self.lj_cutoff_factor, 1.0) * self.max_sigma()

    def force_rc(self) -> float:
        """ForceSystem::new: the largest enabled cutoff (the grid cell size)."""
        rc = self.lj_rc()
        if self.enable_electrostatics:
            rc = max(rc, max(self.coulomb_cutoff, 1.0) * self.max_sigma())
        if self.enable_gravity:
            rc = max(rc, max(self.gravity_cutoff, 1.0) * self.max_sigma())
        return rc

    def grid_dims(self, min_cell: Optional[float] = None) -> tuple:
        """SpatialGrid::new: n = max(floor(L / min_cell), 1) per axis."""
        c = self.force_rc() if min_cell is None else min_cell
        return tuple(max(int(math.floor(L / c)), 1) for L in self.box)
