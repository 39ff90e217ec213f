"""Mirror of the reference's plugin interface for the hot path.

The reference composes a tick from ``System`` objects
(``trait System { name, access, run }``, src/scheduler/system.rs:86-97)
registered in ``build_schedule`` order (src/universe/mod.rs:355-392) and run
sequentially by ``Scheduler::run`` (src/scheduler/mod.rs:83-89).  The classes
below keep those names and that order; each ``run`` is one C-ABI call on the
resident GPU state, so a caller that drives systems one by one sees the same
call sequence as the reference while the fused ``gs_step`` runs the whole
schedule in one resident loop.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List

from .params import Physics
from .universe import GpuUniverse


@dataclass
class SystemContext:
    """SystemContext (src/scheduler/system.rs:75-83): the resident universe
    stands in for ``world``; ``resources`` holds PotentialEnergy/BondEnergy."""

    universe: GpuUniverse
    resources: Dict[str, float] = field(default_factory=dict)
    tick: int = 0


class System:
    def name(self) -> str:  # pragma: no cover - interface
        raise NotImplementedError

    def run(self, ctx: SystemContext) -> None:  # pragma: no cover - interface
        raise NotImplementedError


class VelocityHalfKick(System):
    """src/systems/integrate.rs:25-45."""

    def name(self) -> str:
        return "velocity_half_kick"

    def run(self, ctx: SystemContext) -> None:
        ctx.universe.half_kick()


class PositionDrift(System):
    """src/systems/integrate.rs:48-68."""

    def name(self) -> str:
        return "position_drift"

    def run(self, ctx: SystemContext) -> None:
        ctx.universe.drift()


class BoundarySystem(System):
    """src/systems/boundary.rs:11-35."""

    def name(self) -> str:
        return "boundary"

    def run(self, ctx: SystemContext) -> None:
        ctx.universe.wrap()


class ForceSystem(System):
    """src/systems/forces.rs:71-354: writes Acceleration, PotentialEnergy and
    BondEnergy."""

    def name(self) -> str:
        return "forces"

    def run(self, ctx: SystemContext) -> None:
        v, vb = ctx.universe.forces()
        ctx.resources["PotentialEnergy"] = v
        ctx.resources["BondEnergy"] = vb


class ThermostatSystem(System):
    """src/systems/thermostat.rs:38-71."""

    def name(self) -> str:
        return "thermostat"

    def run(self, ctx: SystemContext) -> None:
        ctx.resources["lambda"] = ctx.universe.thermostat()


class BondObservationSystem(System):
    """src/systems/bonds.rs:73-185: bound pairs, the pair tracker and the
    Bonds adjacency of this tick (read by the next tick's ForceSystem)."""

    def name(self) -> str:
        return "bond_observation"

    def run(self, ctx: SystemContext) -> None:
        ctx.universe.bond_observe(ctx.tick)


class Scheduler:
    """Sequential scheduler in registration order (src/scheduler/mod.rs:34-89)."""

    def __init__(self) -> None:
        self.systems: List[System] = []

    def add_system(self, s: System) -> None:
        self.systems.append(s)

    def names(self) -> List[str]:
        return [s.name() for s in self.systems]

    def run(self, ctx: SystemContext) -> None:
        for s in self.systems:
            s.run(ctx)


def build_schedule(physics: Physics) -> Scheduler:
    """build_schedule with forces on (src/universe/mod.rs:355-388), up to the
    hot path's last system: the observation lenses after it (structure,
    bond structure, stats) read the state and are separate calls."""
    s = Scheduler()
    s.add_system(VelocityHalfKick())
    s.add_system(PositionDrift())
    if physics.enable_boundaries:
        s.add_system(BoundarySystem())
    s.add_system(ForceSystem())
    s.add_system(VelocityHalfKick())
    if physics.enable_thermostat:
        s.add_system(ThermostatSystem())
    if physics.enable_bond_observation:
        s.add_system(BondObservationSystem())
    return s


def tick(scheduler: Scheduler, ctx: SystemContext) -> None:
    """Universe::tick (src/universe/mod.rs:140-159): the clock advances, then
    the schedule runs once."""
    ctx.tick += 1
    scheduler.run(ctx)
