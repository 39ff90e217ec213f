"""Stats snapshot of a tick (StatsSystem, src/systems/aggregate.rs:36-143;
StatsSnapshot and StatsCollector, src/stats/mod.rs:28-190) fed from the
resident GPU state.

The per-entity sums the reference takes over Velocity and Mass (kinetic
energy and the speed sum, :87-93) are device reductions
(``gs_speed_stats``); the bond observation summary comes from
``gs_bond_stats_read``; the potential and bond energies are the tick's
observables; structure (friends-of-friends) and chemical (bond-graph)
summaries come from their lenses.  What remains here is the arithmetic of
:95-140 on those scalars.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional


@dataclass
class StatsSnapshot:
    """StatsSnapshot (src/stats/mod.rs:28-77); ``structure`` and ``chemical``
    are dicts with the fields of StructureStats and ChemicalStructure."""
    tick: int = 0
    time: float = 0.0
    entities: int = 0
    energy_total: float = 0.0
    energy_avg: float = 0.0
    energy_potential: float = 0.0
    temperature_avg: float = 0.0
    mean_speed: float = 0.0
    density: float = 0.0
    collisions: int = 0
    systems_run: int = 0
    fps: float = 0.0
    memory_bytes: int = 0
    structure: Optional[dict] = None
    bonded_pairs: int = 0
    bonds_formed: int = 0
    bond_lifetime_ticks: float = 0.0
    bond_energy: float = 0.0
    bond_matrix: List[float] = field(default_factory=list)
    chemical: Optional[dict] = None


def snapshot_from_sums(tick: int, time: float, entities: int, count: int, kinetic: float, speed_sum: float,
                       thermal_constant: float, volume: Optional[float], energy_potential: float = 0.0,
                       bond_energy: float = 0.0, collisions: int = 0, systems_run: int = 0, fps: float = 0.0,
                       memory_bytes: int = 0, structure: Optional[dict] = None,
                       bond_observation: Optional[dict] = None, chemical: Optional[dict] = None) -> StatsSnapshot:
    """StatsSystem::run (src/systems/aggregate.rs:36-143) from its reductions:
    kinetic = sum 0.5 m |v|^2 and speed_sum = sum |v| over the ``count``
    entities that have a Velocity and a Mass; ``entities`` = world.len()."""
    structure = structure if structure is not None and structure.get("tick", 0) > 0 else None
    chemical = chemical if chemical is not None and chemical.get("tick", 0) > 0 else None
    if bond_observation is not None:
        bonded_pairs = int(bond_observation["bonded_pairs"])
        bonds_formed = int(bond_observation["bonds_formed"])
        bond_lifetime = (float(bond_observation["lifetime_sum_ticks"]) / bonds_formed) if bonds_formed > 0 else 0.0
        bond_matrix = [float(c) for c in list(bond_observation["species_matrix"].reshape(-1))]
    else:
        bonded_pairs, bonds_formed, bond_lifetime, bond_matrix = 0, 0, 0.0, []
    energy_avg = kinetic / count if count > 0 else 0.0
    temperature = (2.0 / 3.0) * energy_avg / thermal_constant if thermal_constant > 0.0 else 0.0
    density = entities / volume if volume is not None and volume > 0.0 else 0.0
    return StatsSnapshot(tick=int(tick), time=float(time), entities=int(entities),
                         energy_total=kinetic + energy_potential + bond_energy, energy_avg=energy_avg,
                         energy_potential=energy_potential, temperature_avg=temperature,
                         mean_speed=speed_sum / count if count > 0 else 0.0, density=density,
                         collisions=int(collisions), systems_run=int(systems_run), fps=float(fps),
                         memory_bytes=int(memory_bytes), structure=structure, bonded_pairs=bonded_pairs,
                         bonds_formed=bonds_formed, bond_lifetime_ticks=bond_lifetime, bond_energy=bond_energy,
                         bond_matrix=bond_matrix, chemical=chemical)


class StatsSystem:
    """StatsSystem (src/systems/aggregate.rs:16-143) on a GPU universe."""

    def name(self) -> str:
        return "stats"

    def run(self, universe, tick: int, time: float, energy_potential: float = 0.0, bond_energy: float = 0.0,
            collisions: int = 0, systems_run: int = 0, fps: float = 0.0, memory_bytes: int = 0,
            structure: Optional[dict] = None, chemical: Optional[dict] = None) -> StatsSnapshot:
        ph = universe.physics
        kinetic, speed_sum = universe.speed_stats()
        box = ph.box
        volume = box[0] * box[1] * box[2]
        bond_obs = universe.bond_stats() if ph.enable_bond_observation else None
        return snapshot_from_sums(tick, time, universe.n, universe.n, kinetic, speed_sum, ph.thermal_constant,
                                  volume, energy_potential, bond_energy, collisions, systems_run, fps,
                                  memory_bytes, structure, bond_obs, chemical)


class StatsCollector:
    """StatsCollector (src/stats/mod.rs:147-190): the latest snapshot and a
    bounded history (the oldest entry leaves when the cap is reached)."""

    def __init__(self, history_cap: int):
        self.snapshot = StatsSnapshot()
        self.systems_run = 0
        self.fps = 0.0
        self._history: List[StatsSnapshot] = []
        self.history_cap = int(history_cap)

    def record(self, snapshot: StatsSnapshot) -> None:
        if len(self._history) == self.history_cap:
            self._history.pop(0)
        self.snapshot = snapshot
        self._history.append(snapshot)

    def history(self) -> List[StatsSnapshot]:
        return self._history

