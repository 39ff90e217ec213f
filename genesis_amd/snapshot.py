"""Snapshots of a GPU universe in the reference's save format.

``save_universe`` / ``load_universe`` restate src/serialization/mod.rs:84-120:
the file is ``bincode::serialize(&Snapshot)`` with FORMAT_VERSION 4 (:23),

    Snapshot { format: u32, config: Config, state: UniverseState,
               world: Vec<EntitySnapshot> }                       (:33-39)
    UniverseState { time: Time, rng: Rng, stats: StatsCollector } (:26-31)
    EntitySnapshot { index: u32, generation: u32, position, velocity, mass,
                     charge, atom_type, bonds, acceleration }     (:42-53)

in bincode 1's default encoding: little-endian fixed-width integers, f64 as
its 8 bytes, bool as one byte, usize as u64, strings and sequences with a
u64 length prefix, maps as a u64 count then key/value pairs, Option as a
one-byte tag (0 None, 1 Some) then the value, unit enum variants as their
u32 variant index (AtomType: H = 0 ... Fe = 9, the element index used here),
newtype structs (Position(Vec3), Mass(f64), ...) as their content, fixed
arrays ([u64; 4], the xoshiro state) as their elements.  serde's derive
writes a unit variant's declaration index, not its explicit discriminant
(Hydrogen = 1, ... in src/components/atom_type.rs:14-26).

After the Snapshot the file carries the device image written by
``gs_snapshot_append`` (resident order, Verlet lists, bond tracker, tick):
bincode's ``deserialize`` ignores trailing bytes, so the reference still
reads the file, and a universe loaded here with the image continues bit for
bit.  A file without the image (written by the reference) loads from its
components alone.

The reference's RNG is xoshiro256++ seeded through SplitMix64 (rand_xoshiro
0.6, seed_from_u64); its state is carried through a load/save unchanged, and
a fresh config's state is derived from the seed with SplitMix64 as that
crate does: parity unpinned (no reference vectors; this path never draws
from it).
"""
from __future__ import annotations

import ctypes
import os
import struct
from dataclasses import dataclass, field, fields
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import _lib
from .inputs import State
from .params import (DEFAULT_CHARGE, DEFAULT_EPSILON_K, DEFAULT_SIGMA, GRID_REFERENCE, LJ_CUTOFF_FACTOR, NSPECIES,
                     SYMBOLS, Physics)
from .stats import StatsCollector, StatsSnapshot

FORMAT_VERSION = 4  # src/serialization/mod.rs:23
DEVICE_MAGIC = b"GSDEVIMG"


# ---- bincode 1 (default options) ------------------------------------------------
class Writer:
    def __init__(self):
        self.parts: List[bytes] = []

    def raw(self, b: bytes) -> None:
        self.parts.append(b)

    def u8(self, v: int) -> None:
        self.raw(struct.pack("<B", v))

    def u32(self, v: int) -> None:
        self.raw(struct.pack("<I", v))

    def u64(self, v: int) -> None:
        self.raw(struct.pack("<Q", v))

    def f64(self, v: float) -> None:
        self.raw(struct.pack("<d", v))

    def boolean(self, v: bool) -> None:
        self.u8(1 if v else 0)

    def string(self, s: str) -> None:
        b = s.encode("utf-8")
        self.u64(len(b))
        self.raw(b)

    def option(self, v, put) -> None:
        if v is None:
            self.u8(0)
        else:
            self.u8(1)
            put(v)

    def seq(self, items, put) -> None:
        self.u64(len(items))
        for it in items:
            put(it)

    def bytes(self) -> bytes:
        return b"".join(self.parts)


class Reader:
    def __init__(self, data: bytes):
        self.data = data
        self.pos = 0

    def take(self, n: int) -> bytes:
        if self.pos + n > len(self.data):
            raise ValueError("snapshot truncated")
        b = self.data[self.pos:self.pos + n]
        self.pos += n
        return b

    def u8(self) -> int:
        return struct.unpack("<B", self.take(1))[0]

    def u32(self) -> int:
        return struct.unpack("<I", self.take(4))[0]

    def u64(self) -> int:
        return struct.unpack("<Q", self.take(8))[0]

    def f64(self) -> float:
        return struct.unpack("<d", self.take(8))[0]

    def boolean(self) -> bool:
        v = self.u8()
        if v > 1:
            raise ValueError("invalid bool")
        return v == 1

    def string(self) -> str:
        return self.take(self.u64()).decode("utf-8")

    def option(self, get):
        tag = self.u8()
        if tag == 0:
            return None
        if tag != 1:
            raise ValueError("invalid Option tag")
        return get()

    def seq(self, get) -> list:
        return [get() for _ in range(self.u64())]


# ---- Config (src/config/mod.rs:17-260) --------------------------------------------
@dataclass
class Config:
    """Config with Config::default_config's values (src/config/mod.rs:233-260)."""
    # universe
    name: str = "Genesis"
    size: Tuple[float, float, float] = (128.0, 128.0, 128.0)
    dt: float = 1.0 / 60.0
    initial_atoms: int = 10_000
    elements: List[int] = field(default_factory=lambda: list(range(NSPECIES)))
    stats_history: int = 1024
    # rng
    seed: int = 42
    # physics
    particle_radius: float = 0.4
    initial_temperature: float = 300.0
    speed_limit: float = 1_000.0
    thermal_constant: float = 0.01
    gravity_constant: float = 6.674e-11
    coulomb_constant: float = 1.0
    coulomb_cutoff: float = 2.5
    gravity_cutoff: float = 3.0
    thermostat_temperature: float = 300.0
    thermostat_tau: float = 20.0
    element_overrides: Dict[str, Dict[str, Optional[float]]] = field(default_factory=dict)
    # systems (SystemsConfig::default)
    enable_movement: bool = True
    enable_boundaries: bool = True
    enable_collisions: bool = False
    enable_forces: bool = True
    enable_electrostatics: bool = False
    enable_gravity: bool = False
    enable_thermostat: bool = False
    enable_bond_observation: bool = False
    enable_bond_interaction: bool = False
    # stats (StatsConfig::default)
    histogram_bins: int = 32
    histogram_max_speed: float = 10.0
    structure_interval: int = 100
    csv_path: str = ""
    csv_interval: int = 10
    xyz_prefix: str = ""
    xyz_interval: int = 100
    bond_min_periods: float = 10.0
    bond_k_bind: float = 1.5

    @classmethod
    def from_physics(cls, ph: Physics, n: int, **kw) -> "Config":
        """The Config fields the force path reads, from a Physics; element
        sigma / epsilon / charge that differ from the built-in table become
        overrides (apply_overrides, src/physics/forces.rs:90-109)."""
        over: Dict[str, Dict[str, Optional[float]]] = {}
        for s in range(NSPECIES):
            o = {"sigma": None, "epsilon_k": None, "charge": None}
            if ph.sigma[s] != DEFAULT_SIGMA[s]:
                o["sigma"] = float(ph.sigma[s])
            if ph.epsilon_k[s] != DEFAULT_EPSILON_K[s]:
                o["epsilon_k"] = float(ph.epsilon_k[s])
            if ph.charge[s] != DEFAULT_CHARGE[s]:
                o["charge"] = float(ph.charge[s])
            if any(v is not None for v in o.values()):
                over[SYMBOLS[s]] = o
        cfg = cls(size=tuple(float(b) for b in ph.box), dt=float(ph.dt), initial_atoms=int(n),
                  thermal_constant=ph.thermal_constant, gravity_constant=ph.gravity_constant,
                  coulomb_constant=ph.coulomb_constant, coulomb_cutoff=ph.coulomb_cutoff,
                  gravity_cutoff=ph.gravity_cutoff, thermostat_temperature=ph.thermostat_temperature,
                  thermostat_tau=ph.thermostat_tau, element_overrides=over, enable_forces=True,
                  enable_boundaries=bool(ph.enable_boundaries), enable_electrostatics=bool(ph.enable_electrostatics),
                  enable_gravity=bool(ph.enable_gravity), enable_thermostat=bool(ph.enable_thermostat),
                  enable_bond_observation=bool(ph.enable_bond_observation),
                  enable_bond_interaction=bool(ph.enable_bond_interaction), bond_min_periods=ph.bond_min_periods,
                  bond_k_bind=ph.bond_k_bind)
        for k, v in kw.items():
            setattr(cfg, k, v)
        return cfg

    def to_physics(self, mass, grid_semantics: int = GRID_REFERENCE) -> Physics:
        """Physics of this config; mass: per-species masses (the Mass
        components; the reference has no per-species mass setting)."""
        ph = Physics(box=tuple(self.size), dt=self.dt, mass=tuple(float(m) for m in mass),
                     thermal_constant=self.thermal_constant, lj_cutoff_factor=LJ_CUTOFF_FACTOR,
                     coulomb_constant=self.coulomb_constant, coulomb_cutoff=self.coulomb_cutoff,
                     gravity_constant=self.gravity_constant, gravity_cutoff=self.gravity_cutoff,
                     thermostat_temperature=self.thermostat_temperature, thermostat_tau=self.thermostat_tau,
                     bond_k_bind=self.bond_k_bind, bond_min_periods=self.bond_min_periods,
                     enable_boundaries=self.enable_boundaries, enable_electrostatics=self.enable_electrostatics,
                     enable_gravity=self.enable_gravity, enable_thermostat=self.enable_thermostat,
                     enable_bond_interaction=self.enable_bond_interaction,
                     enable_bond_observation=self.enable_bond_observation, grid_semantics=grid_semantics)
        over = {sym: {k: v for k, v in o.items() if v is not None} for sym, o in self.element_overrides.items()}
        return ph.with_overrides(over)


def _put_config(w: Writer, c: Config) -> None:
    # UniverseConfig
    w.string(c.name)
    for v in c.size:
        w.f64(v)
    w.f64(c.dt)
    w.u64(c.initial_atoms)
    w.seq(c.elements, w.u32)
    w.u64(c.stats_history)
    # RngConfig
    w.u64(c.seed)
    # PhysicsConfig
    for v in (c.particle_radius, c.initial_temperature, c.speed_limit, c.thermal_constant, c.gravity_constant,
              c.coulomb_constant, c.coulomb_cutoff, c.gravity_cutoff, c.thermostat_temperature, c.thermostat_tau):
        w.f64(v)
    w.u64(len(c.element_overrides))
    for sym in sorted(c.element_overrides):  # a HashMap: any order reads back the same
        o = c.element_overrides[sym]
        w.string(sym)
        for k in ("sigma", "epsilon_k", "charge"):
            w.option(o.get(k), w.f64)
    # SystemsConfig
    for v in (c.enable_movement, c.enable_boundaries, c.enable_collisions, c.enable_forces,
              c.enable_electrostatics, c.enable_gravity, c.enable_thermostat, c.enable_bond_observation,
              c.enable_bond_interaction):
        w.boolean(v)
    # StatsConfig
    w.u64(c.histogram_bins)
    w.f64(c.histogram_max_speed)
    w.u64(c.structure_interval)
    w.string(c.csv_path)
    w.u64(c.csv_interval)
    w.string(c.xyz_prefix)
    w.u64(c.xyz_interval)
    w.f64(c.bond_min_periods)
    w.f64(c.bond_k_bind)


def _get_config(r: Reader) -> Config:
    c = Config()
    c.name = r.string()
    c.size = (r.f64(), r.f64(), r.f64())
    c.dt = r.f64()
    c.initial_atoms = r.u64()
    c.elements = r.seq(r.u32)
    c.stats_history = r.u64()
    c.seed = r.u64()
    (c.particle_radius, c.initial_temperature, c.speed_limit, c.thermal_constant, c.gravity_constant,
     c.coulomb_constant, c.coulomb_cutoff, c.gravity_cutoff, c.thermostat_temperature,
     c.thermostat_tau) = (r.f64() for _ in range(10))
    over = {}
    for _ in range(r.u64()):
        sym = r.string()
        over[sym] = {k: r.option(r.f64) for k in ("sigma", "epsilon_k", "charge")}
    c.element_overrides = over
    (c.enable_movement, c.enable_boundaries, c.enable_collisions, c.enable_forces, c.enable_electrostatics,
     c.enable_gravity, c.enable_thermostat, c.enable_bond_observation,
     c.enable_bond_interaction) = (r.boolean() for _ in range(9))
    c.histogram_bins = r.u64()
    c.histogram_max_speed = r.f64()
    c.structure_interval = r.u64()
    c.csv_path = r.string()
    c.csv_interval = r.u64()
    c.xyz_prefix = r.string()
    c.xyz_interval = r.u64()
    c.bond_min_periods = r.f64()
    c.bond_k_bind = r.f64()
    return c


# ---- UniverseState: Time, Rng, StatsCollector ------------------------------------
def splitmix_state(seed: int) -> Tuple[int, int, int, int]:
    """xoshiro256++ state of seed_from_u64 (rand_xoshiro 0.6): four SplitMix64
    outputs (Steele, Lea & Flood, OOPSLA 2014) from state = seed."""
    out, s = [], seed & 0xFFFFFFFFFFFFFFFF
    for _ in range(4):
        s = (s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        out.append(z ^ (z >> 31))
    return tuple(out)


def _put_stats_snapshot(w: Writer, s: StatsSnapshot) -> None:
    w.u64(s.tick)
    w.f64(s.time)
    w.u64(s.entities)
    for v in (s.energy_total, s.energy_avg, s.energy_potential, s.temperature_avg, s.mean_speed, s.density):
        w.f64(v)
    w.u64(s.collisions)
    w.u64(s.systems_run)
    w.f64(s.fps)
    w.u64(s.memory_bytes)

    def structure(st):
        w.u64(st["tick"])
        w.u64(st["monomers"])
        w.u64(st["aggregates"])
        w.u64(st["largest"])
        w.f64(st["mean_size"])
        w.u64(st["bound_pairs"])
    w.option(s.structure, structure)
    w.u64(s.bonded_pairs)
    w.u64(s.bonds_formed)
    w.f64(s.bond_lifetime_ticks)
    w.f64(s.bond_energy)
    w.seq(list(s.bond_matrix), w.f64)

    def chemical(ch):
        w.u64(ch["tick"])
        for k in ("aggregates", "bound_entities", "monomers", "largest"):
            w.u64(ch[k])

        def comp(c):
            w.string(c["formula"])
            w.u64(c["count"])
            w.f64(c["mean_binding"])
        w.seq(ch["compositions"], comp)
        for k in ("appeared", "disappeared", "fusions", "scissions"):
            w.u64(ch[k])

        def reaction(rx):
            w.seq(rx["reactants"], w.string)
            w.seq(rx["products"], w.string)
            w.u64(rx["count"])
        w.seq(ch["reactions"], reaction)
    w.option(s.chemical, chemical)


def _get_stats_snapshot(r: Reader) -> StatsSnapshot:
    s = StatsSnapshot()
    s.tick = r.u64()
    s.time = r.f64()
    s.entities = r.u64()
    (s.energy_total, s.energy_avg, s.energy_potential, s.temperature_avg, s.mean_speed,
     s.density) = (r.f64() for _ in range(6))
    s.collisions = r.u64()
    s.systems_run = r.u64()
    s.fps = r.f64()
    s.memory_bytes = r.u64()
    s.structure = r.option(lambda: {"tick": r.u64(), "monomers": r.u64(), "aggregates": r.u64(),
                                    "largest": r.u64(), "mean_size": r.f64(), "bound_pairs": r.u64()})
    s.bonded_pairs = r.u64()
    s.bonds_formed = r.u64()
    s.bond_lifetime_ticks = r.f64()
    s.bond_energy = r.f64()
    s.bond_matrix = r.seq(r.f64)

    def chemical():
        ch = {"tick": r.u64()}
        for k in ("aggregates", "bound_entities", "monomers", "largest"):
            ch[k] = r.u64()
        ch["compositions"] = r.seq(lambda: {"formula": r.string(), "count": r.u64(), "mean_binding": r.f64()})
        for k in ("appeared", "disappeared", "fusions", "scissions"):
            ch[k] = r.u64()
        ch["reactions"] = r.seq(lambda: {"reactants": r.seq(r.string), "products": r.seq(r.string),
                                         "count": r.u64()})
        return ch
    s.chemical = r.option(chemical)
    return s


@dataclass
class UniverseState:
    """UniverseState (src/serialization/mod.rs:26-31): Time { t, tick, dt }
    (src/universe/time.rs:10-17), the RNG state and the stats collector."""
    t: float = 0.0
    tick: int = 0
    dt: float = 1.0 / 60.0
    rng: Tuple[int, int, int, int] = (0, 0, 0, 0)
    stats: StatsCollector = field(default_factory=lambda: StatsCollector(1024))


def advance_time(t: float, tick: int, to_tick: int, dt: float) -> float:
    """Time::t after Time::advance from tick to to_tick (src/universe/time.rs:25-28):
    dt added once per tick, rounded each time."""
    for _ in range(max(0, to_tick - tick)):
        t += dt
    return t


def _put_state(w: Writer, st: UniverseState) -> None:
    w.f64(st.t)
    w.u64(st.tick)
    w.f64(st.dt)
    for v in st.rng:
        w.u64(v)
    c = st.stats
    _put_stats_snapshot(w, c.snapshot)
    w.u64(c.systems_run)
    w.f64(c.fps)
    w.seq(c.history(), lambda s: _put_stats_snapshot(w, s))
    w.u64(c.history_cap)


def _get_state(r: Reader) -> UniverseState:
    st = UniverseState()
    st.t = r.f64()
    st.tick = r.u64()
    st.dt = r.f64()
    st.rng = tuple(r.u64() for _ in range(4))
    snap = _get_stats_snapshot(r)
    systems_run = r.u64()
    fps = r.f64()
    history = r.seq(lambda: _get_stats_snapshot(r))
    cap = r.u64()
    c = StatsCollector(cap)
    c.snapshot, c.systems_run, c.fps = snap, systems_run, fps
    c._history = history
    st.stats = c
    return st


# ---- World: EntitySnapshot rows (src/serialization/mod.rs:42-53) ------------------
@dataclass
class World:
    """The entities' components as columns, in the order of the file.  A
    missing component is a NaN row (vectors, mass, charge), species 255 or
    bonds None; bonds are lists of (index, generation)."""
    index: np.ndarray
    generation: np.ndarray
    pos: np.ndarray    # (n, 3)
    vel: np.ndarray    # (n, 3)
    mass: np.ndarray
    charge: np.ndarray
    species: np.ndarray  # uint8
    bonds: List[Optional[List[Tuple[int, int]]]]
    acc: np.ndarray    # (n, 3)

    @property
    def n(self) -> int:
        return int(self.index.shape[0])


def _put_world(w: Writer, wd: World) -> None:
    w.u64(wd.n)
    for k in range(wd.n):
        w.u32(int(wd.index[k]))
        w.u32(int(wd.generation[k]))
        for vec in (wd.pos, wd.vel):
            if np.isnan(vec[k]).any():
                w.u8(0)
            else:
                w.u8(1)
                w.raw(struct.pack("<3d", *vec[k]))
        for sc in (wd.mass, wd.charge):
            w.option(None if np.isnan(sc[k]) else float(sc[k]), w.f64)
        w.option(None if wd.species[k] == 255 else int(wd.species[k]), w.u32)
        b = wd.bonds[k]
        if b is None:
            w.u8(0)
        else:
            w.u8(1)
            w.u64(len(b))
            w.raw(b"".join(struct.pack("<II", i, g) for i, g in b))
        if np.isnan(wd.acc[k]).any():
            w.u8(0)
        else:
            w.u8(1)
            w.raw(struct.pack("<3d", *wd.acc[k]))


def _get_world(r: Reader) -> World:
    n = r.u64()
    idx, gen = np.zeros(n, np.uint32), np.zeros(n, np.uint32)
    pos, vel, acc = (np.full((n, 3), np.nan) for _ in range(3))
    mass, charge = np.full(n, np.nan), np.full(n, np.nan)
    species = np.full(n, 255, np.uint8)
    bonds: List[Optional[List[Tuple[int, int]]]] = []

    def vec3():
        return struct.unpack("<3d", r.take(24))
    for k in range(n):
        idx[k], gen[k] = r.u32(), r.u32()
        p = r.option(vec3)
        if p is not None:
            pos[k] = p
        v = r.option(vec3)
        if v is not None:
            vel[k] = v
        m = r.option(r.f64)
        if m is not None:
            mass[k] = m
        q = r.option(r.f64)
        if q is not None:
            charge[k] = q
        t = r.option(r.u32)
        if t is not None:
            if t >= NSPECIES:
                raise ValueError("AtomType variant out of range")
            species[k] = t
        bonds.append(r.option(lambda: r.seq(lambda: (r.u32(), r.u32()))))
        a = r.option(vec3)
        if a is not None:
            acc[k] = a
    return World(idx, gen, pos, vel, mass, charge, species, bonds, acc)


# ---- whole snapshots ----------------------------------------------------------------
def encode_snapshot(config: Config, state: UniverseState, world: World) -> bytes:
    """bincode::serialize(&Snapshot) (src/serialization/mod.rs:101)."""
    w = Writer()
    w.u32(FORMAT_VERSION)
    _put_config(w, config)
    _put_state(w, state)
    _put_world(w, world)
    return w.bytes()


def decode_snapshot(data: bytes):
    """bincode::deserialize of a Snapshot and the format check of
    load_universe (:108-114); returns (config, state, world, end offset)."""
    r = Reader(data)
    fmt = r.u32()
    if fmt != FORMAT_VERSION:
        raise ValueError(f"incompatible snapshot format (v{fmt})")
    config = _get_config(r)
    state = _get_state(r)
    world = _get_world(r)
    return config, state, world, r.pos


def world_of(state: State, mass_table, bonds_csr=None) -> World:
    """EntitySnapshot columns of a State (entity i = slot i, generation 0,
    every component present as spawn_atom inserts them,
    src/universe/mod.rs:247-273)."""
    n = state.n
    pos = np.stack([state.x, state.y, state.z], 1)
    vel = np.stack([state.vx, state.vy, state.vz], 1)
    ax, ay, az = state.accel()
    acc = np.stack([ax, ay, az], 1)
    mass = np.asarray(mass_table, dtype=np.float64)[state.species]
    if bonds_csr is None:
        bonds = [[] for _ in range(n)]
    else:
        ptr, idx = bonds_csr
        bonds = [[(int(m), 0) for m in idx[ptr[e]:ptr[e + 1]]] for e in range(n)]
    return World(np.arange(n, dtype=np.uint32), np.zeros(n, np.uint32), pos, vel, mass,
                 np.asarray(state.charge, dtype=np.float64), np.asarray(state.species, dtype=np.uint8), bonds, acc)


def save_universe(u, path: str, config: Optional[Config] = None, state: Optional[UniverseState] = None) -> None:
    """save_universe (src/serialization/mod.rs:84-105) of a GpuUniverse: the
    components of every entity downloaded from the device, the Bonds
    adjacency, the tick; then the device image."""
    ph = u.physics
    st = u.download()
    ptr, idx = u.bonds() if (ph.enable_bond_observation or ph.enable_bond_interaction) else (None, None)
    world = world_of(st, ph.mass, None if ptr is None else (ptr, idx))
    cfg = config if config is not None else Config.from_physics(ph, st.n)
    if state is None:
        tick = u.tick()
        state = UniverseState(t=advance_time(0.0, 0, tick, ph.dt), tick=tick, dt=ph.dt, rng=splitmix_state(cfg.seed),
                              stats=StatsCollector(cfg.stats_history))
    data = encode_snapshot(cfg, state, world)
    tmp = path + ".tmp"
    with open(tmp, "wb") as f:
        f.write(data)
    _lib.check(u.lib.gs_snapshot_append(u.h, tmp.encode()))
    os.replace(tmp, path)


def read_snapshot(path: str):
    """(config, state, world, device image offset or None) of a file."""
    with open(path, "rb") as f:
        data = f.read()
    config, state, world, end = decode_snapshot(data)
    dev = end if data[end:end + len(DEVICE_MAGIC)] == DEVICE_MAGIC else None
    return config, state, world, dev


def load_universe(path: str, grid_semantics: int = GRID_REFERENCE, use_device_image: bool = True):
    """load_universe (src/serialization/mod.rs:108-120) onto the GPU: the
    Config's physics with the masses of the Mass components, the entities'
    components uploaded (entity index = slot), the Bonds adjacency, the
    tick; then, when present, the device image.  Returns (universe,
    config, state)."""
    from .universe import GpuUniverse
    config, ustate, world, dev = read_snapshot(path)
    n = world.n
    order = np.argsort(world.index, kind="stable")
    if not np.array_equal(world.index[order], np.arange(n)):
        raise ValueError("entity indices must be 0..n-1")
    w = world
    for name, col in (("position", w.pos), ("velocity", w.vel)):
        if np.isnan(col).any():
            raise ValueError(f"every entity needs a {name}")
    if (w.species == 255).any() or np.isnan(w.mass).any():
        raise ValueError("every entity needs an atom type and a mass")
    species = w.species[order]
    mass_table = list(Physics().mass)
    for s in range(NSPECIES):
        ms = w.mass[order][species == s]
        if ms.size:
            if not np.all(ms == ms[0]):
                raise ValueError("entities of one species with different masses")
            mass_table[s] = float(ms[0])
    ph = config.to_physics(mass_table, grid_semantics)
    charge = np.where(np.isnan(w.charge), 0.0, w.charge)[order]
    acc = np.where(np.isnan(w.acc), 0.0, w.acc)[order]
    st = State(w.pos[order, 0].copy(), w.pos[order, 1].copy(), w.pos[order, 2].copy(), w.vel[order, 0].copy(),
               w.vel[order, 1].copy(), w.vel[order, 2].copy(), species.copy(), charge, acc[:, 0].copy(),
               acc[:, 1].copy(), acc[:, 2].copy())
    u = GpuUniverse(ph, st)
    bonds = [w.bonds[k] or [] for k in order]
    if any(bonds) and (ph.enable_bond_observation or ph.enable_bond_interaction):
        ptr = np.zeros(n + 1, dtype=np.int32)
        ptr[1:] = np.cumsum([len(b) for b in bonds])
        idx = np.array([i for b in bonds for i, _ in b], dtype=np.int32)
        u.upload_bonds(ptr, idx)
    if dev is not None and use_device_image:
        _lib.check(u.lib.gs_snapshot_load(u.h, path.encode(), ctypes.c_int64(dev)))
    else:
        u.set_tick(ustate.tick)
    return u, config, ustate
