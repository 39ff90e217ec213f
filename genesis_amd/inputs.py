"""Seeded synthetic inputs for the benchmark workloads.

The reference seeds through rand 0.8 + rand_xoshiro 0.6
(src/universe/mod.rs:195-273), a stream this package does not reproduce:
inputs are generated here and handed to both the GPU path and the checker, so
comparisons are always on identical inputs.

Randomness is SplitMix64 used as a counter-based generator
(Steele, Lea & Flood, "Fast splittable pseudorandom number generators",
OOPSLA 2014): draw i of stream s is mix64(key(seed, s) + (i+1)*golden).
uniform = (x >> 11) * 2^-53; normals by Box-Muller, the method the reference
uses for Maxwell-Boltzmann seeding (src/rng/mod.rs:50-54).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from .params import (DEFAULT_CHARGE, DEFAULT_MASS, DEFAULT_SIGMA, GRID_MINIMAGE, NSPECIES, SPECIES_INDEX,
                     Physics, fs_to_time)

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix64(z: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


class Stream:
    """Counter-based SplitMix64 stream ``(seed, stream_id)``."""

    def __init__(self, seed: int, stream: int):
        key = np.array([(seed * 0x100000001B3 + stream * 0x9E3779B1 + 0x632BE59BD9B4E019) & 0xFFFFFFFFFFFFFFFF],
                       dtype=np.uint64)
        self.key = _mix64(key)[0]
        self.pos = 0

    def raw(self, n: int) -> np.ndarray:
        i = np.arange(self.pos + 1, self.pos + n + 1, dtype=np.uint64)
        self.pos += n
        with np.errstate(over="ignore"):
            return _mix64(self.key + i * _GOLDEN)

    def uniform(self, n: int) -> np.ndarray:
        return (self.raw(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    def normal(self, n: int) -> np.ndarray:
        u1 = np.maximum(self.uniform(n), 2.2250738585072014e-308)
        u2 = self.uniform(n)
        return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)

    def permutation(self, n: int) -> np.ndarray:
        return np.argsort(self.raw(n), kind="stable")


@dataclass
class State:
    """SoA state in entity (slot) order, as the ECS columns hold it."""

    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    vx: np.ndarray
    vy: np.ndarray
    vz: np.ndarray
    species: np.ndarray  # uint8 element index
    charge: np.ndarray
    ax: Optional[np.ndarray] = None
    ay: Optional[np.ndarray] = None
    az: Optional[np.ndarray] = None

    @property
    def n(self) -> int:
        return int(self.x.shape[0])

    def copy(self) -> "State":
        return State(*(None if v is None else v.copy() for v in
                       (self.x, self.y, self.z, self.vx, self.vy, self.vz, self.species, self.charge,
                        self.ax, self.ay, self.az)))

    def accel(self):
        if self.ax is None:
            z = np.zeros(self.n)
            return z, z.copy(), z.copy()
        return self.ax, self.ay, self.az


@dataclass
class Workload:
    name: str
    physics: Physics
    state: State
    steps: int
    extra: dict


MIX10 = (0.20, 0.05, 0.15, 0.10, 0.20, 0.10, 0.05, 0.05, 0.05, 0.05)


def largest_remainder(n: int, fractions) -> List[int]:
    """Exact species counts: floors, then the largest remainders (ties by index)."""
    raw = [n * f for f in fractions]
    counts = [int(math.floor(r)) for r in raw]
    rem = n - sum(counts)
    order = sorted(range(len(raw)), key=lambda i: (-(raw[i] - counts[i]), i))
    for i in order[:rem]:
        counts[i] += 1
    return counts


def mb_velocities(stream: Stream, species: np.ndarray, mass: tuple, temperature: float, k: float,
                  remove_com: bool = True, rescale: bool = True):
    """Maxwell-Boltzmann: each component ~ N(0, sqrt(k T / m)) as
    spawn_atom does (src/universe/mod.rs:265-270); then optional removal of
    the centre-of-mass velocity and an exact rescale to ``temperature``."""
    n = species.shape[0]
    m = np.asarray(mass, dtype=np.float64)[species]
    s = np.sqrt(k * temperature / m)
    v = [stream.normal(n) * s for _ in range(3)]
    if remove_com and n > 0:
        mt = m.sum()
        for a in range(3):
            v[a] = v[a] - (m * v[a]).sum() / mt
    if rescale and n > 0 and temperature > 0:
        kin = 0.5 * (m * (v[0] ** 2 + v[1] ** 2 + v[2] ** 2)).sum()
        t_now = (2.0 / 3.0) * (kin / n) / k
        if t_now > 0:
            f = math.sqrt(temperature / t_now)
            v = [c * f for c in v]
    return v


def lattice_sites(nx: int, ny: int, nz: int, box, count: int):
    """First ``count`` sites of an nx*ny*nz simple-cubic lattice centred in the
    box [-L/2, L/2), x fastest (cell-centred like seed_lattice,
    src/universe/mod.rs:221-243)."""
    i = np.arange(count)
    ix, iy, iz = i % nx, (i // nx) % ny, i // (nx * ny)
    ax = (ix + 0.5) * (box[0] / nx) - 0.5 * box[0]
    ay = (iy + 0.5) * (box[1] / ny) - 0.5 * box[1]
    az = (iz + 0.5) * (box[2] / nz) - 0.5 * box[2]
    return ax.astype(np.float64), ay.astype(np.float64), az.astype(np.float64), (ix, iy, iz)


def _jitter(stream: Stream, n: int, amp: float):
    return [(stream.uniform(n) * 2.0 - 1.0) * amp for _ in range(3)]


def _mean_sigma(fractions, sigma=DEFAULT_SIGMA) -> float:
    return float(sum(f * s for f, s in zip(fractions, sigma)))


def plain_lj(seed: int = 1001, n_side: int = 16, steps: int = 1000) -> Workload:
    """Plain NVE LJ, one species (sigma 3.4, eps/kB 120 K, mass 12 u) on an
    n_side^3 simple-cubic lattice at rho sigma^3 = 0.8, +-0.05 sigma jitter,
    MB 100 K with COM momentum removed, rc = 2.5 sigma with the quintic switch,
    dt = 1 fs, no Coulomb, no thermostat.  n_side = 16 gives 4,096 atoms."""
    sig = 3.4
    n = n_side ** 3
    L = (n / 0.8) ** (1.0 / 3.0) * sig
    box = (L, L, L)
    ph = Physics(box=box, dt=fs_to_time(1.0)).with_overrides({"C": {"sigma": sig, "epsilon_k": 120.0, "mass": 12.0}})
    st = Stream(seed, 0)
    x, y, z, _ = lattice_sites(n_side, n_side, n_side, box, n)
    j = _jitter(st, n, 0.05 * sig)
    x, y, z = x + j[0], y + j[1], z + j[2]
    species = np.full(n, SPECIES_INDEX["C"], dtype=np.uint8)
    v = mb_velocities(Stream(seed, 1), species, ph.mass, 100.0, ph.thermal_constant)
    charge = np.zeros(n)
    return Workload("plain-nve-lj-fp64-%d" % n, ph, State(x, y, z, v[0], v[1], v[2], species, charge), steps,
                    {"n_side": n_side})


def _mixture_species(stream: Stream, n: int, fractions=MIX10) -> np.ndarray:
    counts = largest_remainder(n, fractions)
    species = np.concatenate([np.full(c, t, dtype=np.uint8) for t, c in enumerate(counts)])
    return species[stream.permutation(n)]


def _neutralise(stream: Stream, species: np.ndarray, charge: np.ndarray) -> np.ndarray:
    """Zero the per-entity charge of seeded-random oxygens (net negative) or
    cations (net positive) until the total charge is zero, as far as the
    available charges allow.  Charge is a per-entity component
    (src/systems/forces.rs:196-200), so this is a legal input."""
    q = charge.copy()
    net = q.sum()
    order = stream.permutation(species.shape[0])
    if net < 0:
        for i in order:
            if net > -0.5:
                break
            if q[i] == -1.0:
                q[i] = 0.0
                net += 1.0
    elif net > 0:
        for i in order:
            if net < 0.5:
                break
            if q[i] > 0 and q[i] <= net + 1e-12:
                net -= q[i]
                q[i] = 0.0
    return q


def mixture(seed: int = 1002, n: int = 32768, temperature: float = 60.0, steps: int = 5000,
            thermostat: bool = True, dt_fs: float = 0.5, tau_fs: float = 100.0,
            lattice: Optional[tuple] = None) -> Workload:
    """10-element mixture (H 20%, He 5%, C 15%, N 10%, O 20%, Na 10%, Si 5%,
    P 5%, S 5%, Fe 5%) with the default table and Lorentz-Berthelot mixing,
    charges O -1 / Na +1 / Si +0.5 / Fe +1 neutralised by seeded zeroing,
    Coulomb and LJ sharing rc = 2.5 max sigma, lattice +-0.05 mean-sigma jitter
    at rho <sigma>^3 = 0.8, Berendsen at ``temperature`` with tau = 100 fs.
    n = 32,768 on a 32^3 lattice is the 32k-atom mixture workload."""
    sbar = _mean_sigma(MIX10)
    rho = 0.8 / sbar ** 3
    L = (n / rho) ** (1.0 / 3.0)
    box = (L, L, L)
    m = lattice[0] if lattice else int(math.ceil(round(n ** (1.0 / 3.0), 9)))
    dt = fs_to_time(dt_fs)
    ph = Physics(box=box, dt=dt, enable_electrostatics=True, enable_thermostat=thermostat,
                 thermostat_temperature=temperature, thermostat_tau=tau_fs / dt_fs)
    st = Stream(seed, 0)
    x, y, z, _ = lattice_sites(m, m, m, box, n)
    j = _jitter(st, n, 0.05 * sbar)
    x, y, z = x + j[0], y + j[1], z + j[2]
    species = _mixture_species(Stream(seed, 2), n)
    charge = np.asarray(DEFAULT_CHARGE, dtype=np.float64)[species]
    charge = _neutralise(Stream(seed, 3), species, charge)
    v = mb_velocities(Stream(seed, 1), species, ph.mass, temperature, ph.thermal_constant)
    return Workload("mix10-coulomb-berendsen-fp64-%d" % n, ph,
                    State(x, y, z, v[0], v[1], v[2], species, charge), steps,
                    {"gr_bins": 200, "gr_rmax": L / 4.0, "gr_every": 100})


def replica_sweep(seed: int = 3000, n_rep: int = 64, lo: int = 512, hi: int = 16384, steps: int = 2000):
    """Batched sweep: replica r has N_r = round(exp(U[ln lo, ln hi])) atoms of
    the 10-element mix at the mixture density and T_r = 20 + 100 r/(n_rep-1) K
    (Berendsen tau = 100 fs, dt = 0.5 fs).  Replicas too small for 3 cells per
    axis count every pair once (min-image semantics), not with the
    reference grid's 2-cell multiplicity."""
    u = Stream(seed, 99).uniform(n_rep)
    sizes = np.rint(np.exp(np.log(lo) + u * (np.log(hi) - np.log(lo)))).astype(int)
    out = []
    for r in range(n_rep):
        t = 20.0 + 100.0 * r / max(n_rep - 1, 1)
        w = mixture(seed=seed + r, n=int(sizes[r]), temperature=t, steps=steps)
        w.physics.grid_semantics = GRID_MINIMAGE
        out.append(w)
    return out


def adjoint_mixture(seed: int = 1004, n: int = 8192, steps: int = 1000) -> Workload:
    """8,192 atoms of the 10-element mixture, NVE, dt = 0.5 fs (first 8,192
    sites of a 21^3 lattice)."""
    w = mixture(seed=seed, n=n, temperature=60.0, steps=steps, thermostat=False, lattice=(21,))
    w.name = "adjoint-mix10-fp64-%d" % n
    return w


def ionic_melt(seed: int = 1005, nx: int = 128, ny: int = 128, nz: int = 64, steps: int = 20000) -> Workload:
    """Na/O rock salt (alternating by lattice parity, exactly neutral) on an
    nx*ny*nz simple-cubic lattice at rho sigma_NaO^3 = 0.7 in a box of the
    lattice's aspect ratio, +-0.05 sigma_NaO jitter, MB 300 K, Berendsen to
    80 K (tau = 100 fs), dt = 0.5 fs, bond observation and interaction on.
    128*128*64 = 1,048,576 atoms."""
    n = nx * ny * nz
    s_nao = 0.5 * (DEFAULT_SIGMA[SPECIES_INDEX["Na"]] + DEFAULT_SIGMA[SPECIES_INDEX["O"]])
    rho = 0.7 / s_nao ** 3
    a = (n / rho / (nx * ny * nz)) ** (1.0 / 3.0)  # lattice spacing
    box = (nx * a, ny * a, nz * a)
    ph = Physics(box=box, dt=fs_to_time(0.5), enable_electrostatics=True, enable_thermostat=True,
                 thermostat_temperature=80.0, thermostat_tau=200.0, enable_bond_observation=True,
                 enable_bond_interaction=True)
    st = Stream(seed, 0)
    x, y, z, (ix, iy, iz) = lattice_sites(nx, ny, nz, box, n)
    j = _jitter(st, n, 0.05 * s_nao)
    x, y, z = x + j[0], y + j[1], z + j[2]
    species = np.where((ix + iy + iz) % 2 == 0, SPECIES_INDEX["Na"], SPECIES_INDEX["O"]).astype(np.uint8)
    charge = np.asarray(DEFAULT_CHARGE, dtype=np.float64)[species]
    v = mb_velocities(Stream(seed, 1), species, ph.mass, 300.0, ph.thermal_constant)
    return Workload("ionic-melt-bonds-fp64-%d" % n, ph, State(x, y, z, v[0], v[1], v[2], species, charge),
                    steps, {"gr_bins": 256, "gr_rmax": min(box) / 4.0, "components_every": 500})


def random_gas(seed: int, n: int, box, species_pool=(0,), temperature: float = 100.0, k: float = 0.01,
               mass=DEFAULT_MASS, min_sep: float = 0.0) -> State:
    """Uniform random positions in [-L/2, L/2) (Rng::in_box shape), uniform
    species from a pool, MB velocities.  Optional rejection of pairs closer
    than min_sep (O(n^2), small n only)."""
    st = Stream(seed, 7)
    pts = []
    if min_sep > 0:
        out = np.zeros((0, 3))
        while out.shape[0] < n:
            cand = (st.uniform(3) - 0.5) * np.asarray(box)
            if out.shape[0]:
                d = out - cand
                d -= np.asarray(box) * np.round(d / np.asarray(box))
                if (np.einsum("ij,ij->i", d, d) < min_sep * min_sep).any():
                    continue
            out = np.vstack([out, cand])
        x, y, z = out[:, 0].copy(), out[:, 1].copy(), out[:, 2].copy()
    else:
        x = (st.uniform(n) - 0.5) * box[0]
        y = (st.uniform(n) - 0.5) * box[1]
        z = (st.uniform(n) - 0.5) * box[2]
    pool = np.asarray(species_pool, dtype=np.uint8)
    species = pool[(Stream(seed, 8).raw(n) % np.uint64(len(pool))).astype(np.int64)]
    v = mb_velocities(Stream(seed, 9), species, mass, temperature, k, remove_com=False, rescale=False)
    charge = np.asarray(DEFAULT_CHARGE, dtype=np.float64)[species]
    return State(x, y, z, v[0], v[1], v[2], species, charge)
