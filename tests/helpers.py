"""Shared helpers for the tests (inputs shaped like the reference's tests)."""
import math

import numpy as np

from genesis_amd.inputs import State, Stream, lattice_sites
from genesis_amd.params import DEFAULT_CHARGE, DEFAULT_MASS, Physics


def reference_like(n_side, L, temperature, seed, species_pool=tuple(range(10)), k=0.01):
    """The reference's seed_lattice shape (src/universe/mod.rs:221-273):
    n_side^3 cell-centred sites, jitter uniform in +-0.1 cell per axis,
    species uniform from a pool, per-species charge, MB velocities
    N(0, sqrt(k T / m)) with no COM removal, a = 0."""
    n = n_side ** 3
    box = (L, L, L)
    x, y, z, _ = lattice_sites(n_side, n_side, n_side, box, n)
    st = Stream(seed, 0)
    cell = L / n_side
    x = x + (st.uniform(n) * 2 - 1) * 0.1 * cell
    y = y + (st.uniform(n) * 2 - 1) * 0.1 * cell
    z = z + (st.uniform(n) * 2 - 1) * 0.1 * cell
    pool = np.asarray(species_pool, dtype=np.uint8)
    sp = pool[(Stream(seed, 1).raw(n) % np.uint64(len(pool))).astype(np.int64)]
    m = np.asarray(DEFAULT_MASS)[sp]
    s = np.sqrt(k * temperature / m)
    g = Stream(seed, 2)
    vx, vy, vz = g.normal(n) * s, g.normal(n) * s, g.normal(n) * s
    q = np.asarray(DEFAULT_CHARGE, dtype=np.float64)[sp]
    return State(x, y, z, vx, vy, vz, sp, q)


def force_tolerance_ok(ax_g, ay_g, az_g, ref, mass_of_atom, rel=1e-10):
    """Per-atom force agreement: |a_gpu - a_cpu|_inf <= rel * max(|a_cpu|_inf, S_i/m_i)."""
    a_ref = np.stack([ref["ax"], ref["ay"], ref["az"]], 1)
    a_g = np.stack([ax_g, ay_g, az_g], 1)
    scale = np.maximum(np.abs(a_ref).max(1), ref["sabs"] / mass_of_atom)
    err = np.abs(a_g - a_ref).max(1)
    worst = float((err / np.maximum(scale, 1e-300)).max()) if len(err) else 0.0
    return bool((err <= rel * scale + 1e-300).all()), worst
