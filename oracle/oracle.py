"""ctypes wrapper of the CPU oracle (liboracle.so built from oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_D = ctypes.POINTER(ctypes.c_double)
_U8 = ctypes.POINTER(ctypes.c_uint8)
_U32 = ctypes.POINTER(ctypes.c_uint32)
_U64 = ctypes.POINTER(ctypes.c_uint64)
_I32 = ctypes.POINTER(ctypes.c_int32)
_I64 = ctypes.POINTER(ctypes.c_int64)

_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-B", "-C", HERE], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        from genesis_amd.params import GsParams
        P = ctypes.POINTER(GsParams)
        sig = {
            "or_lj_raw": (None, [ctypes.c_double] * 3 + [_D, _D]),
            "or_smooth_cutoff": (None, [ctypes.c_double] * 3 + [_D, _D]),
            "or_switched": (None, [ctypes.c_double] * 4 + [_D, _D]),
            "or_coulomb_raw": (None, [ctypes.c_double] * 4 + [_D, _D]),
            "or_gravity_raw": (None, [ctypes.c_double] * 4 + [_D, _D]),
            "or_bond_harmonic_raw": (None, [ctypes.c_double] * 3 + [_D, _D]),
            "or_well_curvature": (ctypes.c_double, [ctypes.c_double] * 2),
            "or_vib_period": (ctypes.c_double, [ctypes.c_double] * 3),
            "or_equilibrium_distance": (ctypes.c_double, [ctypes.c_double]),
            "or_lj_table": (None, [P, _D, _D, _D]),
            "or_force_switched_scalar": (None, [P, ctypes.c_int, ctypes.c_int, ctypes.c_double, _D, _D]),
            "or_min_image": (None, [_D, _D, _D]),
            "or_grid_cells": (None, [ctypes.c_int64, _D, _D, _D, _D, ctypes.c_double, _U32, _U32]),
            "or_pairs": (ctypes.c_int64, [ctypes.c_int64, _D, _D, _D, _D, ctypes.c_double, ctypes.c_int,
                                          ctypes.c_int64, _U32, _U32]),
            "or_forces": (ctypes.c_int, [P, ctypes.c_int64, _D, _D, _D, _U8, _D, _I32, _I32, ctypes.c_int,
                                         _D, _D, _D, _D, _D, _D, _D, _I64]),
            "or_half_kick": (None, [ctypes.c_int64, _D, _D, _D, _D, _D, _D, ctypes.c_double]),
            "or_drift": (None, [ctypes.c_int64, _D, _D, _D, _D, _D, _D, ctypes.c_double]),
            "or_wrap": (None, [ctypes.c_int64, _D, _D, _D, _D]),
            "or_kinetic": (ctypes.c_double, [ctypes.c_int64, _D, _D, _D, _U8, _D]),
            "or_thermostat": (ctypes.c_double, [P, ctypes.c_int64, _D, _D, _D, _U8]),
            "or_run": (ctypes.c_int, [P, ctypes.c_int64] + [_D] * 9 + [_U8, _D, ctypes.c_int64, ctypes.c_int, _D]),
            "or_radial_distribution": (ctypes.c_int, [ctypes.c_int64, _D, _D, _D, _U8, _D, ctypes.c_double,
                                                      ctypes.c_int, ctypes.c_int, ctypes.c_int, _U64, _D, _D, _D,
                                                      ctypes.c_int]),
            "or_clusters": (ctypes.c_int, [ctypes.c_int64, _D, _D, _D, _U8, _D, _D, ctypes.c_double, _U32, _D]),
            "or_pair_derivs": (None, [ctypes.c_double] * 5 + [ctypes.c_int, _D]),
            "or_adj_accel": (ctypes.c_double, [P, ctypes.c_int64, _D, _D, _D, _U8, _U8, _D, _D, _D, _D]),
            "or_adj_pass": (None, [P, ctypes.c_int64, _D, _D, _D, _U8, _U8, _D, _D, _D, _D, ctypes.c_double,
                                   _D, _D, _D, _D]),
            "or_adjoint": (ctypes.c_int, [P, ctypes.c_int64] + [_D] * 6 + [_U8, _U8, _D, ctypes.c_int64,
                                          ctypes.c_int64, ctypes.c_int, ctypes.c_double, ctypes.c_double, _D,
                                          _D, _D, _D]),
            "or_bond_components": (ctypes.c_int, [ctypes.c_int64, _U32, _U8, ctypes.c_int64, _U32, _U32, _I64, _U32,
                                                  ctypes.c_char_p, ctypes.c_int64]),
            "or_run_observed": (ctypes.c_int, [P, ctypes.c_int64] + [_D] * 9 + [_U8, _D, ctypes.c_int64, ctypes.c_int,
                                               _D, ctypes.c_int64, _D, _I32, _I32, ctypes.c_int64]),
            "or_observe_frames": (ctypes.c_int, [P, ctypes.c_int64, ctypes.c_int64, _D, _D, _D, _U8, ctypes.c_int64,
                                                 _D, _I32, _I32, ctypes.c_int64]),
            "or_bound_pairs": (ctypes.c_int64, [P, ctypes.c_int64, _D, _D, _D, _U8, _U32, _U32, ctypes.c_int64]),
            "or_pair_potential_raw": (ctypes.c_double, [P, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                                        ctypes.c_int]),
            "or_bond_binding": (None, [P, ctypes.c_int64, _D, _D, _D, _U8, ctypes.c_int64, _U32, _U32, _D]),
            "or_persistence_ticks": (ctypes.c_uint64, [P, ctypes.c_int, ctypes.c_int]),
            "or_track": (ctypes.c_int64, [ctypes.c_int64, _I64, _U32, _U32, ctypes.c_uint64, _U32, _U32, _U64,
                                          ctypes.c_int64]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _d(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def _p(a, t=ctypes.c_double):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(t))


def out2(fn, *args):
    a, b = ctypes.c_double(), ctypes.c_double()
    fn(*args, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def lj_raw(sigma, eps, r):
    return out2(lib().or_lj_raw, sigma, eps, r)


def smooth_cutoff(r, r_on, rc):
    return out2(lib().or_smooth_cutoff, r, r_on, rc)


def switched(f, v, sw, dsw):
    return out2(lib().or_switched, f, v, sw, dsw)


def coulomb_raw(k, qa, qb, r):
    return out2(lib().or_coulomb_raw, k, qa, qb, r)


def gravity_raw(g, ma, mb, r):
    return out2(lib().or_gravity_raw, g, ma, mb, r)


def bond_harmonic_raw(k, r_eq, r):
    return out2(lib().or_bond_harmonic_raw, k, r_eq, r)


def well_curvature(eps, sigma):
    return lib().or_well_curvature(eps, sigma)


def vib_period(eps, sigma, mu):
    return lib().or_vib_period(eps, sigma, mu)


def equilibrium_distance(sigma):
    return lib().or_equilibrium_distance(sigma)


def lj_table(physics):
    p = physics.to_c()
    sig, eps, cut = np.zeros(100), np.zeros(100), np.zeros(6)
    lib().or_lj_table(ctypes.byref(p), _p(sig), _p(eps), _p(cut))
    return sig.reshape(10, 10), eps.reshape(10, 10), dict(zip(
        ("lj_rc", "lj_ron", "rc", "coulomb_rc", "coulomb_ron", "gravity_rc"), cut.tolist()))


def force_switched_scalar(physics, ta, tb, r):
    p = physics.to_c()
    return out2(lib().or_force_switched_scalar, ctypes.byref(p), ta, tb, r)


def min_image(d, size):
    d, s, o = _d(d), _d(size), np.zeros(3)
    lib().or_min_image(_p(d), _p(s), _p(o))
    return o


def grid_cells(x, y, z, size, min_cell):
    n = len(x)
    x, y, z, s = _d(x), _d(y), _d(z), _d(size)
    dims, cell = np.zeros(3, dtype=np.uint32), np.zeros(n, dtype=np.uint32)
    lib().or_grid_cells(n, _p(x), _p(y), _p(z), _p(s), min_cell, _p(dims, ctypes.c_uint32), _p(cell, ctypes.c_uint32))
    return cell, tuple(int(v) for v in dims)


def pairs(x, y, z, size, cutoff, brute=False):
    n = len(x)
    x, y, z, s = _d(x), _d(y), _d(z), _d(size)
    total = lib().or_pairs(n, _p(x), _p(y), _p(z), _p(s), cutoff, 1 if brute else 0, 0, None, None)
    a, b = np.zeros(max(total, 1), dtype=np.uint32), np.zeros(max(total, 1), dtype=np.uint32)
    lib().or_pairs(n, _p(x), _p(y), _p(z), _p(s), cutoff, 1 if brute else 0, total, _p(a, ctypes.c_uint32),
                   _p(b, ctypes.c_uint32))
    return a[:total], b[:total]


def forces(physics, state, nthreads=1, bonds_csr=None, extras=True):
    """ForceSystem::run on ``state``.  Returns dict with ax, ay, az, V, Vb,
    pe, sabs, npairs (slot order)."""
    n = state.n
    p = physics.to_c()
    x, y, z, q = _d(state.x), _d(state.y), _d(state.z), _d(state.charge)
    sp = np.ascontiguousarray(state.species, dtype=np.uint8)
    ax, ay, az = np.zeros(n), np.zeros(n), np.zeros(n)
    pe = np.zeros(n) if extras else None
    sa = np.zeros(n) if extras else None
    V, Vb, npairs = ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
    bp = bi = None
    if bonds_csr is not None:
        bp = np.ascontiguousarray(bonds_csr[0], dtype=np.int32)
        bi = np.ascontiguousarray(bonds_csr[1], dtype=np.int32)
    lib().or_forces(ctypes.byref(p), n, _p(x), _p(y), _p(z), _p(sp, ctypes.c_uint8), _p(q), _p(bp, ctypes.c_int32),
                    _p(bi, ctypes.c_int32), nthreads, _p(ax), _p(ay), _p(az), ctypes.byref(V), ctypes.byref(Vb),
                    _p(pe), _p(sa), ctypes.byref(npairs))
    return {"ax": ax, "ay": ay, "az": az, "V": V.value, "Vb": Vb.value, "pe": pe, "sabs": sa,
            "npairs": npairs.value}


def run(physics, state, nsteps, nthreads=1, observe=True):
    """Universe::tick x nsteps on a copy of ``state``; returns (state, obs)."""
    s = state.copy()
    n = s.n
    if s.ax is None:
        s.ax, s.ay, s.az = np.zeros(n), np.zeros(n), np.zeros(n)
    arrs = [_d(a) for a in (s.x, s.y, s.z, s.vx, s.vy, s.vz, s.ax, s.ay, s.az)]
    sp = np.ascontiguousarray(s.species, dtype=np.uint8)
    q = _d(s.charge)
    obs = np.zeros((max(nsteps, 1), 6)) if observe else None
    p = physics.to_c()
    lib().or_run(ctypes.byref(p), n, *[_p(a) for a in arrs], _p(sp, ctypes.c_uint8), _p(q), nsteps, nthreads,
                 _p(obs))
    s.x, s.y, s.z, s.vx, s.vy, s.vz, s.ax, s.ay, s.az = arrs
    return s, (obs[:nsteps] if observe else None)


BOND_STAT_NAMES = ("tick", "bonded_pairs", "bonded_entities", "bonds_formed", "lifetime_sum_ticks",
                   "mean_coordination")


def run_observed(physics, state, nsteps, nthreads=1, tick0=0, bond_cap=None):
    """run() with the bond observation of every tick (physics.enable_bond_observation):
    returns (state, obs, bstats, (ptr, idx)) where bstats is (nsteps, 106): the six
    BOND_STAT_NAMES then the 10x10 species matrix, and (ptr, idx) the final Bonds CSR."""
    s = state.copy()
    n = s.n
    if s.ax is None:
        s.ax, s.ay, s.az = np.zeros(n), np.zeros(n), np.zeros(n)
    arrs = [_d(a) for a in (s.x, s.y, s.z, s.vx, s.vy, s.vz, s.ax, s.ay, s.az)]
    sp = np.ascontiguousarray(s.species, dtype=np.uint8)
    q = _d(s.charge)
    obs = np.zeros((max(nsteps, 1), 6))
    bst = np.zeros((max(nsteps, 1), 106))
    cap = 32 * max(n, 1) if bond_cap is None else bond_cap
    ptr = np.zeros(n + 1, dtype=np.int32)
    idx = np.zeros(max(cap, 1), dtype=np.int32)
    p = physics.to_c()
    rc = lib().or_run_observed(ctypes.byref(p), n, *[_p(a) for a in arrs], _p(sp, ctypes.c_uint8), _p(q), nsteps,
                               nthreads, _p(obs), tick0, _p(bst), _p(ptr, ctypes.c_int32), _p(idx, ctypes.c_int32),
                               cap)
    if rc != 0:
        raise RuntimeError("bond adjacency exceeds bond_cap")
    s.x, s.y, s.z, s.vx, s.vy, s.vz, s.ax, s.ay, s.az = arrs
    return s, obs[:nsteps], bst[:nsteps], (ptr, idx[:ptr[n]])


def observe_frames(physics, frames, species, tick0=0, with_bonds=False):
    """BondObservationSystem alone over position frames [(x, y, z), ...]; returns (F, 106)
    stats (and the final Bonds CSR when with_bonds)."""
    nf = len(frames)
    n = len(frames[0][0]) if nf else 0
    X = np.ascontiguousarray(np.concatenate([f[0] for f in frames]) if nf else np.zeros(1))
    Y = np.ascontiguousarray(np.concatenate([f[1] for f in frames]) if nf else np.zeros(1))
    Z = np.ascontiguousarray(np.concatenate([f[2] for f in frames]) if nf else np.zeros(1))
    sp = np.ascontiguousarray(species, dtype=np.uint8)
    bst = np.zeros((max(nf, 1), 106))
    cap = 32 * max(n, 1)
    ptr = np.zeros(n + 1, dtype=np.int32)
    idx = np.zeros(cap, dtype=np.int32)
    p = physics.to_c()
    rc = lib().or_observe_frames(ctypes.byref(p), n, nf, _p(X), _p(Y), _p(Z), _p(sp, ctypes.c_uint8), tick0,
                                 _p(bst), _p(ptr, ctypes.c_int32), _p(idx, ctypes.c_int32), cap)
    if rc != 0:
        raise RuntimeError("bond adjacency exceeds its capacity")
    if with_bonds:
        return bst[:nf], (ptr, idx[:ptr[n]])
    return bst[:nf]


def bound_pairs(physics, x, y, z, species):
    """bound_pairs_with_grid on one state: sorted (a, b) pairs, a < b."""
    x, y, z = _d(x), _d(y), _d(z)
    sp = np.ascontiguousarray(species, dtype=np.uint8)
    n = len(x)
    p = physics.to_c()
    cap = 64 * max(n, 1)
    pa = np.zeros(cap, dtype=np.uint32)
    pb = np.zeros(cap, dtype=np.uint32)
    k = lib().or_bound_pairs(ctypes.byref(p), n, _p(x), _p(y), _p(z), _p(sp, ctypes.c_uint8),
                             _p(pa, ctypes.c_uint32), _p(pb, ctypes.c_uint32), cap)
    return np.stack([pa[:k], pb[:k]], 1)


def pair_potential_raw(physics, a, b, r, with_coulomb=None, with_bond=None):
    """pair_potential_raw (src/physics/forces.rs:360-384); the switches default to the physics'."""
    c = physics.to_c()
    wc = physics.enable_electrostatics if with_coulomb is None else with_coulomb
    wb = physics.enable_bond_interaction if with_bond is None else with_bond
    return lib().or_pair_potential_raw(ctypes.byref(c), int(a), int(b), float(r), int(bool(wc)), int(bool(wb)))


def bond_binding(physics, x, y, z, species, edges):
    """Binding energy of each edge (a, b) as BondStructureSystem::run evaluates it
    (src/systems/bond_structure.rs:101-115)."""
    c = physics.to_c()
    x, y, z = _d(x), _d(y), _d(z)
    sp = np.ascontiguousarray(species, dtype=np.uint8)
    e = np.asarray(edges, dtype=np.uint32).reshape(-1, 2)
    ea, eb = np.ascontiguousarray(e[:, 0]), np.ascontiguousarray(e[:, 1])
    out = np.zeros(len(e))
    lib().or_bond_binding(ctypes.byref(c), len(x), _p(x), _p(y), _p(z), _p(sp, ctypes.c_uint8), len(e),
                          _p(ea, ctypes.c_uint32), _p(eb, ctypes.c_uint32), _p(out))
    return out


def velocity_histogram(vx, vy, vz, max_speed, bins):
    """velocity_histogram (src/stats/mod.rs:269-297) in numpy: bins clamped to
    1..512, width max(max_speed, 1e-9) / bins, |v| = sqrt((x^2 + y^2) + z^2),
    speeds >= max_speed counted as overflow, the others in
    min((s / width) as usize, bins - 1)."""
    bins = min(max(int(bins), 1), 512)
    max_speed = max(float(max_speed), 1e-9)
    width = max_speed / bins
    vx, vy, vz = (np.asarray(a, dtype=np.float64) for a in (vx, vy, vz))
    s = np.sqrt(vx * vx + vy * vy + vz * vz)
    over = s >= max_speed
    q = s[~over] / width
    # `as usize`: NaN -> 0, saturating above; then the last bin
    idx = np.where(q > 0.0, np.minimum(q, 2.0 ** 63), 0.0).astype(np.uint64)
    idx = np.minimum(idx, np.uint64(bins - 1))
    counts = np.bincount(idx.astype(np.int64), minlength=bins).astype(np.uint64)
    return {"max_speed": max_speed, "bin_width": width, "bins": counts, "samples": int(s.size),
            "overflow": int(over.sum())}


def stats_sums(vx, vy, vz, mass):
    """StatsSystem's sums over Velocity and Mass (src/systems/aggregate.rs:87-93)
    in the reference's order, entity after entity: kinetic += (0.5 m) |v|^2
    and speed_sum += sqrt(|v|^2), |v|^2 = (x^2 + y^2) + z^2
    (Vec3::length_squared).  np.add.accumulate adds left to right."""
    vx, vy, vz, mass = (np.asarray(a, dtype=np.float64) for a in (vx, vy, vz, mass))
    if vx.size == 0:
        return 0.0, 0.0
    l2 = (vx * vx + vy * vy) + vz * vz
    ke = (0.5 * mass) * l2
    sp = np.sqrt(l2)
    return float(np.add.accumulate(ke)[-1]), float(np.add.accumulate(sp)[-1])


def persistence_ticks(physics, a, b):
    p = physics.to_c()
    return int(lib().or_persistence_ticks(ctypes.byref(p), a, b))


def track(ticks, debounce=2):
    """PairTracker fed with per-tick lists of (a, b) pairs; returns {(a, b): open ticks}."""
    counts = np.array([len(t) for t in ticks], dtype=np.int64)
    flat = [p for t in ticks for p in t]
    pa = np.array([p[0] for p in flat] or [0], dtype=np.uint32)
    pb = np.array([p[1] for p in flat] or [0], dtype=np.uint32)
    cap = max(len(flat), 1)
    oa = np.zeros(cap, dtype=np.uint32)
    ob = np.zeros(cap, dtype=np.uint32)
    ot = np.zeros(cap, dtype=np.uint64)
    m = lib().or_track(len(ticks), _p(counts, ctypes.c_int64), _p(pa, ctypes.c_uint32), _p(pb, ctypes.c_uint32),
                       debounce, _p(oa, ctypes.c_uint32), _p(ob, ctypes.c_uint32), _p(ot, ctypes.c_uint64), cap)
    return {(int(oa[k]), int(ob[k])): int(ot[k]) for k in range(m)}


def half_kick(state, dt):
    lib().or_half_kick(state.n, _p(state.vx), _p(state.vy), _p(state.vz), _p(state.ax), _p(state.ay), _p(state.az),
                       dt)


def drift(state, dt):
    lib().or_drift(state.n, _p(state.x), _p(state.y), _p(state.z), _p(state.vx), _p(state.vy), _p(state.vz), dt)


def wrap(state, box):
    b = _d(box)
    lib().or_wrap(state.n, _p(state.x), _p(state.y), _p(state.z), _p(b))


def kinetic(state, mass):
    m = _d(mass)
    sp = np.ascontiguousarray(state.species, dtype=np.uint8)
    return lib().or_kinetic(state.n, _p(state.vx), _p(state.vy), _p(state.vz), _p(sp, ctypes.c_uint8), _p(m))


def thermostat(physics, state):
    p = physics.to_c()
    sp = np.ascontiguousarray(state.species, dtype=np.uint8)
    return lib().or_thermostat(ctypes.byref(p), state.n, _p(state.vx), _p(state.vy), _p(state.vz),
                               _p(sp, ctypes.c_uint8))


def radial_distribution(x, y, z, size, r_max, nbins, species=None, ta=-1, tb=-1, full=False):
    """full: every atom (no subsample above 8192 atoms)."""
    n = len(x)
    x, y, z, s = _d(x), _d(y), _d(z), _d(size)
    sp = None if species is None else np.ascontiguousarray(species, dtype=np.uint8)
    nb = max(1, min(nbins, 1 << 16))
    counts, g = np.zeros(nb, dtype=np.uint64), np.zeros(nb)
    rm, dr = ctypes.c_double(), ctypes.c_double()
    lib().or_radial_distribution(n, _p(x), _p(y), _p(z), _p(sp, ctypes.c_uint8), _p(s), r_max, nbins, ta, tb,
                                 _p(counts, ctypes.c_uint64), _p(g), ctypes.byref(rm), ctypes.byref(dr),
                                 1 if full else 0)
    return counts, g, rm.value, dr.value


def clusters(x, y, z, species, sigma_table, size, bond_factor=1.5):
    n = len(x)
    x, y, z, s, st = _d(x), _d(y), _d(z), _d(size), _d(sigma_table)
    sp = np.ascontiguousarray(species, dtype=np.uint8)
    labels, stats = np.zeros(max(n, 1), dtype=np.uint32), np.zeros(5)
    lib().or_clusters(n, _p(x), _p(y), _p(z), _p(sp, ctypes.c_uint8), _p(st), _p(s), bond_factor,
                      _p(labels, ctypes.c_uint32), _p(stats))
    return labels[:n], {"monomers": int(stats[0]), "aggregates": int(stats[1]), "largest": int(stats[2]),
                        "mean_size": float(stats[3]), "bound_pairs": int(stats[4])}


def bond_components(entities, edges):
    """entities: list of (id, species); edges: list of (a, b)."""
    ne = len(entities)
    ids = np.array([e[0] for e in entities], dtype=np.uint32) if ne else np.zeros(1, dtype=np.uint32)
    sp = np.array([e[1] for e in entities], dtype=np.uint8) if ne else np.zeros(1, dtype=np.uint8)
    ea = np.array([e[0] for e in edges], dtype=np.uint32) if edges else np.zeros(1, dtype=np.uint32)
    eb = np.array([e[1] for e in edges], dtype=np.uint32) if edges else np.zeros(1, dtype=np.uint32)
    sizes = np.zeros(max(ne, 1), dtype=np.int64)
    members = np.zeros(max(ne, 1), dtype=np.uint32)
    stride = 64
    fbuf = ctypes.create_string_buffer(stride * max(ne, 1))
    nc = lib().or_bond_components(ne, _p(ids, ctypes.c_uint32), _p(sp, ctypes.c_uint8), len(edges),
                                  _p(ea, ctypes.c_uint32), _p(eb, ctypes.c_uint32), _p(sizes, ctypes.c_int64),
                                  _p(members, ctypes.c_uint32), fbuf, stride)
    out, off = [], 0
    for k in range(nc):
        sz = int(sizes[k])
        f = fbuf.raw[k * stride:(k + 1) * stride].split(b"\0", 1)[0].decode()
        out.append({"size": sz, "members": members[off:off + sz].tolist(), "formula": f})
        off += sz
    return out


# ---- adjoint reference (oracle/adjoint.cpp) ----------------------------------

PAIR_DERIV_NAMES = ("phi", "d1", "d2", "phi_s", "phi_e", "phi_q", "phi_rc", "f_s", "f_e", "f_q", "f_rc")


def pair_derivs(d, sigma, eps, qq, rc, coul=True):
    out = np.zeros(11)
    lib().or_pair_derivs(d, sigma, eps, qq, rc, 1 if coul else 0, _p(out))
    return dict(zip(PAIR_DERIV_NAMES, out.tolist()))


def theta_of(physics):
    return np.array(list(physics.sigma) + list(physics.epsilon_k) + list(physics.charge), dtype=np.float64)


def adj_accel(physics, state, keep, theta):
    n = state.n
    p = physics.to_c()
    ax, ay, az = np.zeros(n), np.zeros(n), np.zeros(n)
    sp = np.ascontiguousarray(state.species, dtype=np.uint8)
    kp = np.ascontiguousarray(keep, dtype=np.uint8)
    th = _d(theta)
    V = lib().or_adj_accel(ctypes.byref(p), n, _p(_d(state.x)), _p(_d(state.y)), _p(_d(state.z)),
                           _p(sp, ctypes.c_uint8), _p(kp, ctypes.c_uint8), _p(th), _p(ax), _p(ay), _p(az))
    return V, np.stack([ax, ay, az], 1)


def adj_pass(physics, state, keep, theta, lam, vcoef=0.0):
    n = state.n
    p = physics.to_c()
    sp = np.ascontiguousarray(state.species, dtype=np.uint8)
    kp = np.ascontiguousarray(keep, dtype=np.uint8)
    lx, ly, lz = (np.ascontiguousarray(lam[:, k]) for k in range(3))
    ox, oy, oz, g = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(30)
    lib().or_adj_pass(ctypes.byref(p), n, _p(_d(state.x)), _p(_d(state.y)), _p(_d(state.z)),
                      _p(sp, ctypes.c_uint8), _p(kp, ctypes.c_uint8), _p(_d(theta)), _p(lx), _p(ly), _p(lz), vcoef,
                      _p(ox), _p(oy), _p(oz), _p(g))
    return np.stack([ox, oy, oz], 1), g


def adjoint(physics, state, keep, theta, nsteps, win_from, gr=None, g0=None, grad=True):
    """Loss (mean V over ticks >= win_from + smoothed-g(r) deviation) and its
    theta gradient.  gr = (nbins, r_max, sigma_b)."""
    n = state.n
    p = physics.to_c()
    sp = np.ascontiguousarray(state.species, dtype=np.uint8)
    kp = np.ascontiguousarray(keep, dtype=np.uint8)
    nb, rmax, sb = gr if gr else (0, 0.0, 0.0)
    loss = ctypes.c_double()
    gout = np.zeros(30) if grad else None
    g_final = np.zeros(max(nb, 1))
    g0a = None if g0 is None else _d(g0)
    lib().or_adjoint(ctypes.byref(p), n, *[_p(_d(a)) for a in (state.x, state.y, state.z, state.vx, state.vy,
                                                                 state.vz)],
                     _p(sp, ctypes.c_uint8), _p(kp, ctypes.c_uint8), _p(_d(theta)), nsteps, win_from, nb, rmax, sb,
                     _p(g0a), ctypes.byref(loss), _p(gout), _p(g_final))
    return loss.value, gout, g_final[:nb]
