"""Host-side handle over the resident GPU state (one ``gs_universe``).

``GpuUniverse`` owns one C-ABI handle; its methods map one-to-one onto the
entry points of include/genesis_gpu.h, with numpy arrays in entity (slot)
order on the host side.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Tuple

import numpy as np

from . import _lib
from .inputs import State
from .params import Physics


def _dp(a: Optional[np.ndarray]):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


OBS_FIELDS = ("kinetic", "potential", "bond_energy", "temperature", "lambda", "energy_total")


class GpuUniverse:
    """Resident universe on the current HIP device."""

    def __init__(self, physics: Physics, state: State):
        self.lib = _lib.load()
        self.physics = physics
        self.n = state.n
        self._params = physics.to_c()
        h = ctypes.c_void_p()
        _lib.check(self.lib.gs_create(ctypes.byref(self._params), self.n, ctypes.byref(h)))
        self.h = h
        self.upload(state)

    # -- lifecycle ---------------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.gs_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- state transfer ------------------------------------------------------------
    def upload(self, s: State) -> None:
        ax, ay, az = (None, None, None) if s.ax is None else (_f64(s.ax), _f64(s.ay), _f64(s.az))
        arrs = [_f64(s.x), _f64(s.y), _f64(s.z), _f64(s.vx), _f64(s.vy), _f64(s.vz)]
        sp = np.ascontiguousarray(s.species, dtype=np.uint8)
        q = _f64(s.charge)
        self._keep = (arrs, ax, ay, az, sp, q)
        _lib.check(self.lib.gs_upload_state(self.h, *[_dp(a) for a in arrs], _dp(ax), _dp(ay), _dp(az),
                                            sp.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), _dp(q)))
        self.species = sp.copy()
        self.charge = q.copy()

    def download(self) -> State:
        out = [np.empty(self.n) for _ in range(9)]
        _lib.check(self.lib.gs_download_state(self.h, *[_dp(a) for a in out]))
        return State(out[0], out[1], out[2], out[3], out[4], out[5], self.species.copy(), self.charge.copy(),
                     out[6], out[7], out[8])

    # -- systems -------------------------------------------------------------------
    def forces(self) -> Tuple[float, float]:
        v, vb = ctypes.c_double(), ctypes.c_double()
        _lib.check(self.lib.gs_forces(self.h, ctypes.byref(v), ctypes.byref(vb)))
        return v.value, vb.value

    def force_extras(self) -> Tuple[np.ndarray, np.ndarray]:
        pe, sa = np.empty(self.n), np.empty(self.n)
        _lib.check(self.lib.gs_download_force_extras(self.h, _dp(pe), _dp(sa)))
        return pe, sa

    def cells(self) -> Tuple[np.ndarray, tuple]:
        c = np.empty(self.n, dtype=np.uint32)
        d = np.zeros(3, dtype=np.uint32)
        _lib.check(self.lib.gs_download_cells(self.h, c.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                              d.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))))
        return c, tuple(int(v) for v in d)

    def half_kick(self) -> None:
        _lib.check(self.lib.gs_half_kick(self.h))

    def drift(self) -> None:
        _lib.check(self.lib.gs_drift(self.h))

    def wrap(self) -> None:
        _lib.check(self.lib.gs_wrap(self.h))

    def thermostat(self) -> float:
        lam = ctypes.c_double()
        _lib.check(self.lib.gs_thermostat(self.h, ctypes.byref(lam)))
        return lam.value

    def kinetic(self) -> float:
        k = ctypes.c_double()
        _lib.check(self.lib.gs_kinetic(self.h, ctypes.byref(k)))
        return k.value

    def speed_stats(self) -> Tuple[float, float]:
        """(sum 0.5 m |v|^2, sum |v|) over all entities, reduced on the device."""
        k, sp = ctypes.c_double(), ctypes.c_double()
        _lib.check(self.lib.gs_speed_stats(self.h, ctypes.byref(k), ctypes.byref(sp)))
        return k.value, sp.value

    def step(self, n_steps: int, observe: bool = True) -> Optional[np.ndarray]:
        """Universe::tick x n_steps on the device; returns (n_steps, 6) observables."""
        if not observe:
            _lib.check(self.lib.gs_step(self.h, n_steps, None))
            return None
        buf = (_lib.Observables * max(1, n_steps))()
        _lib.check(self.lib.gs_step(self.h, n_steps, buf))
        return np.ctypeslib.as_array(ctypes.cast(buf, ctypes.POINTER(ctypes.c_double)),
                                     shape=(max(1, n_steps), 6))[:n_steps].copy()

    def tuning(self) -> dict:
        t = _lib.Tuning()
        _lib.check(self.lib.gs_get_tuning(self.h, ctypes.byref(t)))
        return {f: getattr(t, f) for f, _ in _lib.Tuning._fields_}

    def set_tuning(self, pair_mode: int, lanes_per_atom: int = 4, skin: Optional[float] = None,
                   sort_every: int = 0) -> None:
        """pair_mode 0 = per-tick cell tiles, 1 = Verlet list (lanes_per_atom, skin, and
        sort_every = ticks between full re-sorts of the resident state; 0 keeps it),
        2 = the same lists evaluated cell by cell from LDS (lanes_per_atom unused)."""
        cur = self.tuning()
        t = _lib.Tuning(pair_mode, lanes_per_atom, cur["skin"] if skin is None else skin, sort_every, 0, 0, 0.0)
        _lib.check(self.lib.gs_set_tuning(self.h, ctypes.byref(t)))

    def set_bound_search(self, aux_grid_only: bool) -> None:
        """Where gs_step finds bond observation's bound pairs: False (default)
        lets the Verlet-list pass do it when it can, True always uses the aux
        grid; the results are identical."""
        _lib.check(self.lib.gs_set_bound_search(self.h, 1 if aux_grid_only else 0))

    def autotune(self, trial_steps: int = 40) -> dict:
        t = _lib.Tuning()
        _lib.check(self.lib.gs_autotune(self.h, trial_steps, ctypes.byref(t)))
        return {f: getattr(t, f) for f, _ in _lib.Tuning._fields_}

    def timing(self, enable=None, every: int = 1) -> dict:
        """Event timing of step(): enable=True or "phases" times every phase (four
        events per timed tick), "pair" only the pair kernel (two), False stops;
        every k-th tick of each step() call is timed.  Enabling resets the counters."""
        if enable is not None:
            mode = {False: 0, True: 2, "phases": 2, "pair": 1}[enable]
            _lib.check(self.lib.gs_timing_enable(self.h, mode, every))
        t = _lib.Timing()
        _lib.check(self.lib.gs_timing_read(self.h, ctypes.byref(t)))
        return {f: getattr(t, f) for f, _ in _lib.Timing._fields_}

    # -- differentiable trajectory ---------------------------------------------------
    def hvp(self, theta, keep, lam, vcoef: float = 0.0):
        """(da/dx)^T lam - vcoef dV/dx and (da/dtheta)^T lam + vcoef dV/dtheta."""
        th = _f64(theta)
        kp = np.ascontiguousarray(keep, dtype=np.uint8)
        lx, ly, lz = (_f64(lam[:, k]) for k in range(3))
        ox, oy, oz, g = np.empty(self.n), np.empty(self.n), np.empty(self.n), np.empty(30)
        _lib.check(self.lib.gs_hvp(self.h, _dp(th), kp.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), _dp(lx),
                                   _dp(ly), _dp(lz), vcoef, _dp(ox), _dp(oy), _dp(oz), _dp(g)))
        return np.stack([ox, oy, oz], 1), g

    def adjoint_forward(self, theta, keep, n_steps: int, ckpt_every: int, win_from: int, gr=None, g0=None):
        """Loss of an n_steps trajectory (mean V over ticks >= win_from + smoothed
        g(r) deviation); gr = (nbins, r_max, sigma_b).  Returns (loss, g(x_T))."""
        th = _f64(theta)
        kp = np.ascontiguousarray(keep, dtype=np.uint8)
        self._adj_keep = (th, kp)
        spec = None
        nb = 0
        if gr is not None:
            nb = int(gr[0])
            spec = _lib.GrSpec(nb, 0, float(gr[1]), float(gr[2]))
        g0a = None if g0 is None else _f64(g0)
        gout = np.zeros(max(nb, 1))
        loss = ctypes.c_double()
        _lib.check(self.lib.gs_adjoint_forward(self.h, _dp(th), kp.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                               n_steps, ckpt_every, win_from,
                                               None if spec is None else ctypes.byref(spec), _dp(g0a),
                                               ctypes.byref(loss), _dp(gout) if nb else None))
        return loss.value, gout[:nb]

    def adjoint_backward(self) -> np.ndarray:
        g = np.empty(30)
        _lib.check(self.lib.gs_adjoint_backward(self.h, _dp(g)))
        return g

    # -- lenses --------------------------------------------------------------------
    def velocity_histogram(self, max_speed: float, nbins: int) -> dict:
        """velocity_histogram (src/stats/mod.rs:255-297) of the resident velocities:
        {max_speed, bin_width, bins (u64), samples, overflow}."""
        nb = min(max(int(nbins), 1), 512)
        counts = np.zeros(nb, dtype=np.uint64)
        samples, overflow, bw = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_double()
        _lib.check(self.lib.gs_velocity_histogram(self.h, float(max_speed), nb,
                                                  counts.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                                  ctypes.byref(samples), ctypes.byref(overflow), ctypes.byref(bw)))
        return {"max_speed": max(float(max_speed), 1e-9), "bin_width": bw.value, "bins": counts,
                "samples": samples.value, "overflow": overflow.value}

    def radial_distribution(self, r_max: float, nbins: int):
        counts = np.zeros(max(1, min(nbins, 1 << 16)), dtype=np.uint64)
        g = np.zeros(counts.shape[0])
        rm, dr = ctypes.c_double(), ctypes.c_double()
        _lib.check(self.lib.gs_radial_distribution(self.h, r_max, nbins,
                                                   counts.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), _dp(g),
                                                   ctypes.byref(rm), ctypes.byref(dr)))
        return counts, g, rm.value, dr.value

    def radial_distribution_between(self, ta: int, tb: int, r_max: float, nbins: int):
        counts = np.zeros(max(1, min(nbins, 1 << 16)), dtype=np.uint64)
        g = np.zeros(counts.shape[0])
        rm, dr = ctypes.c_double(), ctypes.c_double()
        _lib.check(self.lib.gs_radial_distribution_between(
            self.h, ta, tb, r_max, nbins, counts.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), _dp(g),
            ctypes.byref(rm), ctypes.byref(dr)))
        return counts, g, rm.value, dr.value

    def radial_distribution_full(self, r_max: float, nbins: int, ta: int = -1, tb: int = -1):
        """g(r) (ta < 0) or g_ab(r) over every atom, without the subsample."""
        counts = np.zeros(max(1, min(nbins, 1 << 16)), dtype=np.uint64)
        g = np.zeros(counts.shape[0])
        rm, dr = ctypes.c_double(), ctypes.c_double()
        _lib.check(self.lib.gs_radial_distribution_full(
            self.h, ta, tb, r_max, nbins, counts.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), _dp(g),
            ctypes.byref(rm), ctypes.byref(dr)))
        return counts, g, rm.value, dr.value

    def clusters(self, bond_factor: float = 1.5):
        labels = np.zeros(self.n, dtype=np.uint32)
        st = _lib.ClusterStats()
        _lib.check(self.lib.gs_clusters(self.h, bond_factor, labels.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                        ctypes.byref(st)))
        stats = {"monomers": st.monomers, "aggregates": st.aggregates, "largest": st.largest,
                 "mean_size": st.mean_size, "bound_pairs": st.bound_pairs}
        return labels, stats

    # -- observed bonds ---------------------------------------------------------
    def bond_observe(self, tick: int) -> None:
        """BondObservationSystem::run on the resident positions as tick ``tick``."""
        _lib.check(self.lib.gs_bond_observe(self.h, int(tick)))

    def bond_stats(self) -> dict:
        """The BondObservation summary of the last observed tick."""
        b = _lib.BondStats()
        _lib.check(self.lib.gs_bond_stats_read(self.h, ctypes.byref(b)))
        return {"tick": b.tick, "bonded_pairs": b.bonded_pairs, "bonded_entities": b.bonded_entities,
                "mean_coordination": b.mean_coordination, "bonds_formed": b.bonds_formed,
                "lifetime_sum_ticks": b.lifetime_sum_ticks,
                "species_matrix": np.array(b.species_matrix[:], dtype=np.int64).reshape(10, 10)}

    def bonds(self):
        """The Bonds adjacency as CSR (ptr[n+1], idx) over entity ids, neighbours ascending."""
        ptr = np.zeros(self.n + 1, dtype=np.int32)
        nnz = ctypes.c_int64()
        i32 = ctypes.POINTER(ctypes.c_int32)
        cap = 32 * max(self.n, 1)
        idx = np.zeros(cap, dtype=np.int32)
        _lib.check(self.lib.gs_download_bonds(self.h, ptr.ctypes.data_as(i32), idx.ctypes.data_as(i32), cap,
                                              ctypes.byref(nnz)))
        return ptr, idx[:nnz.value].copy()

    def upload_bonds(self, ptr, idx) -> None:
        ptr = np.ascontiguousarray(ptr, dtype=np.int32)
        idx = np.ascontiguousarray(idx if len(idx) else np.zeros(1), dtype=np.int32)
        i32 = ctypes.POINTER(ctypes.c_int32)
        _lib.check(self.lib.gs_upload_bonds(self.h, ptr.ctypes.data_as(i32), idx.ctypes.data_as(i32)))

    def bound_pairs(self) -> np.ndarray:
        """bound_pairs_with_grid on the resident positions: (k, 2) entity-id pairs, a < b, ascending."""
        cnt = ctypes.c_int64()
        u32 = ctypes.POINTER(ctypes.c_uint32)
        _lib.check(self.lib.gs_bound_pairs(self.h, None, None, 0, ctypes.byref(cnt)))
        k = cnt.value
        a = np.zeros(max(k, 1), dtype=np.uint32)
        b = np.zeros(max(k, 1), dtype=np.uint32)
        _lib.check(self.lib.gs_bound_pairs(self.h, a.ctypes.data_as(u32), b.ctypes.data_as(u32), k,
                                           ctypes.byref(cnt)))
        return np.stack([a[:k], b[:k]], 1)

    def bond_labels(self) -> np.ndarray:
        """Component label per entity (smallest id of its Bonds component; 0xFFFFFFFF when unbonded)."""
        labels = np.zeros(max(self.n, 1), dtype=np.uint32)
        nc = ctypes.c_int64()
        _lib.check(self.lib.gs_bond_components(self.h, labels.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                               ctypes.byref(nc)))
        return labels[:self.n]

    def bond_binding(self) -> np.ndarray:
        """Observed binding energy per entity (gs_bond_binding): the raw pair potential
        summed over the entity's bonded partners with a larger id, on the resident
        positions; a component's binding energy is the sum over its members."""
        out = np.zeros(max(self.n, 1))
        _lib.check(self.lib.gs_bond_binding(self.h, _dp(out)))
        return out[:self.n]

    def bond_components(self, species) -> list:
        """bond_components (src/analysis/mod.rs:337-412) of the Bonds graph: one dict per
        component (size, composition, formula, members ascending, edges) ordered by size
        descending, then formula, then smallest member (the reference leaves that last tie
        to its hash order)."""
        labels = self.bond_labels()
        ptr, idx = self.bonds()
        return components_from_labels(labels, np.asarray(species), ptr, idx)

    def bond_reset(self) -> None:
        _lib.check(self.lib.gs_bond_reset(self.h))

    def set_tick(self, tick: int) -> None:
        """Time::tick of the universe; the next step is tick + 1."""
        _lib.check(self.lib.gs_set_tick(self.h, int(tick)))

    def tick(self) -> int:
        t = ctypes.c_int64()
        _lib.check(self.lib.gs_get_tick(self.h, ctypes.byref(t)))
        return t.value

    def save(self, path: str, config=None, state=None) -> None:
        """save_universe: a bincode v4 Snapshot followed by the device image."""
        from .snapshot import save_universe
        save_universe(self, path, config, state)


def formula(composition) -> str:
    """Stoichiometric formula (src/analysis/mod.rs:417-435): symbols by descending
    count, then alphabetically, '-' separated, the count shown when > 1."""
    from .params import SYMBOLS
    parts = [(int(c), SYMBOLS[i]) for i, c in enumerate(composition) if c > 0]
    parts.sort(key=lambda t: (-t[0], t[1]))
    return "-".join(f"{sym}{c}" if c > 1 else sym for c, sym in parts)


def components_from_labels(labels, species, ptr, idx) -> list:
    """ChemicalComponent records from per-entity component labels and the Bonds CSR."""
    no = np.uint32(0xFFFFFFFF)
    comps = {}
    for e in np.nonzero(labels != no)[0]:
        comps.setdefault(int(labels[e]), []).append(int(e))
    out = []
    for root, members in comps.items():
        comp = np.zeros(10, dtype=np.int64)
        edges = []
        for e in members:
            comp[int(species[e])] += 1
            for m in idx[ptr[e]:ptr[e + 1]]:
                if int(m) > e:
                    edges.append((e, int(m)))
        out.append({"size": len(members), "composition": comp.tolist(), "formula": formula(comp),
                    "members": sorted(members), "edges": sorted(edges)})
    out.sort(key=lambda c: (-c["size"], c["formula"], c["members"][0]))
    return out

