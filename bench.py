#!/usr/bin/env python3
"""Benchmark of the MI355X force-and-integration path.

Headline workload: the 1,048,576-atom Na/O ionic melt with bonds (rock salt
on a 128 x 128 x 64 lattice at rho sigma_NaO^3 = 0.7, FP64, LJ + truncated
Coulomb at rc = 2.5 max sigma, Maxwell-Boltzmann 300 K quenched by Berendsen
to 80 K with tau = 100 fs, dt = 0.5 fs, bond observation every tick with the
harmonic bond interaction on, g(r) with 256 bins to L_min/4 every 100 ticks,
the Bonds graph's components every 500 ticks).  The quench runs untimed until
persistent bonds exist (the persistence threshold is about 11,600 ticks), so
the timed window is in the bonded regime; it ends on a tick that is a
multiple of 500, so it always holds its g(r) and bond-component samples.  A
step is one full tick of the reference schedule (kick, drift, wrap, list
maintenance, forces with the bond term, kick, Berendsen, bond observation)
over the resident state; inputs are synthetic (seeded) and resident in HBM
before the timed region; gs_step returns only after its stream has drained.

Prints one JSON line.  ``roofline`` is the dominant kernel's FP64 roofline
(its binding resource: the FP64 vector pipe; the MI355X FP64 vector peak is
78.6 TFLOP/s), ``roofline_hbm`` the same kernel against HBM,
``cpu_baseline`` the oracle's restatement of the reference CPU algorithm
timed on this host on a one-tick sample of the same workload, with all the
host cores this job may use and with one thread.  Extra legs: the 32,768-atom
mixture, the 64-replica temperature sweep, the 8,192-atom adjoint.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "atom-steps/s (force+Verlet), batched vs plain, at stated force tol; % HBM BW"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
# Algorithmic work per atom-step (compulsory HBM bytes of one tick, and the pair law's flops):
STEP_BYTES = 161.0          # x, v, a read+write (144) + species (1) + cell key (8) + charge (8)
# list pair kernel, per atom: x, y, z, q, species read (33) + a written (24) + v read and written
# by the fused second kick (48)
PAIR_BYTES = 105.0
FLOP_PER_PAIR = 82.0        # LJ + quintic switch + accumulate (70) + Coulomb (12), per ordered pair
CPU_SHARE = 16              # host cores one GPU's job may use on the GPU box
# HBM traffic per launch of the dominant kernel, from rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE (KiB per dispatch) of `bench.py --leg ionic` at the default setting, scaled by
# what the same counters read for a 256 MiB device copy (`bench.py --copy-probe`).  Only
# reported for the setting it was measured at.
PAIR_TRAFFIC = {"setting": {"lanes": 2, "skin_frac": 0.035}, "fetch_kib": 376322.7, "write_kib": 112371.0,
                "fetch_scale": 2.0, "write_scale": 1.0,
                "source": "profiles/k_pair_list_bound__ionic_melt1048576_pmc.md (ticks 12001..12500)"}
# the melt is a solid in the timed window: its resident order keeps its locality, so a full
# re-sort every 300 ticks (1.256e9 atom-steps/s against 1.241e9 at 100 and 1.255e9 at 1000)
# skin 0.035 rc: as fast per tick as 0.04 (more rebuilds, fewer list entries; 0.03 is slower)
DEFAULT_LANES, DEFAULT_SKIN_FRAC, DEFAULT_SORT_EVERY = 2, 0.035, 300
GR_EVERY, COMP_EVERY = 100, 500


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


class Sampler:
    """Steps a universe in chunks that end on the sampling ticks: g(r) on every
    tick that is a multiple of gr_every, the Bonds graph's components on every
    multiple of comp_every (labels by entity id, copied to the host)."""

    def __init__(self, u, tick, gr_rmax, gr_bins, gr_every=GR_EVERY, comp_every=COMP_EVERY):
        self.u, self.tick = u, tick
        self.gr_rmax, self.gr_bins = gr_rmax, gr_bins
        self.gr_every, self.comp_every = gr_every, comp_every
        self.n_gr = self.n_comp = 0

    def run(self, k):
        done = 0
        while done < k:
            chunk = min(self.gr_every - self.tick % self.gr_every, k - done)
            self.u.step(chunk, observe=False)
            self.tick += chunk
            done += chunk
            if self.tick % self.gr_every == 0:
                self.u.radial_distribution(self.gr_rmax, self.gr_bins)
                self.n_gr += 1
            if self.comp_every and self.tick % self.comp_every == 0:
                self.u.bond_labels()
                self.n_comp += 1


def count_pairs(s, ph) -> int:
    """Ordered pairs within the force cutoff of a state (periodic k-d tree count)."""
    from scipy.spatial import cKDTree
    box = np.asarray(ph.box, dtype=np.float64)
    p = np.mod(np.stack([s.x, s.y, s.z], 1) + 0.5 * box, box)
    p[p >= box] = 0.0
    t = cKDTree(p, boxsize=box)
    c = t.query_ball_point(p, ph.force_rc(), return_length=True, workers=min(CPU_SHARE, os.cpu_count() or 1))
    return int(c.sum()) - s.n


def pair_roofline(kernel, pair_ms, npairs, n, traffic=None):
    tflops = FLOP_PER_PAIR * npairs / (pair_ms * 1e-3) / 1e12
    gbs = PAIR_BYTES * n / (pair_ms * 1e-3) / 1e9
    fp64 = {"bound": "fp64-valu", "pipe": "FP64 vector ALU (pair forces are not a dense contraction: no MFMA)",
            "kernel": kernel, "achieved": tflops, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": tflops / FP64_PEAK_TFLOPS, "traffic": traffic,
            "per_launch": {"ordered_pairs_in_range": npairs, "flop_per_pair": FLOP_PER_PAIR, "avg_ms": pair_ms,
                           "algorithmic_flop": FLOP_PER_PAIR * npairs}}
    hbm = {"bound": "hbm", "kernel": kernel, "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": gbs / HBM_PEAK_GBS, "bytes_per_atom": PAIR_BYTES, "traffic": traffic}
    return fp64, hbm


def ionic_headline(steps: int, warmup: int, lanes: int, skin_frac: float, setup_max: int = 20000,
                   aux_bound_search: bool = False, sort_every: int = DEFAULT_SORT_EVERY):
    from genesis_amd.inputs import ionic_melt
    from genesis_amd.universe import GpuUniverse
    w = ionic_melt()
    ph, st = w.physics, w.state
    n = st.n
    skin = skin_frac * ph.force_rc()
    log(f"workload {w.name}: n={n} box={tuple(round(b, 3) for b in ph.box)} dims={ph.grid_dims()} "
        f"steps={steps} warmup={warmup} lanes={lanes} skin={skin:.4f}")
    u = GpuUniverse(ph, st)
    u.set_tuning(1, lanes, skin=skin, sort_every=sort_every)
    u.set_bound_search(aux_bound_search)
    # Setup (untimed): the quench until persistent bonds exist.
    t0 = time.perf_counter()
    smp = Sampler(u, 0, w.extra["gr_rmax"], w.extra["gr_bins"])
    bonded = 0
    while smp.tick < setup_max:
        smp.run(500)
        bonded = u.bond_stats()["bonded_pairs"]
        if bonded > 0:
            break
    if bonded == 0:
        raise RuntimeError(f"no persistent bonds after {smp.tick} ticks")
    # the window [first, last] ends on a multiple of 500
    smp.run((-(smp.tick + warmup + steps)) % COMP_EVERY)
    setup_ticks, setup_s = smp.tick, time.perf_counter() - t0
    log(f"setup: {setup_ticks} ticks in {setup_s:.1f} s, {bonded} bonded pairs at the first bonded sample")
    smp.run(warmup)
    b0 = u.bond_stats()
    tun0 = u.tuning()
    gr0, comp0 = smp.n_gr, smp.n_comp
    first = smp.tick + 1
    u.timing(enable="pair", every=10)
    # the window is bracketed by two gs_speed_stats calls (their k_speed_parts
    # launches mark its ends in a kernel trace of this run)
    k_first = u.speed_stats()[0]
    t0 = time.perf_counter()
    smp.run(steps)
    t1 = time.perf_counter()
    k_last = u.speed_stats()[0]
    tm = u.timing()
    u.timing(enable=False)
    elapsed = t1 - t0
    tun = u.tuning()
    b1 = u.bond_stats()
    value = n * steps / elapsed
    pair_ms = tm["pair_ms"] / max(tm["steps"], 1)
    log(f"{value:.4e} atom-steps/s, {1e3 * elapsed / steps:.4f} ms/tick over ticks {first}..{smp.tick}; "
        f"pair kernel {pair_ms:.4f} ms; bonded pairs {b0['bonded_pairs']} -> {b1['bonded_pairs']}; "
        f"g(r) samples {smp.n_gr - gr0}, component samples {smp.n_comp - comp0}")
    # untimed: phases, fixed per-call costs, pair count of the final state
    u.timing(enable="phases")
    t0 = time.perf_counter()
    u.step(200, observe=False)
    t_tick = (time.perf_counter() - t0) / 200
    pt = u.timing()
    u.timing(enable=False)
    phases = {k: pt[k] / max(pt["steps"], 1) for k in ("rebuild_ms", "pair_ms", "kick_ms")}
    phases["observe_and_rest_ms"] = 1e3 * t_tick - sum(phases.values())
    fixed = fixed_costs(u, w.extra["gr_rmax"], w.extra["gr_bins"], 1e3 * t_tick)
    fin = u.download()
    bfin = u.bond_stats()
    u.close()
    npairs = count_pairs(fin, ph)
    kernel = (f"k_pair_list<{lanes} lanes{'' if aux_bound_search else ', bound-pair search'}> "
              f"(pair forces + bond term + second kick + Berendsen)")
    traffic = None
    if PAIR_TRAFFIC["fetch_kib"] is not None and (lanes, skin_frac) == (PAIR_TRAFFIC["setting"]["lanes"],
                                                                       PAIR_TRAFFIC["setting"]["skin_frac"]):
        traffic = 1024.0 * (PAIR_TRAFFIC["fetch_kib"] * PAIR_TRAFFIC["fetch_scale"] +
                            PAIR_TRAFFIC["write_kib"] * PAIR_TRAFFIC["write_scale"])
    roof, roof_hbm = pair_roofline(kernel, pair_ms, npairs, n, traffic)
    deg = 2.0 * bfin["bonded_pairs"] / n
    step_bytes = STEP_BYTES + 4.0 * (1.0 + deg)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": 1e3 * elapsed / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded rock-salt lattice + Maxwell-Boltzmann, no dataset)",
        "config": {"workload": w.name, "atoms": n, "lattice": "128x128x64 rock salt, Na +1 / O -1",
                   "box": list(ph.box), "rc": ph.force_rc(), "cells": list(ph.grid_dims()), "dt_fs": 0.5,
                   "thermostat": "berendsen, 300 K start -> 80 K, tau 200 ticks", "coulomb": True,
                   "bonds": {"k_bind": ph.bond_k_bind, "min_periods": ph.bond_min_periods, "debounce_ticks": 2,
                             "observation": "every tick", "interaction": True},
                   "gr": {"bins": w.extra["gr_bins"], "r_max": w.extra["gr_rmax"], "every": GR_EVERY},
                   "bond_components_every": COMP_EVERY,
                   "window": {"first_tick": first, "last_tick": smp.tick, "untimed_setup_ticks": setup_ticks,
                              "bonded_pairs_at_start": b0["bonded_pairs"], "bonded_pairs_at_end": b1["bonded_pairs"],
                              "bonded_entities_at_end": b1["bonded_entities"],
                              "gr_samples_timed": smp.n_gr - gr0, "component_samples_timed": smp.n_comp - comp0,
                              "kinetic_at_ends": [k_first, k_last],
                              "marked_by": "one gs_speed_stats call (k_speed_parts) at each end"},
                   "parallelism": "single GPU"},
        "roofline": roof,
        "roofline_hbm": roof_hbm,
        "step_roofline": {"hbm_frac": value * step_bytes / (HBM_PEAK_GBS * 1e9),
                          "fp64_frac": value * (FLOP_PER_PAIR * npairs / n) / (FP64_PEAK_TFLOPS * 1e12),
                          "bytes_per_atom_step": step_bytes, "flop_per_atom_step": FLOP_PER_PAIR * npairs / n},
        "phases_ms_per_step": phases,
        "fixed_costs_ms": fixed,
        "tuning": {"pair_mode": tun["pair_mode"], "lanes_per_atom": tun["lanes_per_atom"], "skin": tun["skin"],
                   "sort_every": tun["sort_every"]},
        "list_builds_in_timed_region": tun["list_builds"] - tun0["list_builds"],
        "list_replays_in_timed_region": tun["list_replays"] - tun0["list_replays"],
    }
    return out, w, fin


def fixed_costs(u, gr_rmax, gr_bins, tick_ms) -> dict:
    """Per-call costs outside the ticks: a g(r) sample, a bond-components
    sample, and the fixed cost of a gs_step call (snapshot, flag reads,
    synchronisation) from 20 one-tick calls against one 20-tick call."""
    t0 = time.perf_counter()
    for _ in range(3):
        u.radial_distribution(gr_rmax, gr_bins)
    gr = (time.perf_counter() - t0) / 3
    t0 = time.perf_counter()
    for _ in range(3):
        u.bond_labels()
    comp = (time.perf_counter() - t0) / 3
    t0 = time.perf_counter()
    u.step(20, observe=False)
    t20 = time.perf_counter() - t0
    t0 = time.perf_counter()
    for _ in range(20):
        u.step(1, observe=False)
    t1 = time.perf_counter() - t0
    return {"gr_sample": 1e3 * gr, "bond_components_sample": 1e3 * comp,
            "step_call_overhead": 1e3 * (t1 - t20) / 19, "tick": tick_ms}


def mixture_leg(steps: int = 2000, warmup: int = 100) -> dict:
    """The 32,768-atom 10-element mixture (Coulomb, Berendsen 60 K, dt 0.5 fs,
    g(r) every 100 ticks) at the list setting of 4 lanes per atom, skin 0.10 rc,
    with the autotuner's pick beside it."""
    from genesis_amd.inputs import mixture
    from genesis_amd.universe import GpuUniverse
    w = mixture()
    ph, st = w.physics, w.state
    n = st.n
    with GpuUniverse(ph, st) as u:
        u.set_tuning(1, 4, skin=0.10 * ph.force_rc())
        smp = Sampler(u, 0, w.extra["gr_rmax"], w.extra["gr_bins"], comp_every=0)
        smp.run(warmup)
        u.timing(enable="pair", every=10)
        u.speed_stats()  # window marker (k_speed_parts) for kernel traces and counters
        t0 = time.perf_counter()
        smp.run(steps)
        el = time.perf_counter() - t0
        u.speed_stats()
        tm = u.timing()
        u.timing(enable=False)
        fin = u.download()
    with GpuUniverse(ph, st) as au:
        t = au.autotune(40)
    pair_ms = tm["pair_ms"] / max(tm["steps"], 1)
    npairs = count_pairs(fin, ph)
    roof, roof_hbm = pair_roofline("k_pair_list<4 lanes>", pair_ms, npairs, n)
    v = n * steps / el
    log(f"mixture: {v:.4e} atom-steps/s ({1e3 * el / steps:.4f} ms/tick), pair kernel {pair_ms:.4f} ms, "
        f"{roof['achieved']:.2f} TFLOP/s; autotune pick {t}")
    return {"workload": w.name, "atoms": n, "ticks": steps, "atom_steps_per_s": v, "ms_per_tick": 1e3 * el / steps,
            "roofline": roof, "roofline_hbm": roof_hbm,
            "autotune": {k: t[k] for k in ("pair_mode", "lanes_per_atom", "skin", "ms_per_tick")}}


def batched_leg(ticks: int = 200, segment: int = 100, lanes: int = 0) -> dict:
    """Temperature sweep: 64 replicas of log-uniform sizes in [512, 16384]
    (20 K .. 120 K), stepped (a) as one batch (one launch per phase, hipGraph
    segments of up to 100 ticks, autotuned pair path: Verlet lists or cell
    tiles) and (b) one universe after another through gs_step with its own
    default setting (the per-replica-launch plain path).  lanes > 0 fixes the
    batch on Verlet lists with that many lanes per atom instead of autotuning
    (for profiles of one setting)."""
    from genesis_amd.batch import GpuBatch
    from genesis_amd.inputs import replica_sweep
    from genesis_amd.universe import GpuUniverse
    ws = replica_sweep()
    total = sum(w.state.n for w in ws)
    # window markers for kernel traces and counters: k_speed_parts of a
    # small universe before and after the timed batch steps
    marker = GpuUniverse(ws[0].physics, ws[0].state)
    with GpuBatch([w.physics for w in ws], [w.state for w in ws]) as b:
        if lanes > 0:
            b.set_tuning(1, lanes)
            tune = {"cell_scale": 1.0}
        else:
            tune = b.autotune(20)
        b.step(segment, graph_segment=segment)
        builds0 = b.get_tuning()["list_builds"]
        marker.speed_stats()
        t0 = time.perf_counter()
        b.step(ticks, graph_segment=segment)
        tb = time.perf_counter() - t0
        marker.speed_stats()
        btun = b.get_tuning()
    marker.close()
    us = [GpuUniverse(w.physics, w.state) for w in ws]
    for u in us:
        u.step(20, observe=False)
    t0 = time.perf_counter()
    for u in us:
        u.step(ticks, observe=False)
    tp = time.perf_counter() - t0
    for u in us:
        u.close()
    vb, vp = total * ticks / tb, total * ticks / tp
    log(f"batched sweep: {vb:.4e} atom-steps/s vs per-replica {vp:.4e} ({vb / vp:.2f}x), {total} atoms")
    return {"workload": f"temperature sweep, 64 replicas, {total} atoms, sizes log-uniform in [512, 16384]",
            "batched_atom_steps_per_s": vb, "per_replica_atom_steps_per_s": vp, "speedup": vb / vp,
            "ticks": ticks, "graph_segment": segment,
            "tuning": {"pair_mode": btun["pair_mode"], "lanes_per_atom": btun["lanes_per_atom"],
                       "skin": btun["skin"], "cell_scale": tune["cell_scale"],
                       "ticks_with_list_builds": btun["list_builds"] - builds0,
                       "list_replays": btun["list_replays"]}}


def adjoint_leg(steps: int = 1000, ckpt: int = 50) -> dict:
    """Adjoint workload: 8,192 atoms of the mixture, NVE, dt = 0.5 fs, 1000
    ticks, checkpoints every 50; loss = mean V over the last 200 ticks +
    smoothed g(r) deviation; gradient w.r.t. the 30 species parameters.
    Reports the reverse pass per tick against the adjoint's own forward tick
    and against a plain gs_step tick (Verlet-list path) of the same state."""
    from genesis_amd.adjoint import theta_of
    from genesis_amd.inputs import adjoint_mixture
    from genesis_amd.universe import GpuUniverse
    w = adjoint_mixture()
    ph, st = w.physics, w.state
    keep = ((np.abs(st.charge) > 0) | (np.asarray(ph.charge)[st.species] == 0)).astype(np.uint8)
    th0 = theta_of(ph).numpy()
    gr = (200, ph.box[0] / 4.0, 0.05)
    with GpuUniverse(ph, st) as u:
        win = max(1, steps - 199)  # the last 200 ticks
        _, g0 = u.adjoint_forward(th0, keep, steps, ckpt, win, gr=gr)
        th = th0 * (1.0 + 0.005 * np.cos(np.arange(30)))
        u.upload(st)
        u.speed_stats()  # window marker (k_speed_parts) for kernel traces and counters
        t0 = time.perf_counter()
        loss, _ = u.adjoint_forward(th, keep, steps, ckpt, win, gr=gr, g0=g0)
        tf = time.perf_counter() - t0
        t0 = time.perf_counter()
        grad = u.adjoint_backward()
        tb = time.perf_counter() - t0
        u.speed_stats()
    with GpuUniverse(ph, st) as u:
        u.step(50, observe=False)
        t0 = time.perf_counter()
        u.step(steps, observe=False)
        tplain = time.perf_counter() - t0
        tun = u.tuning()
    log(f"adjoint: forward {1e3 * tf / steps:.4f} ms/tick, reverse {1e3 * tb / steps:.4f} ms/tick "
        f"({tb / tf:.2f}x); plain gs_step {1e3 * tplain / steps:.4f} ms/tick (reverse {tb / tplain:.2f}x)")
    return {"workload": f"adjoint mixture, {st.n} atoms, {steps} ticks, checkpoint every {ckpt}",
            "forward_ms_per_tick": 1e3 * tf / steps, "reverse_ms_per_tick": 1e3 * tb / steps,
            "reverse_over_forward": tb / tf, "plain_step_ms_per_tick": 1e3 * tplain / steps,
            "reverse_over_plain_step": tb / tplain,
            "plain_step_tuning": {k: tun[k] for k in ("pair_mode", "lanes_per_atom", "skin")},
            "loss": loss, "grad_norm": float(np.linalg.norm(grad))}


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(w, state) -> dict:
    """The oracle's restatement of the reference CPU algorithm (serial grid +
    half-shell pair list, threaded fold with per-thread buffers merged in
    order, serial kinetic sum, the bond tracker) timed on this host on a
    bounded sample of the headline workload: one full tick with bond
    observation from the state the GPU reached, with all the cores this job
    may use and with one thread.  Beside it, the force pass alone on the
    32,768-atom mixture at 1, 8 and all cores: the reference's published claim
    is '~3x on 8 cores' for the parallel force pass (README.md:55-56)."""
    from genesis_amd.inputs import mixture
    from oracle import oracle as O
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    cores = max(1, min(avail, CPU_SHARE))
    legs = {}
    for th in (cores, 1):
        t0 = time.perf_counter()
        O.run_observed(w.physics, state, 1, nthreads=th)
        legs[th] = time.perf_counter() - t0
        log(f"cpu baseline: one tick of {state.n} atoms with {th} threads: {legs[th]:.2f} s")
    m = mixture()
    fp = {}
    for th in sorted({1, min(8, cores), cores}):
        O.forces(m.physics, m.state, nthreads=th, extras=False)
        t0 = time.perf_counter()
        for _ in range(3):
            O.forces(m.physics, m.state, nthreads=th, extras=False)
        fp[th] = (time.perf_counter() - t0) / 3
    log(f"cpu force pass, 32768 atoms: " + ", ".join(f"{k} threads {v * 1e3:.1f} ms" for k, v in fp.items()))
    return {"value": state.n / legs[cores], "unit": "atom-steps/s", "cores": cores, "kind": "port",
            "sample": f"1 full tick (forces, bond observation, Berendsen) of the {state.n}-atom ionic melt from "
                      f"the GPU's bonded state after the timed window (tracker episodes start empty), "
                      f"C++ restatement of the reference CPU algorithm, {legs[cores]:.1f} s with {cores} threads",
            "one_thread": {"value": state.n / legs[1], "seconds": legs[1]},
            "speedup_all_over_one": legs[1] / legs[cores],
            "cpu_model": cpu_model(), "host_cpus_visible": os.cpu_count(),
            "force_pass_mixture32768_ms": {str(k): 1e3 * v for k, v in fp.items()},
            "force_pass_speedup_8_over_1": fp[1] / fp[min(8, cores)],
            "reference_claim": "~3x parallel force pass on 8 cores (README.md:55-56)"}


def copy_probe() -> int:
    """A plain device copy of known size, for scaling rocprofv3's FETCH_SIZE /
    WRITE_SIZE counters (read 256 MiB, write 256 MiB per launch)."""
    import torch
    n = 256 * 1024 * 1024 // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": 8 * n, "copy_probe_bytes_written": 8 * n, "launches": 5}))
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--lanes", type=int, default=DEFAULT_LANES)
    ap.add_argument("--skin-frac", type=float, default=DEFAULT_SKIN_FRAC, help="list skin as a fraction of rc")
    ap.add_argument("--no-extras", action="store_true", help="skip the mixture, batched-replica and adjoint legs")
    ap.add_argument("--sort-every", type=int, default=DEFAULT_SORT_EVERY,
                    help="ticks between full re-sorts of the resident state")
    ap.add_argument("--aux-bound-search", action="store_true",
                    help="bond observation's bound pairs from the aux-grid search instead of the list pass")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a 256 MiB device copy (HBM counter calibration) and exit")
    ap.add_argument("--leg", choices=("ionic", "mixture", "batched", "adjoint"), default=None,
                    help="run only that leg and print its JSON record, for profiling")
    ap.add_argument("--batch-lanes", type=int, default=0,
                    help="batched leg: Verlet lists with this many lanes per atom instead of the autotuner's pick")
    args = ap.parse_args()
    if args.copy_probe:
        return copy_probe()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only: the path would shard by spatial "
                                   "domain with halo exchange, but no multi-GPU figure is produced; run with --gpus 1"}))
        return 2
    if args.leg == "mixture":
        print(json.dumps(mixture_leg(steps=args.steps, warmup=args.warmup)), flush=True)
        return 0
    if args.leg == "batched":
        print(json.dumps(batched_leg(ticks=args.steps, segment=min(100, args.steps), lanes=args.batch_lanes)),
              flush=True)
        return 0
    if args.leg == "adjoint":
        print(json.dumps(adjoint_leg(steps=args.steps, ckpt=min(50, args.steps))), flush=True)
        return 0
    out, w, fin = ionic_headline(args.steps, args.warmup, args.lanes, args.skin_frac,
                                 aux_bound_search=args.aux_bound_search, sort_every=args.sort_every)
    if args.leg == "ionic":
        print(json.dumps(out), flush=True)
        return 0
    if not args.no_extras:
        out["mixture"] = mixture_leg()
        out["batched"] = batched_leg()
        out["adjoint"] = adjoint_leg()
    if not args.no_cpu_baseline:
        st = w.state.copy()
        st.x, st.y, st.z, st.vx, st.vy, st.vz = fin.x, fin.y, fin.z, fin.vx, fin.vy, fin.vz
        st.ax, st.ay, st.az = fin.ax, fin.ay, fin.az
        out["cpu_baseline"] = cpu_baseline(w, st)
        out["gpu_over_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
