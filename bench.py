#!/usr/bin/env python3
"""Benchmark of the MI355X force-and-integration path.

Workload: the 10-element mixture (32,768 atoms on a 32^3 lattice, FP64,
Lorentz-Berthelot LJ + truncated Coulomb sharing rc = 2.5 max sigma,
Berendsen NVT at 60 K with tau = 100 fs, dt = 0.5 fs, g(r) with 200 bins
to L/4 sampled every 100 ticks).  A step is one full tick of the reference
schedule (kick, drift, wrap, cell-list rebuild, forces, kick, Berendsen) over
the resident state; inputs are synthetic (seeded) and resident in HBM before
the timed region.

Prints one JSON line.  ``roofline`` is the pair kernel's FP64 roofline (its
binding resource; MI355X FP64 vector and matrix peaks are both 78.6 TFLOP/s),
``roofline_hbm`` the same kernel against HBM, ``cpu_baseline`` the oracle's
restatement of the reference CPU algorithm timed on this host.
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

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
# Algorithmic work per atom-step (compulsory HBM bytes of one tick, and the pair law's flops):
STEP_BYTES = 161.0          # x, v, a read+write (144) + species (1) + cell key (8) + charge (8)
PAIR_BYTES = 57.0           # pair kernel: x, y, z, q, species read (33) + a written (24)
FLOP_PER_PAIR = 82.0        # LJ + quintic switch + accumulate (70) + Coulomb (12), per ordered pair
# HBM traffic of one k_pair_list launch of the default setting (4 lanes, skin
# 0.10 rc) on this workload, from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE in
# KiB per dispatch, scaled by what the same counters read for a 256 MiB
# device copy (FETCH_SIZE 131084.5 KiB -> x2.0, WRITE_SIZE 262144 KiB -> x1.0);
# profiles/k_pair_list__mixture32768_pmc.md holds the runs.
PAIR_TRAFFIC_KIB = {"fetch": 12218.0, "write": 1584.4, "fetch_scale": 2.0, "write_scale": 1.0}
DEFAULT_LANES, DEFAULT_SKIN_FRAC = 4, 0.10


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--pair-mode", type=int, default=1, help="-1 autotune, 0 cell tiles, 1 Verlet list")
    ap.add_argument("--lanes", type=int, default=DEFAULT_LANES)
    ap.add_argument("--skin", type=float, default=None,
                    help=f"list skin (length units); default {DEFAULT_SKIN_FRAC} rc")
    ap.add_argument("--sort-every", type=int, default=0, help="list mode: ticks between full re-sorts")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the batched-replica and adjoint legs")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a 256 MiB device copy (HBM counter calibration) and exit")
    ap.add_argument("--leg", choices=("batched", "adjoint", "ionic"), default=None,
                    help="run only that leg (ticks from --steps) and print its JSON record, for profiling")
    args = ap.parse_args()
    if args.copy_probe:
        return copy_probe()
    if args.leg:
        if args.leg == "batched":
            rec = batched_leg(ticks=args.steps, segment=min(100, args.steps))
        elif args.leg == "adjoint":
            rec = adjoint_leg(steps=args.steps, ckpt=min(50, args.steps))
        else:
            rec = ionic_leg(ticks=args.steps, gr_every=min(100, args.steps))
        print(json.dumps(rec), flush=True)
        return 0
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only: the path would shard by spatial "
                                   "domain with halo exchange, but no multi-GPU figure is produced; run with --gpus 1"}))
        return 2

    from genesis_amd.inputs import mixture
    from genesis_amd.universe import GpuUniverse

    w = mixture()
    ph, st = w.physics, w.state
    n = st.n
    gr_every, gr_bins, gr_rmax = w.extra["gr_every"], w.extra["gr_bins"], w.extra["gr_rmax"]
    log(f"workload {w.name}: n={n} box={ph.box[0]:.4f} dims={ph.grid_dims()} steps={args.steps} "
        f"warmup={args.warmup}")

    u = GpuUniverse(ph, st)
    skin = DEFAULT_SKIN_FRAC * ph.force_rc() if args.skin is None else args.skin
    default_setting = args.pair_mode == 1 and args.lanes == DEFAULT_LANES and args.skin is None
    if args.pair_mode < 0:
        tuning = u.autotune(40)
        log(f"autotuned: {tuning}")
    else:
        u.set_tuning(args.pair_mode, args.lanes, skin=skin, sort_every=args.sort_every)
        tuning = u.tuning()
        log(f"tuning: {tuning}")

    def run(k, observe=False):
        done = 0
        while done < k:
            chunk = min(gr_every, k - done)
            u.step(chunk, observe=observe)
            done += chunk
            u.radial_distribution(gr_rmax, gr_bins)
    if args.warmup:
        run(args.warmup)
    builds0, replays0 = u.tuning()["list_builds"], u.tuning()["list_replays"]
    # timed region: on every 10th tick the pair kernel is launched with start
    # and stop HIP events bound to its dispatch (its average duration feeds
    # the roofline); sampling keeps the events' cost out of the rate
    u.timing(enable="pair", every=10)
    t0 = time.perf_counter()
    run(args.steps)
    t1 = time.perf_counter()
    tm = u.timing()
    builds = u.tuning()["list_builds"] - builds0
    replays = u.tuning()["list_replays"] - replays0
    elapsed = t1 - t0
    value = n * args.steps / elapsed
    ms_per_step = 1e3 * elapsed / args.steps
    # phase breakdown from a separate, untimed pass with four events per tick
    u.timing(enable="phases")
    run(min(args.steps, 500))
    ph_tm = u.timing()
    u.timing(enable=False)
    phases = {k: ph_tm[k] / max(ph_tm["steps"], 1) for k in ("rebuild_ms", "pair_ms", "kick_ms")}
    log(f"{value:.4e} atom-steps/s, {ms_per_step:.4f} ms/step; pair kernel {tm['pair_ms'] / tm['steps']:.4f} ms; "
        f"phases per step: rebuild {phases['rebuild_ms']:.4f} ms, pair {phases['pair_ms']:.4f} ms, "
        f"kick {phases['kick_ms']:.4f} ms; list builds {builds}, replays {replays}")
    kernel = {1: f"k_pair_list<{tuning['lanes_per_atom']} lanes> (pair forces + second kick + Berendsen)",
              2: "k_pair_cell (cell-staged lists; pair forces + second kick + Berendsen)"}.get(
                  tuning["pair_mode"], "k_pair")

    # pair count of the final state (ordered pairs within rc)
    fin = u.download()
    pair_ms = tm["pair_ms"] / max(tm["steps"], 1)
    # in-range ordered pairs per atom from a neighbour count on the final state
    npairs = count_pairs(fin, ph)
    flops_per_launch = FLOP_PER_PAIR * npairs
    tflops = flops_per_launch / (pair_ms * 1e-3) / 1e12
    gbs = PAIR_BYTES * n / (pair_ms * 1e-3) / 1e9
    out = {
        "metric": "atom-steps/s (force+Verlet), batched vs plain, at stated force tol; % HBM BW",
        "value": value,
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded lattice + Maxwell-Boltzmann, no dataset)",
        "config": {"workload": w.name, "atoms": n, "species": 10, "box": ph.box[0], "rc": ph.force_rc(),
                   "cells": list(ph.grid_dims()), "dt_fs": 0.5, "thermostat": "berendsen 60 K tau 200 ticks",
                   "coulomb": True, "gr": {"bins": gr_bins, "r_max": gr_rmax, "every": gr_every},
                   "parallelism": "single GPU"},
        "roofline": {"bound": "mfma", "pipe": "fp64 VALU (pair forces are not a dense contraction; the MI355X "
                                              "FP64 vector peak equals its FP64 matrix peak)",
                     "kernel": kernel, "achieved": tflops, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": tflops / FP64_PEAK_TFLOPS,
                     "traffic": (1024.0 * (PAIR_TRAFFIC_KIB["fetch"] * PAIR_TRAFFIC_KIB["fetch_scale"] +
                                           PAIR_TRAFFIC_KIB["write"] * PAIR_TRAFFIC_KIB["write_scale"])
                                 if default_setting else None),
                     "per_launch": {"ordered_pairs": npairs, "flop_per_pair": FLOP_PER_PAIR,
                                    "avg_ms": pair_ms}},
        "roofline_hbm": {"bound": "hbm", "kernel": kernel, "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": gbs / HBM_PEAK_GBS, "bytes_per_atom": PAIR_BYTES},
        "step_roofline": {"hbm_frac": value * STEP_BYTES / (HBM_PEAK_GBS * 1e9),
                          "fp64_frac": value * (FLOP_PER_PAIR * npairs / n) / (FP64_PEAK_TFLOPS * 1e12),
                          "bytes_per_atom_step": STEP_BYTES, "flop_per_atom_step": FLOP_PER_PAIR * npairs / n},
        "phases_ms_per_step": phases,
        "tuning": {k: tuning[k] for k in ("pair_mode", "lanes_per_atom", "skin", "sort_every", "ms_per_tick")},
        "list_builds_in_timed_region": builds,
        "list_replays_in_timed_region": replays,
    }
    u.close()
    if args.pair_mode >= 0 and not args.no_extras:
        # the project autotuner on the same workload, for comparison with the
        # fixed setting timed above
        with GpuUniverse(ph, st) as au:
            t = au.autotune(40)
        out["autotune"] = {k: t[k] for k in ("pair_mode", "lanes_per_atom", "skin", "ms_per_tick")}
        log(f"autotune pick: {out['autotune']}")
    if not args.no_extras:
        out["batched"] = batched_leg()
        out["adjoint"] = adjoint_leg()
        out["ionic_melt"] = ionic_leg()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(w, args.cpu_seconds)
        out["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)
    return 0


def batched_leg(ticks: int = 200, segment: int = 100) -> dict:
    """Temperature sweep: 64 replicas of log-uniform sizes in [512, 16384]
    (20 K .. 120 K), stepped (a) as one batch (one launch per phase, hipGraph
    segments of up to 100 ticks, autotuned pair path: Verlet lists or cell
    tiles) and (b) one universe after another through gs_step with its own
    default setting (the per-replica-launch plain path)."""
    from genesis_amd.batch import GpuBatch
    from genesis_amd.inputs import replica_sweep
    from genesis_amd.universe import GpuUniverse
    ws = replica_sweep()
    total = sum(w.state.n for w in ws)
    with GpuBatch([w.physics for w in ws], [w.state for w in ws]) as b:
        tune = b.autotune(20)
        b.step(segment, graph_segment=segment)
        builds0 = b.get_tuning()["list_builds"]
        t0 = time.perf_counter()
        b.step(ticks, graph_segment=segment)
        tb = time.perf_counter() - t0
        btun = b.get_tuning()
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
    Reports the reverse pass per tick relative to the forward tick."""
    import numpy as np
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
        t0 = time.perf_counter()
        loss, _ = u.adjoint_forward(th, keep, steps, ckpt, win, gr=gr, g0=g0)
        tf = time.perf_counter() - t0
        t0 = time.perf_counter()
        grad = u.adjoint_backward()
        tb = time.perf_counter() - t0
    log(f"adjoint: forward {1e3 * tf / steps:.4f} ms/tick, reverse {1e3 * tb / steps:.4f} ms/tick "
        f"({tb / tf:.2f}x)")
    return {"workload": f"adjoint mixture, {st.n} atoms, {steps} ticks, checkpoint every {ckpt}",
            "forward_ms_per_tick": 1e3 * tf / steps, "reverse_ms_per_tick": 1e3 * tb / steps,
            "reverse_over_forward": tb / tf, "loss": loss, "grad_norm": float(np.linalg.norm(grad))}


def ionic_leg(ticks: int = 200, gr_every: int = 100) -> dict:
    """The 1,048,576-atom Na/O rock-salt melt (128 x 128 x 64, Coulomb,
    Berendsen 300 K -> 80 K, dt = 0.5 fs) with bond observation and the bond
    interaction on every tick, g(r) (256 bins) every 100 ticks, and the
    components of the Bonds graph once at the end (timed apart)."""
    from genesis_amd.inputs import ionic_melt
    from genesis_amd.structure import BondStructureSystem
    from genesis_amd.universe import GpuUniverse
    w = ionic_melt()
    ph, st = w.physics, w.state
    n = st.n
    with GpuUniverse(ph, st) as u:
        tune = u.autotune(20)
        u.step(20, observe=False)
        t0 = time.perf_counter()
        done = 0
        while done < ticks:
            k = min(gr_every, ticks - done)
            u.step(k, observe=False)
            done += k
            u.radial_distribution(w.extra["gr_rmax"], w.extra["gr_bins"])
        te = time.perf_counter() - t0
        t0 = time.perf_counter()
        comps = u.bond_labels()
        tc = time.perf_counter() - t0
        # the bond-graph stoichiometry lens (components, formulas, binding
        # energies, lifecycle against the previous sample), as sampled every
        # 500 ticks in this workload
        lens = BondStructureSystem(1)
        t0 = time.perf_counter()
        chem = lens.run(u, ticks, st.species)
        ts = time.perf_counter() - t0
        bs = u.bond_stats()
    v = n * ticks / te
    log(f"ionic melt: {v:.4e} atom-steps/s ({1e3 * te / ticks:.3f} ms/tick), {n} atoms, tuning {tune}, "
        f"bound episodes tracked, bonded pairs {bs['bonded_pairs']}")
    return {"workload": f"{w.name}: Na/O rock-salt melt, bonds observed every tick, interaction on, "
                        f"g(r) every {gr_every}",
            "atoms": n, "ticks": ticks, "atom_steps_per_s": v, "ms_per_tick": 1e3 * te / ticks,
            "bond_components_ms": 1e3 * tc, "bond_structure_ms": 1e3 * ts,
            "aggregates_at_end": chem["aggregates"], "bonded_pairs_at_end": bs["bonded_pairs"],
            "tuning": {k: tune[k] for k in ("pair_mode", "lanes_per_atom", "skin")}}


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


def count_pairs(s, ph) -> int:
    """Ordered pairs within the force cutoff (cell-binned numpy count)."""
    rc = ph.force_rc()
    box = np.asarray(ph.box)
    p = np.stack([s.x, s.y, s.z], 1)
    p = p - box * np.floor(p / box + 0.5)
    dims = np.maximum(np.floor(box / rc).astype(int), 1)
    cell = np.floor((p + 0.5 * box) / (box / dims)).astype(int) % dims
    key = cell[:, 0] + dims[0] * (cell[:, 1] + dims[1] * cell[:, 2])
    order = np.argsort(key, kind="stable")
    ks = key[order]
    starts = np.searchsorted(ks, np.arange(dims.prod() + 1))
    total = 0
    for c in range(int(dims.prod())):
        a = order[starts[c]:starts[c + 1]]
        if a.size == 0:
            continue
        cx, cy, cz = c % dims[0], (c // dims[0]) % dims[1], c // (dims[0] * dims[1])
        neigh = []
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    nc = ((cx + dx) % dims[0]) + dims[0] * (((cy + dy) % dims[1]) + dims[1] * ((cz + dz) % dims[2]))
                    neigh.append(order[starts[nc]:starts[nc + 1]])
        b = np.unique(np.concatenate(neigh))
        d = p[a][:, None, :] - p[b][None, :, :]
        d -= box * np.round(d / box)
        r2 = (d * d).sum(-1)
        total += int(((r2 < rc * rc) & (r2 > 0)).sum())
    return total


def cpu_baseline(w, seconds: float) -> dict:
    """The oracle's restatement of the reference CPU algorithm (serial grid +
    half-shell pair list, threaded fold with per-thread buffers merged in
    order, serial kinetic sum) timed on this host on a bounded sample of the
    same workload: as many full ticks as fit in ~``seconds``."""
    from oracle import oracle as O
    cores = max(1, min(os.cpu_count() or 1, 16))
    st = w.state
    O.run(w.physics, st, 1, nthreads=cores, observe=False)  # warm caches
    steps, t_used = 0, 0.0
    s = st
    while t_used < seconds and steps < 400:
        k = 1 if steps < 2 else max(1, min(20, int((seconds - t_used) / max(t_used / steps, 1e-3))))
        t0 = time.perf_counter()
        s, _ = O.run(w.physics, s, k, nthreads=cores, observe=False)
        t_used += time.perf_counter() - t0
        steps += k
    v = st.n * steps / t_used
    log(f"cpu baseline: {v:.4e} atom-steps/s over {steps} ticks ({t_used:.1f} s, {cores} threads)")
    return {"value": v, "unit": "atom-steps/s", "cores": cores, "kind": "port",
            "sample": f"{steps} full ticks of the {st.n}-atom mixture workload ({t_used:.1f} s), "
                      f"C++ restatement of the reference rayon force pass with {cores} threads"}


if __name__ == "__main__":
    sys.exit(main())
