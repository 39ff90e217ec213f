"""The System mirror driven call by call on the GPU (genesis_amd/systems.py:
one C-ABI call per system, in build_schedule order) against the fused
gs_step on the cell tiles: the same kernels' arithmetic in the same order,
so the trajectories agree bit for bit, bonds included."""
import numpy as np
import pytest

from genesis_amd import systems as S
from genesis_amd.inputs import ionic_melt, mixture
from genesis_amd.params import Physics

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-systems] {msg}", flush=True)


def _bonded_melt():
    w = ionic_melt(seed=1105, nx=8, ny=8, nz=8, steps=60)
    return Physics(**{**w.physics.__dict__, "bond_min_periods": 0.1}), w.state


def _thermostatted_mixture():
    w = mixture(n=4096, seed=4097, steps=1)
    return w.physics, w.state


@pytest.mark.parametrize("case,ticks", [("mixture", 40), ("bonded-melt", 150)])
def test_schedule_call_by_call_equals_gs_step(case, ticks):
    from genesis_amd.universe import GpuUniverse
    ph, st = _bonded_melt() if case == "bonded-melt" else _thermostatted_mixture()
    sched = S.build_schedule(ph)
    say(f"{case} n={st.n} ticks={ticks}: {sched.names()}")
    with GpuUniverse(ph, st) as u:
        ctx = S.SystemContext(u)
        for _ in range(ticks):
            S.tick(sched, ctx)
        a = u.download()
        ba, pa = u.bond_stats(), u.bonds()
        assert ctx.tick == ticks
    with GpuUniverse(ph, st) as v:
        v.set_tuning(0)  # the per-tick cell tiles, the force pass gs_forces runs
        obs = v.step(ticks)
        b = v.download()
        bb, pb = v.bond_stats(), v.bonds()
    for f in ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f
    assert ctx.resources["PotentialEnergy"] == obs[-1, 1]
    if ph.enable_thermostat:
        assert ctx.resources["lambda"] == obs[-1, 4]
    if ph.enable_bond_observation:
        assert bb["bonded_pairs"] > 0
        assert ctx.resources["BondEnergy"] == obs[-1, 2]
        for k in ("tick", "bonded_pairs", "bonded_entities", "bonds_formed", "lifetime_sum_ticks"):
            assert ba[k] == bb[k], k
        assert np.array_equal(pa[0], pb[0]) and np.array_equal(pa[1], pb[1])


@pytest.mark.parametrize("mode,lanes", [(0, 0), (1, 2)])
def test_call_splitting_and_segments_are_bitwise_invisible(mode, lanes):
    # one call of 1030 ticks runs as segments of 1024 and 6 ticks (each with
    # its own replay snapshot); two calls of 500 and 530 ticks split it
    # elsewhere: the bonded melt ends bit for bit the same either way
    from genesis_amd.universe import GpuUniverse
    w = ionic_melt(seed=1106, nx=16, ny=16, nz=16, steps=60)
    ph = Physics(**{**w.physics.__dict__, "bond_min_periods": 0.1})
    st = w.state
    outs = []
    for split in ((1030,), (500, 530)):
        say(f"split {split} mode={mode} lanes={lanes} n={st.n}")
        with GpuUniverse(ph, st) as u:
            u.set_tuning(mode, lanes or 4)
            obs = np.concatenate([u.step(k) for k in split])
            outs.append((u.download(), obs, u.bond_stats(), u.bonds(), u.tick(), u.tuning()))
    (a, oa, ba, pa, ta, ua), (b, ob, bb, pb, tb, ub) = outs
    assert ta == tb == 1030 and ba["bonded_pairs"] > 0
    assert np.array_equal(oa, ob)
    for f in ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f
    for k in ("tick", "bonded_pairs", "bonded_entities", "bonds_formed", "lifetime_sum_ticks"):
        assert ba[k] == bb[k], k
    assert np.array_equal(pa[0], pb[0]) and np.array_equal(pa[1], pb[1])
    assert ua["pair_mode"] == ub["pair_mode"] == mode
