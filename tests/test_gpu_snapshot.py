"""save_universe / load_universe on the GPU (genesis_amd/snapshot.py over
gs_snapshot_append / gs_snapshot_load).

* save -> load -> continue against an uninterrupted run of the same ticks:
  bit for bit (positions, velocities, accelerations, every tick's
  observables, the bond tracker's counts and the Bonds adjacency) on the
  per-tick cell tiles and on the Verlet lists, in the bonded regime of a
  small Na/O melt (bonds persist after about 116 ticks at 0.1 periods).
* The cell-staged lists are not in the image: that universe re-sorts on its
  first tick and continues to within rounding.
* The reference's own round trip (binary_roundtrip,
  src/serialization/mod.rs:188-207): a universe saved after 10 ticks loads
  with the same entity count, tick, time, name and components; a file that
  carries only the reference's Snapshot (no device image) loads too.
"""
import ctypes

import numpy as np
import pytest

from genesis_amd import snapshot as S
from genesis_amd.inputs import ionic_melt, random_gas
from genesis_amd.params import Physics, SPECIES_INDEX

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-snapshot] {msg}", flush=True)


def melt(n_side=12, seed=1105):
    w = ionic_melt(seed=seed, nx=n_side, ny=n_side, nz=n_side, steps=60)
    ph = Physics(**{**w.physics.__dict__, "bond_min_periods": 0.1})
    return ph, w.state


def _tune(u, mode, lanes):
    if mode:
        u.set_tuning(mode, lanes)


def _leg(u, n):
    obs = u.step(n)
    st = u.download()
    return obs, st, u.bond_stats(), u.bonds(), u.tick()


@pytest.mark.parametrize("mode,lanes", [(0, 0), (1, 2), (1, 4)])
def test_save_load_continue_is_bitwise_identical(tmp_path, mode, lanes):
    from genesis_amd.universe import GpuUniverse
    ph, st = melt()
    n1, n2 = 130, 60
    path = str(tmp_path / f"melt_{mode}_{lanes}.bin")
    say(f"uninterrupted n={st.n} mode={mode} lanes={lanes}: {n1} + {n2} ticks")
    with GpuUniverse(ph, st) as u:
        _tune(u, mode, lanes)
        u.step(n1)
        ref = _leg(u, n2)
        ref_tun = u.tuning()
    say("interrupted: save after the first leg, load, continue")
    with GpuUniverse(ph, st) as u:
        _tune(u, mode, lanes)
        u.step(n1)
        assert u.bond_stats()["bonded_pairs"] > 0
        u.save(path)
    v, cfg, ustate = S.load_universe(path)
    with v:
        assert ustate.tick == n1 and v.tick() == n1
        assert cfg.enable_bond_observation and cfg.enable_bond_interaction
        got = _leg(v, n2)
        got_tun = v.tuning()
    (o1, s1, b1, (p1, i1), t1), (o2, s2, b2, (p2, i2), t2) = ref, got
    assert t1 == t2 == n1 + n2
    assert np.array_equal(o1, o2)
    for a in ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az"):
        assert np.array_equal(getattr(s1, a), getattr(s2, a)), a
    assert b1["bonded_pairs"] > 0
    for k in ("tick", "bonded_pairs", "bonded_entities", "bonds_formed", "lifetime_sum_ticks"):
        assert b1[k] == b2[k], k
    assert np.array_equal(b1["species_matrix"], b2["species_matrix"])
    assert np.array_equal(p1, p2) and np.array_equal(i1, i2)
    assert (got_tun["pair_mode"], got_tun["lanes_per_atom"]) == (ref_tun["pair_mode"], ref_tun["lanes_per_atom"])
    assert got_tun["list_builds"] == ref_tun["list_builds"]


def test_cell_staged_lists_continue_within_rounding(tmp_path):
    from genesis_amd.universe import GpuUniverse
    ph, st = melt()
    n1, n2 = 130, 40
    path = str(tmp_path / "melt_staged.bin")
    say(f"cell-staged lists n={st.n}: {n1} + {n2} ticks")
    with GpuUniverse(ph, st) as u:
        u.set_tuning(2, 0)
        u.step(n1)
        ref = _leg(u, n2)
    with GpuUniverse(ph, st) as u:
        u.set_tuning(2, 0)
        u.step(n1)
        u.save(path)
    v, _, _ = S.load_universe(path)
    with v:
        assert v.tuning()["pair_mode"] == 2
        got = _leg(v, n2)
    np.testing.assert_allclose(got[0][:, 5], ref[0][:, 5], rtol=1e-9)
    np.testing.assert_allclose(got[1].x, ref[1].x, rtol=0, atol=1e-8)
    assert got[2]["bonded_pairs"] == ref[2]["bonded_pairs"]


def test_binary_roundtrip_like_the_reference(tmp_path):
    from genesis_amd.universe import GpuUniverse
    L = 40.0
    sp = tuple(SPECIES_INDEX[s] for s in ("H", "C", "O", "Na"))
    st = random_gas(seed=188, n=2000, box=(L, L, L), species_pool=sp)
    ph = Physics(box=(L, L, L))
    path = str(tmp_path / "roundtrip.bin")
    say(f"binary round trip n={st.n}")
    with GpuUniverse(ph, st) as u:
        u.step(10)
        saved = u.download()
        u.save(path)
    cfg, ustate, world, dev = S.read_snapshot(path)
    assert dev is not None and world.n == st.n
    assert ustate.tick == 10 and ustate.t == S.advance_time(0.0, 0, 10, ph.dt)
    assert cfg.name == "Genesis"
    # the file without the device image: what the reference writes
    with open(path, "rb") as f:
        plain = f.read()[:dev]
    ref_path = str(tmp_path / "reference_only.bin")
    with open(ref_path, "wb") as f:
        f.write(plain)
    for p, img in ((path, True), (ref_path, False)):
        v, cfg2, st2 = S.load_universe(p)
        with v:
            assert v.n == st.n and v.tick() == 10 and st2.t == ustate.t and cfg2.name == cfg.name
            back = v.download()
            for a in ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az"):
                assert np.array_equal(getattr(back, a), getattr(saved, a)), (img, a)
            assert np.array_equal(back.species, st.species)
            v.step(1)
            assert v.tick() == 11


def test_an_image_of_other_parameters_is_rejected(tmp_path):
    from genesis_amd.universe import GpuUniverse
    L = 30.0
    st = random_gas(seed=5, n=500, box=(L, L, L))
    path = str(tmp_path / "gas.bin")
    with GpuUniverse(Physics(box=(L, L, L)), st) as u:
        u.step(3)
        u.save(path)
    _, _, _, dev = S.read_snapshot(path)
    say("device image into a universe with another dt")
    with GpuUniverse(Physics(box=(L, L, L), dt=0.01), st) as w:
        rc = w.lib.gs_snapshot_load(w.h, path.encode(), ctypes.c_int64(dev))
        assert rc != 0 and b"other parameters" in w.lib.gs_last_error()
        rc = w.lib.gs_snapshot_load(w.h, path.encode(), ctypes.c_int64(0))
        assert rc != 0 and b"no device image" in w.lib.gs_last_error()
