"""Stats snapshot and exports fed from the device, through the C ABI.

Tolerance: the device reduces kinetic energy and the speed sum in a fixed
tree order, the reference (and the oracle) entity after entity; the two
agree to 1e-12 relative.  Integer fields (bond counts, the species matrix)
and the exported XYZ text are exact.
"""
import sys

import numpy as np
import pytest

from genesis_amd.export import CSV_HEADER, bond_species, csv_row, frame_text, write_frame
from genesis_amd.inputs import State, ionic_melt, mixture
from genesis_amd.params import Physics
from genesis_amd.stats import StatsSystem
from oracle import oracle as O

pytestmark = pytest.mark.gpu

REL = 1e-12


def say(msg):
    print(f"[gpu-stats] {msg}", flush=True)
    sys.stdout.flush()


def test_speed_stats_match_the_oracle_sums():
    from genesis_amd.universe import GpuUniverse
    w = mixture(n=32768, seed=91, steps=1)
    ph, st = w.physics, w.state
    say("speed stats, mixture 32768, after 20 ticks")
    with GpuUniverse(ph, st) as u:
        u.step(20, observe=False)
        kin, sp = u.speed_stats()
        k_step = u.kinetic()
        s = u.download()
    ref_k, ref_s = O.stats_sums(s.vx, s.vy, s.vz, np.asarray(ph.mass)[s.species])
    say(f"kinetic gpu {kin:.17g} oracle {ref_k:.17g}; speed sum gpu {sp:.17g} oracle {ref_s:.17g}")
    assert kin == pytest.approx(ref_k, rel=REL)
    assert sp == pytest.approx(ref_s, rel=REL)
    assert kin == pytest.approx(k_step, rel=REL)


def test_speed_stats_of_an_empty_universe():
    from genesis_amd.universe import GpuUniverse
    z = np.zeros(0)
    st = State(z, z, z, z, z, z, np.zeros(0, dtype=np.uint8), z)
    with GpuUniverse(Physics(box=(10.0, 10.0, 10.0)), st) as u:
        assert u.speed_stats() == (0.0, 0.0)


def test_stats_snapshot_and_csv_row_of_an_observed_melt():
    from genesis_amd.universe import GpuUniverse
    w = ionic_melt(nx=16, ny=16, nz=16, steps=1)
    ph, st = w.physics, w.state
    assert ph.enable_bond_observation
    say(f"stats snapshot, ionic melt n={st.n}, 30 ticks with bond observation")
    with GpuUniverse(ph, st) as u:
        obs = u.step(30)
        snap = StatsSystem().run(u, tick=30, time=30 * ph.dt, energy_potential=float(obs[-1, 1]),
                                 bond_energy=float(obs[-1, 2]))
        bs = u.bond_stats()
        s = u.download()
    ref_k, ref_s = O.stats_sums(s.vx, s.vy, s.vz, np.asarray(ph.mass)[s.species])
    n = st.n
    assert snap.entities == n
    assert snap.energy_avg == pytest.approx(ref_k / n, rel=REL)
    assert snap.mean_speed == pytest.approx(ref_s / n, rel=REL)
    assert snap.temperature_avg == pytest.approx((2.0 / 3.0) * (ref_k / n) / ph.thermal_constant, rel=REL)
    assert snap.density == n / (ph.box[0] * ph.box[1] * ph.box[2])
    assert snap.energy_total == pytest.approx(ref_k + obs[-1, 1] + obs[-1, 2], rel=1e-10)
    assert snap.bonded_pairs == bs["bonded_pairs"] and snap.bonds_formed == bs["bonds_formed"]
    assert snap.bond_matrix == [float(c) for c in bs["species_matrix"].reshape(-1)]
    row = csv_row(snap)
    cols = row.split(",")
    say(f"bonded pairs {snap.bonded_pairs}, species cell {cols[22]!r}")
    assert len(cols) == len(CSV_HEADER.split(","))
    assert cols[0] == "30" and cols[2] == str(n) and cols[18] == str(bs["bonded_pairs"])
    assert cols[22] == bond_species(bs["species_matrix"].reshape(-1))
    # the cell holds each unlike pair once and each like pair twice (the
    # matrix is incremented at [a, b] and [b, a], src/systems/bonds.rs:321)
    total = 0.0
    for e in (e for e in cols[22].split(";") if e):
        pair, count = e.split(":")
        a, b = pair.split("-")
        total += int(count) / (2.0 if a == b else 1.0)
    assert total == bs["bonded_pairs"]


def test_xyz_frame_of_the_resident_positions(tmp_path):
    from genesis_amd.universe import GpuUniverse
    w = mixture(n=4096, seed=92, steps=1)
    ph, st = w.physics, w.state
    with GpuUniverse(ph, st) as u:
        u.step(10, observe=False)
        path = write_frame(str(tmp_path / "mix"), 10, u)
        s = u.download()
    text = open(path).read()
    assert path.endswith("mix_00000010.xyz")
    assert text == frame_text(10, s.species, s.x, s.y, s.z)
    lines = text.splitlines()
    assert lines[0] == "4096" and lines[1] == "frame 10" and len(lines) == 4098
    xs = np.array([float(l.split()[1]) for l in lines[2:]])
    assert np.abs(xs - s.x).max() <= 5e-7
