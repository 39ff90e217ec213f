"""Stats snapshot arithmetic and the CSV/XYZ exports (host side).

The expected strings are the reference's own test vectors
(src/export/mod.rs:186-273: csv_row_matches_header_column_count and
xyz_frame_has_header_and_atoms); the snapshot arithmetic and the oracle's
stats sums are pinned by closed forms.
"""
import numpy as np
import pytest

from genesis_amd.export import (CSV_HEADER, append_csv, bond_species, chemical_compositions, chemical_reactions,
                                csv_row, fixed, frame_text, write_frame)
from genesis_amd.inputs import State
from genesis_amd.stats import StatsCollector, StatsSnapshot, snapshot_from_sums
from oracle import oracle as O


def reference_snapshot():
    """The snapshot of the reference's csv_row test (src/export/mod.rs:186-257)."""
    m = [0.0] * 100
    m[54] = 3.0  # (Na, O)
    return StatsSnapshot(
        tick=7, time=0.42, entities=3, energy_total=5.0, energy_avg=1.5, energy_potential=2.0,
        temperature_avg=300.0, mean_speed=1.1, density=0.02, collisions=1, systems_run=6, fps=10.0,
        memory_bytes=2048,
        structure={"tick": 7, "monomers": 1, "aggregates": 1, "largest": 2, "mean_size": 1.5, "bound_pairs": 1},
        bonded_pairs=3, bonds_formed=4, bond_lifetime_ticks=25.0, bond_energy=0.5, bond_matrix=m,
        chemical={"tick": 7, "aggregates": 2, "bound_entities": 5, "monomers": 1, "largest": 3,
                  "compositions": [{"formula": "Na-O", "count": 2, "mean_binding": -0.53},
                                   {"formula": "Na2-O", "count": 1, "mean_binding": -0.71}],
                  "appeared": 2, "disappeared": 0, "fusions": 1, "scissions": 1,
                  "reactions": [{"reactants": ["Na-O", "Na-O"], "products": ["Na2-O2"], "count": 2},
                                {"reactants": ["Na2-O2"], "products": ["Na-O", "Na-O"], "count": 1}]})


def test_csv_row_matches_header_column_count_and_reference_vector():
    s = reference_snapshot()
    row = csv_row(s)
    assert len(CSV_HEADER.split(",")) == len(row.split(",")) == 30
    assert row.startswith("7,0.420000,3,")
    assert row.endswith("Na-O:3,2,2,0,1,1,Na-O:2@-0.53;Na2-O:1@-0.71,Na-O+Na-O->Na2-O2x2;Na2-O2->Na-O+Na-O")


def test_csv_row_middle_columns():
    # energy_kinetic = total - potential; fps {:.3}; memory in KiB {:.1};
    # structure columns aggregates, largest, monomers, bound_pairs; bond
    # lifetime {:.1}
    cols = csv_row(reference_snapshot()).split(",")
    assert cols[3:10] == ["5.000000", "3.000000", "2.000000", "1.500000", "300.000000", "1.100000", "0.020000"]
    assert cols[10:14] == ["1", "6", "10.000", "2.0"]
    assert cols[14:22] == ["1", "2", "1", "1", "3", "4", "25.0", "0.500000"]


def test_csv_row_without_lenses_has_empty_cells():
    row = csv_row(StatsSnapshot(tick=1))
    cols = row.split(",")
    assert len(cols) == 30
    assert cols[14:18] == ["0", "0", "0", "0"] and cols[22] == "" and cols[28] == "" and cols[29] == ""


def test_bond_species_order_and_lower_triangle():
    m = np.zeros((10, 10))
    m[6, 4] = m[4, 6] = 12  # Si-O
    m[5, 4] = m[4, 5] = 3   # Na-O
    m[4, 4] = 3             # O-O (diagonal once)
    assert bond_species(m.reshape(-1)) == "Si-O:12;Na-O:3;O-O:3"
    assert bond_species(np.zeros(100)) == ""
    assert bond_species([]) == ""


def test_chemical_cells_when_the_lens_is_off():
    s = StatsSnapshot(tick=3)
    assert chemical_compositions(s) == "" and chemical_reactions(s) == ""


def test_fixed_point_formatting():
    assert fixed(0.42, 6) == "0.420000"
    assert fixed(-0.53, 2) == "-0.53"
    assert fixed(2048 / 1024.0, 1) == "2.0"
    assert fixed(float("nan"), 6) == "NaN"
    assert fixed(float("inf"), 6) == "inf" and fixed(float("-inf"), 2) == "-inf"


def test_xyz_frame_reference_vector(tmp_path):
    # src/export/mod.rs:259-273: two atoms, frame 0
    text = frame_text(0, [0, 2], [1.0, -1.0], [2.0, 0.0], [3.0, 0.5])
    lines = text.splitlines()
    assert lines[0] == "2" and lines[1] == "frame 0"
    assert sorted(lines[2:]) == ["C -1.000000 0.000000 0.500000", "H 1.000000 2.000000 3.000000"]

    class FakeUniverse:
        def download(self):
            z = np.zeros(2)
            return State(np.array([1.0, -1.0]), np.array([2.0, 0.0]), np.array([3.0, 0.5]), z, z, z,
                         np.array([0, 2], dtype=np.uint8), z)

    path = write_frame(str(tmp_path / "frame"), 0, FakeUniverse())
    assert path.endswith("frame_00000000.xyz")
    assert open(path).read() == text


def test_append_csv_writes_the_header_once(tmp_path):
    p = str(tmp_path / "m.csv")
    append_csv(p, reference_snapshot())
    append_csv(p, StatsSnapshot(tick=8))
    lines = open(p).read().splitlines()
    assert lines[0] == CSV_HEADER and len(lines) == 3 and lines[2].startswith("8,")


def test_snapshot_arithmetic_closed_form():
    # kinetic 3, two entities, k = 0.01: <K> = 1.5, T = (2/3) 1.5 / 0.01 = 100
    s = snapshot_from_sums(tick=5, time=0.25, entities=2, count=2, kinetic=3.0, speed_sum=4.0,
                           thermal_constant=0.01, volume=8.0, energy_potential=-1.25, bond_energy=0.5,
                           structure={"tick": 0, "aggregates": 9},
                           bond_observation={"bonded_pairs": 2, "bonds_formed": 4, "lifetime_sum_ticks": 10.0,
                                             "species_matrix": np.eye(10, dtype=np.int64)},
                           chemical={"tick": 5, "aggregates": 1, "compositions": [], "reactions": [],
                                     "appeared": 0, "disappeared": 0, "fusions": 0, "scissions": 0})
    assert s.energy_avg == 1.5 and s.temperature_avg == pytest.approx(100.0, rel=1e-15)
    assert s.mean_speed == 2.0 and s.density == 0.25
    assert s.energy_total == 3.0 - 1.25 + 0.5
    assert s.structure is None  # a structure summary of tick 0 was never sampled
    assert s.chemical is not None and s.bond_lifetime_ticks == 2.5 and s.bonded_pairs == 2
    assert s.bond_matrix[0] == 1.0 and s.bond_matrix[1] == 0.0 and len(s.bond_matrix) == 100
    z = snapshot_from_sums(0, 0.0, 0, 0, 0.0, 0.0, 0.0, 0.0)
    assert z.energy_avg == 0.0 and z.temperature_avg == 0.0 and z.mean_speed == 0.0 and z.density == 0.0


def test_collector_keeps_a_bounded_history():
    c = StatsCollector(3)
    for t in range(5):
        c.record(StatsSnapshot(tick=t))
    assert [s.tick for s in c.history()] == [2, 3, 4] and c.snapshot.tick == 4


def test_oracle_stats_sums_closed_forms():
    # |(3, 4, 0)| = 5 and |(1, 2, 2)| = 3 exactly; kinetic 0.5 m |v|^2
    kin, sp = O.stats_sums([3.0, 1.0], [4.0, 2.0], [0.0, 2.0], [2.0, 4.0])
    assert sp == 8.0 and kin == 0.5 * 2.0 * 25.0 + 0.5 * 4.0 * 9.0
    assert O.stats_sums([], [], [], []) == (0.0, 0.0)
    # left to right: 1e16 + 1 + 1 rounds back to 1e16 at each step
    kin, sp = O.stats_sums([1e8, 1.0, 1.0, 0.0], [0.0] * 4, [0.0] * 4, [2.0, 2.0, 2.0, 0.0])
    assert kin == 1e16 and sp == 1e8 + 2.0
