"""The snapshot format (genesis_amd/snapshot.py) on the CPU.

The reference holds no snapshot fixture (its one test, binary_roundtrip in
src/serialization/mod.rs:188-207, saves and reloads in memory), so the
layout is pinned against bincode 1's published encoding, byte by byte, on
inputs small enough to write out by hand: fixed-width little-endian
integers and f64, u64 lengths, a one-byte Option tag, u32 variant indices.
The xoshiro seeding is pinned on SplitMix64's published outputs for seed 0.
"""
import numpy as np
import pytest

from genesis_amd import snapshot as S
from genesis_amd.params import Physics, SPECIES_INDEX
from genesis_amd.stats import StatsCollector, StatsSnapshot


def _one_entity():
    nan3 = np.full((1, 3), np.nan)
    return S.World(index=np.array([3], np.uint32), generation=np.array([1], np.uint32),
                   pos=np.array([[1.0, 2.0, 3.0]]), vel=nan3.copy(), mass=np.array([1.5]),
                   charge=np.array([np.nan]), species=np.array([SPECIES_INDEX["O"]], np.uint8),
                   bonds=[[(7, 0)]], acc=nan3.copy())


def test_entity_snapshot_bytes_follow_the_bincode_layout():
    w = S.Writer()
    S._put_world(w, _one_entity())
    expected = bytes.fromhex(
        "0100000000000000"                                          # Vec<EntitySnapshot> length
        "03000000" "01000000"                                       # index, generation
        "01" "000000000000f03f" "0000000000000040" "0000000000000840"  # Some(Position(Vec3 1, 2, 3))
        "00"                                                        # velocity: None
        "01" "000000000000f83f"                                     # Some(Mass(1.5))
        "00"                                                        # charge: None
        "01" "04000000"                                             # Some(AtomType::Oxygen), variant 4
        "01" "0100000000000000" "07000000" "00000000"               # Some(Bonds { [EntityId 7, gen 0] })
        "00")                                                       # acceleration: None
    assert w.bytes() == expected
    back = S._get_world(S.Reader(expected))
    assert back.n == 1 and back.index[0] == 3 and back.generation[0] == 1
    assert np.array_equal(back.pos[0], [1.0, 2.0, 3.0]) and np.isnan(back.vel).all()
    assert back.mass[0] == 1.5 and np.isnan(back.charge[0]) and back.species[0] == SPECIES_INDEX["O"]
    assert back.bonds == [[(7, 0)]] and np.isnan(back.acc).all()


def test_default_config_is_288_bytes_and_opens_with_the_name():
    # UniverseConfig: name 8 + 7, size 24, dt 8, initial_atoms 8, elements
    # 8 + 10 x 4, stats_history 8 = 111; RngConfig 8; PhysicsConfig 10 x 8 +
    # an empty map's 8 = 88; SystemsConfig 9 bools; StatsConfig 8 + 8 + 8 +
    # (8 + 0) + 8 + (8 + 0) + 8 + 8 + 8 = 72
    w = S.Writer()
    S._put_config(w, S.Config())
    b = w.bytes()
    assert len(b) == 111 + 8 + 88 + 9 + 72
    assert b[:15] == bytes.fromhex("0700000000000000") + b"Genesis"
    assert b[15:23] == bytes.fromhex("0000000000006040")  # size.x = 128.0
    # the seed, 42, right after the universe section
    assert b[111:119] == bytes.fromhex("2a00000000000000")
    # SystemsConfig::default: movement, boundaries, forces on
    assert b[111 + 8 + 88:111 + 8 + 88 + 9] == bytes([1, 1, 0, 1, 0, 0, 0, 0, 0])


def test_element_override_encodes_options_per_field():
    w = S.Writer()
    S._put_config(w, S.Config(element_overrides={"Na": {"sigma": 2.5, "epsilon_k": None, "charge": -1.0}}))
    b = w.bytes()
    entry = (bytes.fromhex("0100000000000000") + bytes.fromhex("0200000000000000") + b"Na" + b"\x01" +
             bytes.fromhex("0000000000000440") + b"\x00" + b"\x01" + bytes.fromhex("000000000000f0bf"))
    assert entry in b
    assert len(b) == 288 + len(entry) - 8


def test_splitmix_seeding_matches_published_outputs():
    # SplitMix64 from state 0 (Steele, Lea & Flood; Vigna's splitmix64.c)
    assert S.splitmix_state(0) == (0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F, 0xF88BB8A8724C81EC)


def test_time_accumulates_dt_once_per_tick():
    dt = 1.0 / 60.0
    t = S.advance_time(0.0, 0, 10, dt)
    assert t == sum([dt] * 10)  # the same left-to-right sum
    assert S.advance_time(t, 10, 10, dt) == t


def test_snapshot_roundtrip_with_stats_and_overrides():
    ph = Physics(box=(40.0, 40.0, 40.0)).with_overrides({"O": {"charge": -0.8}})
    cfg = S.Config.from_physics(ph, 2)
    assert cfg.element_overrides == {"O": {"sigma": None, "epsilon_k": None, "charge": -0.8}}
    col = StatsCollector(4)
    snap = StatsSnapshot(tick=5, time=5 / 60, entities=2, bond_matrix=[0.0] * 100,
                         structure={"tick": 5, "monomers": 2, "aggregates": 0, "largest": 0, "mean_size": 1.0,
                                    "bound_pairs": 0},
                         chemical={"tick": 5, "aggregates": 1, "bound_entities": 2, "monomers": 0, "largest": 2,
                                   "compositions": [{"formula": "Na-O", "count": 1, "mean_binding": -3.5}],
                                   "appeared": 1, "disappeared": 0, "fusions": 0, "scissions": 0,
                                   "reactions": [{"reactants": ["Na", "O"], "products": ["Na-O"], "count": 1}]})
    col.record(snap)
    st = S.UniverseState(t=5 / 60, tick=5, dt=1 / 60, rng=S.splitmix_state(42), stats=col)
    wd = _one_entity()
    data = S.encode_snapshot(cfg, st, wd)
    assert data[:4] == bytes.fromhex("04000000")
    cfg2, st2, wd2, end = S.decode_snapshot(data + S.DEVICE_MAGIC)
    assert end == len(data)
    assert cfg2 == cfg
    assert (st2.t, st2.tick, st2.dt, st2.rng) == (st.t, st.tick, st.dt, st.rng)
    assert st2.stats.history_cap == 4 and len(st2.stats.history()) == 1
    assert st2.stats.snapshot == snap and st2.stats.history()[0] == snap
    assert wd2.bonds == wd.bonds
    ph2 = cfg2.to_physics(ph.mass)
    assert ph2.charge == ph.charge and ph2.box == ph.box and ph2.sigma == ph.sigma


def test_other_format_versions_are_rejected():
    data = S.encode_snapshot(S.Config(), S.UniverseState(), _one_entity())
    bad = bytes.fromhex("03000000") + data[4:]
    with pytest.raises(ValueError, match="v3"):
        S.decode_snapshot(bad)
    with pytest.raises(ValueError, match="truncated"):
        S.decode_snapshot(data[:-3])


def test_out_of_range_atom_type_is_rejected():
    w = S.Writer()
    S._put_world(w, _one_entity())
    b = bytearray(w.bytes())
    k = b.index(bytes.fromhex("0104000000"))
    b[k + 1] = 10
    with pytest.raises(ValueError, match="AtomType"):
        S._get_world(S.Reader(bytes(b)))
