"""BondStructureSystem (src/systems/bond_structure.rs) on the CPU: the
lifecycle and composition logic of genesis_amd.structure against the
reference's own tests (src/systems/bond_structure.rs:351-510), and the
oracle's pair_potential_raw (src/physics/forces.rs:360-384) against closed
forms and the reference's affinity-table test."""
import math

import numpy as np
import pytest

from genesis_amd.params import Physics, SPECIES_INDEX
from genesis_amd.structure import chemical_structure
from oracle import oracle as O

NA, O_ = SPECIES_INDEX["Na"], SPECIES_INDEX["O"]


def dimers():
    # two_dimers_and_monomer (src/systems/bond_structure.rs:276-305)
    x = np.array([0.0, 2.0, 5.0, 7.0, 10.0])
    sp = np.array([NA, O_, NA, O_, NA], dtype=np.uint8)
    return x, sp, [(0, 1), (2, 3)]


def ring():
    # fused_ring (src/systems/bond_structure.rs:307-336)
    x = np.array([0.0, 1.2, 5.0, 6.2, 10.0])
    sp = np.array([NA, O_, NA, O_, NA], dtype=np.uint8)
    return x, sp, [(0, 1), (1, 2), (2, 3), (0, 3)]


def sample(ph, world, prev, tick):
    """One sample with the oracle's components and binding energies."""
    x, sp, edges = world
    z = np.zeros_like(x)
    ents = sorted({e for ed in edges for e in ed})
    comps = O.bond_components([(e, int(sp[e])) for e in ents], edges)
    e_edge = O.bond_binding(ph, x, z, z, sp, edges)
    binding = [float(sum(e for (a, b), e in zip(edges, e_edge) if a in c["members"])) for c in comps]
    return chemical_structure(comps, binding, len(x), prev, tick)


def test_first_sample_records_new_aggregates_with_negative_binding():
    ph = Physics(box=(40.0, 40.0, 40.0), enable_electrostatics=True, enable_bond_observation=True)
    cs, _ = sample(ph, dimers(), [], 1)
    assert (cs["aggregates"], cs["monomers"], cs["largest"]) == (2, 1, 2)
    assert (cs["appeared"], cs["disappeared"], cs["fusions"]) == (2, 0, 0)
    na_o = [c for c in cs["compositions"] if c["formula"] == "Na-O"][0]
    assert na_o["count"] == 2 and na_o["mean_binding"] < 0.0


def test_fusion_from_dimers_to_ring():
    ph = Physics(box=(40.0, 40.0, 40.0), enable_bond_observation=True)
    cs, prev = sample(ph, dimers(), [], 1)
    assert cs["appeared"] == 2
    cs, _ = sample(ph, ring(), prev, 2)
    assert (cs["aggregates"], cs["disappeared"], cs["appeared"], cs["fusions"]) == (1, 2, 1, 1)
    assert cs["reactions"] == [{"reactants": ["Na-O", "Na-O"], "products": ["Na2-O2"], "count": 1}]


def test_scission_from_ring_to_dimers():
    ph = Physics(box=(40.0, 40.0, 40.0), enable_bond_observation=True)
    cs, prev = sample(ph, ring(), [], 1)
    assert cs["aggregates"] == 1
    cs, _ = sample(ph, dimers(), prev, 2)
    assert (cs["aggregates"], cs["disappeared"], cs["appeared"], cs["scissions"]) == (2, 1, 2, 1)
    assert cs["reactions"] == [{"reactants": ["Na2-O2"], "products": ["Na-O", "Na-O"], "count": 1}]


def test_unchanged_graph_has_no_events_and_a_partial_fusion_is_not_one():
    # the same components again: nothing appears; a new component that holds
    # one previous component plus a fresh entity is neither fusion nor scission
    ph = Physics(box=(40.0, 40.0, 40.0))
    cs, prev = sample(ph, dimers(), [], 1)
    cs, prev = sample(ph, dimers(), prev, 2)
    assert (cs["appeared"], cs["disappeared"], cs["fusions"], cs["scissions"], cs["reactions"]) == (0, 0, 0, 0, [])
    x, sp, _ = dimers()
    cs, _ = sample(ph, (x, sp, [(0, 1), (2, 3), (3, 4)]), prev, 3)
    assert (cs["appeared"], cs["disappeared"], cs["fusions"], cs["scissions"]) == (1, 1, 0, 0)
    assert [c["formula"] for c in cs["compositions"]] == ["Na-O", "Na2-O"]


def test_affinity_table_tunes_binding_energy():
    # src/systems/bond_structure.rs:385-431: Na-O at 2.0 with electrostatics;
    # neutral charges make the pair less bound, and LJ alone still attracts
    ph = Physics(enable_electrostatics=True)
    charged = O.pair_potential_raw(ph, NA, O_, 2.0, with_coulomb=True, with_bond=False)
    ch = list(ph.charge)
    ch[NA] = ch[O_] = 0.0
    neutral = O.pair_potential_raw(Physics(enable_electrostatics=True, charge=tuple(ch)), NA, O_, 2.0,
                                   with_coulomb=True, with_bond=False)
    assert charged < neutral < 0.0


@pytest.mark.parametrize("a,b", [(NA, O_), (NA, NA), (SPECIES_INDEX["Si"], SPECIES_INDEX["H"])])
def test_pair_potential_raw_closed_forms(a, b):
    ph = Physics(enable_electrostatics=True)
    sig = 0.5 * (ph.sigma[a] + ph.sigma[b])
    eps = ph.thermal_constant * math.sqrt(ph.epsilon_k[a] * ph.epsilon_k[b])
    r_m = sig * 2.0 ** (1.0 / 6.0)
    # LJ minimum -eps at sigma 2^(1/6); zero crossing at sigma
    assert O.pair_potential_raw(ph, a, b, r_m, False, False) == pytest.approx(-eps, rel=1e-12)
    assert abs(O.pair_potential_raw(ph, a, b, sig, False, False)) < 1e-12 * eps
    # the Coulomb term is k_e q_a q_b / r with the table's charges
    for r in (0.7, 2.0, 4.5):
        lj = O.pair_potential_raw(ph, a, b, r, False, False)
        d = O.pair_potential_raw(ph, a, b, r, True, False) - lj
        # (the difference carries the rounding of the LJ term it cancels)
        assert d == pytest.approx(ph.coulomb_constant * ph.charge[a] * ph.charge[b] / r, rel=1e-12,
                                  abs=1e-15 * max(abs(lj), 1.0))
    # the bond term vanishes at r_eq = sigma 2^(1/6) and is k/2 (r - r_eq)^2 with
    # k = 72 2^(-1/3) eps / sigma^2 (the curvature of LJ at its minimum)
    k = 72.0 * 2.0 ** (-1.0 / 3.0) * eps / sig ** 2
    assert O.pair_potential_raw(ph, a, b, r_m, False, True) == pytest.approx(-eps, rel=1e-12)
    for dr in (-0.2, 0.3):
        d = O.pair_potential_raw(ph, a, b, r_m + dr, False, True) - O.pair_potential_raw(ph, a, b, r_m + dr, False, False)
        assert d == pytest.approx(0.5 * k * dr * dr, rel=1e-12)


def test_bond_binding_uses_the_minimum_image_and_the_floor():
    # an edge across the periodic boundary takes the nearest image; coincident
    # atoms are evaluated at r = 1e-9 (src/systems/bond_structure.rs:114)
    L = 20.0
    ph = Physics(box=(L, L, L))
    x = np.array([-9.5, 9.0, 3.0, 3.0])
    z = np.zeros(4)
    sp = np.array([NA, O_, NA, NA], dtype=np.uint8)
    e = O.bond_binding(ph, x, z, z, sp, [(0, 1), (2, 3)])
    assert e[0] == pytest.approx(O.pair_potential_raw(ph, NA, O_, 1.5), rel=1e-12)
    assert e[1] == O.pair_potential_raw(ph, NA, NA, 1e-9)
