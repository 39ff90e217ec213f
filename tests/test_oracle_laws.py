"""The oracle's scalar laws against the reference's own closed-form tests.

Each test restates one #[test] of src/physics/forces.rs:394-624 (values and
tolerances as there) and runs it against oracle/oracle.cpp.
"""
import math

import pytest

from genesis_amd.params import Physics, SPECIES_INDEX, DEFAULT_MASS
from oracle import oracle as O

H, C, O_, NA, FE = (SPECIES_INDEX[s] for s in ("H", "C", "O", "Na", "Fe"))


def table(ph=None):
    return O.lj_table(ph or Physics())


def test_potential_minimum_at_sigma_root_sixth_of_2():
    # src/physics/forces.rs:394-405
    sig, eps, _ = table()
    s, e = sig[H, H], eps[H, H]
    r_min = s * 2.0 ** (1.0 / 6.0)
    m, v = O.lj_raw(s, e, r_min)
    assert abs(v + e) < 1e-12 * max(e, 1.0)
    assert abs(m) < 1e-9


def test_repulsive_and_attractive_force():
    # src/physics/forces.rs:407-420
    sig, eps, _ = table()
    s, e = sig[C, C], eps[C, C]
    assert O.lj_raw(s, e, 0.8 * s)[0] > 0.0
    assert O.lj_raw(s, e, 1.5 * s)[0] < 0.0
    assert abs(O.lj_raw(s, e, s * 2.0 ** (1.0 / 6.0))[0]) < 1e-9


def test_mixing_rules_are_symmetric():
    # src/physics/forces.rs:422-432
    sig, eps, _ = table()
    assert sig[H, C] == sig[C, H]
    assert eps[H, C] == eps[C, H]
    assert abs(eps[H, C] - 0.01 * math.sqrt(12.0 * 80.0)) < 1e-12


def test_bond_harmonic_switched_is_exact_derivative():
    # src/physics/forces.rs:434-459
    k, r_eq, r_on, rc, h = 5.0, 2.0, 2.0, 2.6, 1e-6

    def v_sw(x):
        _, v = O.bond_harmonic_raw(k, r_eq, x)
        sw, _ = O.smooth_cutoff(x, r_on, rc)
        return v * sw

    for r in (2.0, 2.1, 2.3, 2.5, 2.59):
        m, v = O.bond_harmonic_raw(k, r_eq, r)
        sw, dsw = O.smooth_cutoff(r, r_on, rc)
        f, _ = O.switched(m, v, sw, dsw)
        num = -(v_sw(r + h) - v_sw(r - h)) / (2.0 * h)
        assert abs(f - num) < 1e-4 * max(abs(f), 1.0), (r, f, num)


def test_switch_vanishes_at_cutoff():
    # src/physics/forces.rs:461-475
    ph = Physics()
    sig, eps, cut = table(ph)
    rc, r_on = cut["lj_rc"], cut["lj_ron"]
    assert rc > sig[O_, O_]
    f, v = O.force_switched_scalar(ph, O_, O_, rc)
    assert f == 0.0 and v == 0.0
    r_in = 0.8 * r_on
    f_in, _ = O.force_switched_scalar(ph, O_, O_, r_in)
    m, _ = O.lj_raw(sig[O_, O_], eps[O_, O_], r_in)
    assert abs(f_in - m) < 1e-12


def test_well_curvature_matches_numeric_second_derivative():
    # src/physics/forces.rs:477-490
    eps, s, h = 0.8, 1.9, 1e-5
    r_m = s * 2.0 ** (1.0 / 6.0)
    v = lambda r: O.lj_raw(s, eps, r)[1]
    numeric = (v(r_m + h) + v(r_m - h) - 2.0 * v(r_m)) / (h * h)
    assert abs(numeric - O.well_curvature(eps, s)) / numeric < 1e-4
    # closed form 72 2^(-1/3) eps / sigma^2
    assert O.well_curvature(eps, s) == pytest.approx(72.0 * 2.0 ** (-1.0 / 3.0) * eps / (s * s), rel=1e-15)


def test_vib_period_is_physically_ordered():
    # src/physics/forces.rs:492-521
    sig, eps, _ = table()
    t_cc = O.vib_period(eps[C, C], sig[C, C], DEFAULT_MASS[C] / 2.0)
    t_hh = O.vib_period(eps[H, H], sig[H, H], DEFAULT_MASS[H] / 2.0)
    t_fe = O.vib_period(eps[FE, FE], sig[FE, FE], DEFAULT_MASS[FE] / 2.0)
    for t in (t_cc, t_hh, t_fe):
        assert 1.0 <= t < 20.0
    assert abs(t_cc - t_hh) > 0.5


def test_smooth_cutoff_is_continuous_and_vanishes_at_cutoff():
    # src/physics/forces.rs:523-535
    sw0, dsw0 = O.smooth_cutoff(2.0, 2.0, 3.0)
    assert sw0 == 1.0 and dsw0 == 0.0
    sw1, dsw1 = O.smooth_cutoff(3.0, 2.0, 3.0)
    assert abs(sw1) < 1e-12 and abs(dsw1) < 1e-12
    mid, _ = O.smooth_cutoff(2.5, 2.0, 3.0)
    assert 0.0 < mid < 1.0
    # the polynomial itself at u = 1/2: 1 - 10/8 + 15/16 - 6/32 = 1/2, and
    # dsw = -30 (1/4)(1/4) / w
    assert mid == pytest.approx(0.5, abs=1e-15)
    assert O.smooth_cutoff(2.5, 2.0, 3.0)[1] == pytest.approx(-30.0 / 16.0, rel=1e-15)


def test_switch_derivative_matches_finite_difference():
    # dsw is the derivative of sw (what makes F = -d(V sw)/dr exact).
    for r in (2.05, 2.3, 2.77, 2.95):
        h = 1e-6
        num = (O.smooth_cutoff(r + h, 2.0, 3.0)[0] - O.smooth_cutoff(r - h, 2.0, 3.0)[0]) / (2 * h)
        assert O.smooth_cutoff(r, 2.0, 3.0)[1] == pytest.approx(num, rel=1e-7, abs=1e-9)


def test_coulomb_signs_and_potential():
    # src/physics/forces.rs:537-550
    f, v = O.coulomb_raw(1.0, 1.0, -1.0, 2.0)
    assert f < 0 and v < 0
    assert abs(f + 0.25) < 1e-12 and abs(v + 0.5) < 1e-12
    assert O.coulomb_raw(1.0, 1.0, 1.0, 2.0)[0] > 0
    assert abs(O.coulomb_raw(1.0, 1.0, 1.0, 4.0)[0] - 0.0625) < 1e-12


def test_gravity_is_always_attractive():
    # src/physics/forces.rs:552-558
    f, v = O.gravity_raw(1.0, 2.0, 3.0, 2.0)
    assert f < 0 and v < 0
    assert abs(f + 1.5) < 1e-12 and abs(v + 3.0) < 1e-12


def test_affinity_table_overrides_change_lj_potential():
    # src/physics/forces.rs:568-594
    ph = Physics().with_overrides({"Na": {"sigma": 1.0, "epsilon_k": 50.0}})
    sig, eps, _ = O.lj_table(ph)
    assert sig[NA, NA] == 1.0
    assert abs(eps[NA, NA] - 0.5) < 1e-12
    _, v = O.lj_raw(sig[NA, NA], eps[NA, NA], sig[NA, NA] * 2.0 ** (1.0 / 6.0))
    assert abs(v + eps[NA, NA]) < 1e-12 * eps[NA, NA]


def test_affinity_table_override_keeps_unmentioned_fields_and_rejects_unknown():
    # src/physics/forces.rs:596-624
    ph = Physics().with_overrides({"O": {"charge": -2.0}})
    assert ph.charge[O_] == -2.0
    assert ph.sigma[O_] == Physics().sigma[O_] and ph.epsilon_k[O_] == Physics().epsilon_k[O_]
    with pytest.raises(ValueError):
        Physics().with_overrides({"Xx": {}})


def test_cutoffs_follow_max_sigma_over_the_whole_table():
    # LjTable::new uses max sigma over all 10 elements (src/physics/forces.rs:224-236):
    # default max sigma = 2.3 (Si) -> rc = 5.75, r_on = 0.9 rc.
    _, _, cut = table()
    assert cut["lj_rc"] == 2.5 * 2.3 and cut["lj_ron"] == 0.9 * 2.5 * 2.3
    ph = Physics().with_overrides({"C": {"sigma": 3.4}})
    assert O.lj_table(ph)[2]["lj_rc"] == pytest.approx(8.5, abs=1e-15)
    # ForceSystem::new: the grid cutoff is the largest enabled cutoff.
    ph2 = Physics(enable_electrostatics=True, coulomb_cutoff=3.0)
    assert O.lj_table(ph2)[2]["rc"] == pytest.approx(3.0 * 2.3)


def test_hardened_core_freezes_force_and_potential():
    # force_switched_scalar clamps r at LJ_NUCLEUS sigma (src/physics/forces.rs:171-173,271).
    ph = Physics()
    sig, eps, _ = table(ph)
    f_core, v_core = O.force_switched_scalar(ph, C, C, 0.5 * sig[C, C])
    for r in (0.1 * sig[C, C], 0.3 * sig[C, C], 0.49 * sig[C, C]):
        assert O.force_switched_scalar(ph, C, C, r) == (f_core, v_core)
