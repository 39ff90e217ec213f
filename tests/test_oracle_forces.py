"""The oracle's force pass (ForceSystem::run, src/systems/forces.rs:163-353)
pinned by what the mathematics fixes: finite differences of the total
potential, Newton's third law, brute-force pair enumeration, a two-atom
closed form, and thread-count independence of the accelerations."""
import numpy as np
import pytest

from genesis_amd.inputs import State, mixture, plain_lj
from genesis_amd.params import DEFAULT_MASS, GRID_MINIMAGE, Physics, SPECIES_INDEX
from oracle import oracle as O
from tests.helpers import reference_like

C = SPECIES_INDEX["C"]


def small_mixture(n=512, seed=11):
    w = mixture(seed=seed, n=n, temperature=60.0, steps=1)
    return w.physics, w.state


def test_two_atoms_closed_form():
    ph = Physics(box=(40.0, 40.0, 40.0))
    sig, eps, cut = O.lj_table(ph)
    for r in (1.7, 2.1, 3.0, 5.3, 5.7):
        st = State(np.array([0.0, r]), np.zeros(2), np.zeros(2), np.zeros(2), np.zeros(2), np.zeros(2),
                   np.array([C, C], dtype=np.uint8), np.zeros(2))
        out = O.forces(ph, st)
        f, v = O.force_switched_scalar(ph, C, C, r)
        m = DEFAULT_MASS[C]
        # normal points from b to a; atom 0 is at -r relative to atom 1
        assert out["ax"][0] == pytest.approx(-f / m, rel=1e-15, abs=1e-300)
        assert out["ax"][1] == pytest.approx(f / m, rel=1e-15, abs=1e-300)
        assert out["V"] == v
        assert out["pe"].sum() == pytest.approx(v, rel=1e-15, abs=1e-300)


def test_finite_difference_of_total_potential():
    ph, st = small_mixture()
    base = O.forces(ph, st)
    m = np.asarray(ph.mass)[st.species]
    rng = np.random.default_rng(0)
    for i in rng.choice(st.n, 8, replace=False):
        for axis, arr in (("ax", "x"), ("ay", "y"), ("az", "z")):
            h = 1e-6
            sp, sm = st.copy(), st.copy()
            getattr(sp, arr)[i] += h
            getattr(sm, arr)[i] -= h
            dv = (O.forces(ph, sp, extras=False)["V"] - O.forces(ph, sm, extras=False)["V"]) / (2 * h)
            f = base[axis][i] * m[i]
            assert f == pytest.approx(-dv, rel=2e-6, abs=2e-7 * max(1.0, base["sabs"][i])), (i, axis)


def test_newton_third_law_total_force_vanishes():
    ph, st = small_mixture(n=1000, seed=3)
    out = O.forces(ph, st)
    m = np.asarray(ph.mass)[st.species]
    for a in ("ax", "ay", "az"):
        tot = (out[a] * m).sum()
        assert abs(tot) < 1e-11 * out["sabs"].sum()


def test_grid_and_brute_force_passes_agree():
    ph, st = small_mixture(n=1000, seed=5)
    g = O.forces(ph, st)
    ph_b = Physics(**{**ph.__dict__, "grid_semantics": GRID_MINIMAGE})
    b = O.forces(ph_b, st)
    assert g["npairs"] == b["npairs"]
    for a in ("ax", "ay", "az"):
        np.testing.assert_allclose(g[a], b[a], rtol=0, atol=1e-12 * np.abs(g[a]).max())
    assert g["V"] == pytest.approx(b["V"], rel=1e-12)


def test_accelerations_do_not_depend_on_thread_count():
    # README.md:57-59: accelerations bit-identical across thread counts; V only
    # to rounding (its grouping follows the splits).
    ph, st = small_mixture(n=1728, seed=7)
    one = O.forces(ph, st, nthreads=1)
    four = O.forces(ph, st, nthreads=4)
    for a in ("ax", "ay", "az"):
        assert np.array_equal(one[a], four[a])
    assert four["V"] == pytest.approx(one["V"], rel=1e-12)


def test_per_atom_potential_sums_to_total():
    ph, st = small_mixture(n=512, seed=9)
    out = O.forces(ph, st)
    assert out["pe"].sum() == pytest.approx(out["V"] + out["Vb"], rel=1e-12)


def test_plain_lj_workload_geometry():
    w = plain_lj(n_side=16, steps=1)
    _, _, cut = O.lj_table(w.physics)
    assert cut["lj_rc"] == pytest.approx(8.5)
    assert w.physics.grid_dims() == (6, 6, 6)
    assert w.state.n == 4096
    out = O.forces(w.physics, w.state)
    assert out["V"] < 0.0


NA, OX = SPECIES_INDEX["Na"], SPECIES_INDEX["O"]


def _coulomb_pair(ph, d):
    """Na (+1) at the origin, O (-1) at +d on x: the scalar force along the
    normal from O to Na (positive = repulsive) and V."""
    st = State(np.array([0.0, d]), np.zeros(2), np.zeros(2), np.zeros(2), np.zeros(2), np.zeros(2),
               np.array([NA, OX], dtype=np.uint8), np.array([1.0, -1.0]))
    out = O.forces(ph, st)
    f0 = -out["ax"][0] * ph.mass[NA]
    f1 = out["ax"][1] * ph.mass[OX]
    assert f0 == pytest.approx(f1, rel=1e-15, abs=1e-300)  # Newton's third law
    return f0, out["V"]


def test_coulomb_clamp_is_half_a_length_unit_absolute():
    # src/systems/forces.rs:282-289: r = max(d, 0.5) in absolute units, so for
    # k_e = 1 and q = +-1 the pair feels -1/0.25 = -4 and V = -1/0.5 = -2 at
    # every d <= 0.5 (the switch is 1 this far inside the cutoff), and the
    # plain law above.  LJ is off (epsilon = 0).  A clamp at 0.5 sigma_ij
    # (0.9125 for Na-O) would give -1/0.9125^2 at d = 0.51.
    ph = Physics(box=(40.0, 40.0, 40.0), enable_electrostatics=True, coulomb_constant=1.0,
                 epsilon_k=(0.0,) * 10)
    for d in (0.2, 0.49, 0.5):
        f, v = _coulomb_pair(ph, d)
        assert v == -2.0
        assert f == pytest.approx(-4.0, rel=1e-15)
    f, v = _coulomb_pair(ph, 0.51)
    assert v == pytest.approx(-1.0 / 0.51, rel=1e-15)
    assert f == pytest.approx(-1.0 / (0.51 * 0.51), rel=1e-15)


def test_coulomb_switch_is_evaluated_at_the_unclamped_distance():
    # A table of sigma 0.2 and coulomb_cutoff 2.7 puts the Coulomb switch on
    # [0.486, 0.54), straddling the clamp.  With V(d) = (q/r) sw(d) and
    # F = (q/r^2) sw(d) - (q/r) sw'(d), r = max(d, 0.5):
    #   d < 0.486:        F = -4, V = -2 (switch 1 at d; at r = 0.5 it is not)
    #   0.486 < d < 0.5:  F + V'(d) = V / 0.5 (the clamp breaks F = -V')
    #   0.5 < d < 0.54:   F = -V'(d) (the switched law is conservative)
    # V' by central differences.
    ph = Physics(box=(3.0, 3.0, 3.0), enable_electrostatics=True, coulomb_constant=1.0, coulomb_cutoff=2.7,
                 sigma=(0.2,) * 10, epsilon_k=(0.0,) * 10)
    _, _, cut = O.lj_table(ph)
    assert cut["coulomb_rc"] == pytest.approx(0.54) and cut["coulomb_ron"] == pytest.approx(0.486)
    for d in (0.2, 0.3, 0.45, 0.48):
        f, v = _coulomb_pair(ph, d)
        assert (f, v) == (pytest.approx(-4.0, rel=1e-15), -2.0)
    h = 1e-7
    for d in (0.489, 0.492, 0.495, 0.498, 0.505, 0.51, 0.52, 0.53, 0.535):
        f, v = _coulomb_pair(ph, d)
        dv = (_coulomb_pair(ph, d + h)[1] - _coulomb_pair(ph, d - h)[1]) / (2 * h)
        assert -2.0 < v < 0.0
        expect = v / 0.5 if d < 0.5 else 0.0
        assert f + dv == pytest.approx(expect, rel=1e-6, abs=1e-6), d
    assert _coulomb_pair(ph, 0.541) == (0.0, 0.0)
