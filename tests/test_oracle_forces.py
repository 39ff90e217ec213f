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
