"""The oracle's tick (Universe::tick with build_schedule order) against the
reference's integration tests: NVE energy and momentum conservation
(src/systems/forces.rs:368-431) and Berendsen convergence
(src/systems/thermostat.rs:94-116), plus the closed-form lambda."""
import math

import numpy as np
import pytest

from genesis_amd.inputs import State
from genesis_amd.params import Physics
from oracle import oracle as O
from tests.helpers import reference_like


def momentum(st, ph):
    m = np.asarray(ph.mass)[st.species]
    return np.array([(m * st.vx).sum(), (m * st.vy).sum(), (m * st.vz).sum()])


def test_energy_and_momentum_conserve_with_forces():
    # 512 atoms (8^3 lattice), 24^3 box, 100 K, default table, dt = 1/60.
    ph = Physics(box=(24.0, 24.0, 24.0))
    st = reference_like(8, 24.0, 100.0, seed=42)
    p0 = momentum(st, ph)
    st, obs = O.run(ph, st, 300)
    e_ref = obs[-1, 5]
    assert e_ref > 0.0
    assert obs[-1, 1] < 0.0
    st, obs = O.run(ph, st, 2000)
    rel = abs(obs[-1, 5] - e_ref) / abs(e_ref)
    assert rel < 1e-3, rel
    dp = np.linalg.norm(momentum(st, ph) - p0)
    assert dp < 1e-6 * max(np.linalg.norm(p0), 1.0)


def test_energy_conserves_with_electrostatics_and_gravity():
    ph = Physics(box=(24.0, 24.0, 24.0), enable_electrostatics=True, enable_gravity=True, gravity_constant=1e-3)
    st = reference_like(8, 24.0, 100.0, seed=43)
    p0n = np.linalg.norm(momentum(st, ph))
    st, obs = O.run(ph, st, 300)
    e_ref = obs[-1, 5]
    st, obs = O.run(ph, st, 2000)
    rel = abs(obs[-1, 5] - e_ref) / abs(e_ref)
    assert rel < 1e-3, rel
    assert abs(np.linalg.norm(momentum(st, ph)) - p0n) < 1e-6 * max(p0n, 1.0)


@pytest.mark.parametrize("target,start", [(60.0, 300.0), (180.0, 20.0)])
def test_thermostat_drives_temperature_to_target(target, start):
    ph = Physics(box=(24.0, 24.0, 24.0), enable_thermostat=True, thermostat_temperature=target,
                 thermostat_tau=10.0)
    st = reference_like(8, 24.0, start, seed=44)
    _, obs = O.run(ph, st, 2500)
    assert abs(obs[-1, 3] - target) < 10.0, obs[-1, 3]


def test_thermostat_lambda_closed_form():
    ph = Physics(box=(24.0, 24.0, 24.0), thermostat_temperature=80.0, thermostat_tau=20.0)
    st = reference_like(4, 24.0, 300.0, seed=1)
    k = O.kinetic(st, ph.mass)
    t = (2.0 / 3.0) * (k / st.n) / ph.thermal_constant
    v0 = st.vx.copy()
    lam = O.thermostat(ph, st)
    assert lam == pytest.approx(math.sqrt(1.0 + (80.0 / t - 1.0) / 20.0), rel=1e-15)
    np.testing.assert_array_equal(st.vx, v0 * lam)


def test_integrator_steps_are_the_reference_expressions():
    ph = Physics(box=(10.0, 10.0, 10.0))
    st = reference_like(3, 10.0, 100.0, seed=2)
    st.ax, st.ay, st.az = st.vx * 3.0, st.vy * -2.0, st.vz * 0.5
    v0, x0 = st.vx.copy(), st.x.copy()
    O.half_kick(st, 0.25)
    np.testing.assert_array_equal(st.vx, v0 + st.ax * (0.5 * 0.25))
    O.drift(st, 0.25)
    np.testing.assert_array_equal(st.x, x0 + st.vx * 0.25)
    st.x += 13.7
    O.wrap(st, ph.box)
    assert (st.x >= -5.0).all() and (st.x <= 5.0).all()


def test_wrap_is_a_half_box_translation_modulo_the_box():
    # BoundarySystem maps v to v.rem_euclid(2h) - h (src/systems/boundary.rs:30),
    # which is mod(v, L) - L/2: every coordinate moves by half a box (mod L),
    # so the origin goes to -L/2.  A wrap into [-L/2, L/2) without that
    # translation would map 0 to 0 and fail here.
    L = np.array([10.0, 7.5, 3.0])
    vals = np.array([0.0, -0.0, 1.0, -1.0, 4.999, -4.999, 5.0, -5.0, 10.0, -10.0, 20.0, -20.0, 30.0, 12.5, -12.5,
                     -1e-300, 1e-300, 7.25, -7.25, 1e6 + 0.5, -1e6 - 0.5, 2.5, -2.5])
    n = len(vals)
    z = np.zeros(n)
    st = State(vals.copy(), vals.copy(), vals.copy(), z, z, z, np.zeros(n, dtype=np.uint8), z)
    O.wrap(st, tuple(L))
    for got, Lk in ((st.x, L[0]), (st.y, L[1]), (st.z, L[2])):
        expect = np.mod(vals, Lk) - 0.5 * Lk
        np.testing.assert_array_equal(got, expect)
    assert st.x[0] == -5.0 and st.x[2] == -4.0 and st.x[3] == 4.0 and st.x[8] == -5.0
    # -1e-300 mod L rounds to L itself: the reference's result is then +L/2
    assert st.x[15] == 5.0


def test_wrap_moves_the_whole_universe_rigidly():
    # relative positions (minimum image) are unchanged by the wrap, absolute
    # ones move by L/2 modulo L
    ph = Physics(box=(24.0, 24.0, 24.0))
    st = reference_like(5, 24.0, 100.0, seed=9)
    x0, y0 = st.x.copy(), st.y.copy()
    O.wrap(st, ph.box)
    d_before = x0[:, None] - x0[None, :]
    d_after = st.x[:, None] - st.x[None, :]
    img = lambda d: d - 24.0 * np.round(d / 24.0)
    np.testing.assert_allclose(img(d_after), img(d_before), rtol=0, atol=1e-12)
    shift = np.mod(st.y - y0, 24.0)
    np.testing.assert_allclose(np.minimum(np.abs(shift - 12.0), np.abs(shift + 12.0)), 0.0, atol=1e-12)
