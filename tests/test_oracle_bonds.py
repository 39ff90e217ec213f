"""The oracle's bond rows: the bond term of the force pass
(src/systems/forces.rs:305-317), the bound-pair search and the PairTracker
(src/analysis/pairs.rs:62-174) and BondObservationSystem
(src/systems/bonds.rs:73-185).  Pinned by the reference's own unit tests
(re-run here on the oracle), brute force, closed forms and finite
differences."""
import math

import numpy as np
import pytest

from genesis_amd.inputs import State, random_gas
from genesis_amd.params import DEBOUNCE, GRID_MINIMAGE, Physics, SPECIES_INDEX
from oracle import oracle as O

H, NA, O_, SI = SPECIES_INDEX["H"], SPECIES_INDEX["Na"], SPECIES_INDEX["O"], SPECIES_INDEX["Si"]


def two(pos_a, pos_b, ta, tb):
    x = np.array([pos_a[0], pos_b[0]], dtype=float)
    y = np.array([pos_a[1], pos_b[1]], dtype=float)
    z = np.array([pos_a[2], pos_b[2]], dtype=float)
    return x, y, z, np.array([ta, tb], dtype=np.uint8)


# ---- PairTracker: src/analysis/pairs.rs:218-253 ------------------------------

def test_tracker_counts_episodes_and_debounce():
    p = (1, 2)
    # debounce 1: five bound ticks, one absent tick closes the episode
    assert O.track([[p]] * 5 + [[]], debounce=1) == {}
    # a rebound pair starts a new episode: three ticks
    assert O.track([[p]] * 5 + [[]] + [[p]] * 3, debounce=1) == {p: 3}
    # debounce 3: one absent tick keeps the episode open with its count
    assert O.track([[p]] * 5 + [[]], debounce=3) == {p: 5}
    assert O.track([[p]] * 5 + [[]] * 2, debounce=3) == {p: 5}
    assert O.track([[p]] * 5 + [[]] * 3, debounce=3) == {}
    # the engine's debounce (2 ticks): a one-tick gap does not reset the count
    assert O.track([[p]] * 4 + [[]] + [[p]] * 2, debounce=DEBOUNCE) == {p: 6}


# ---- bound pairs: src/analysis/pairs.rs:195-216 ------------------------------

def test_bound_pairs_respect_the_per_pair_threshold():
    ph = Physics(box=(64.0, 64.0, 64.0))
    # H-H mixed sigma 1.6: 2.0 < 1.5 * 1.6 = 2.4 is bound
    assert len(O.bound_pairs(ph, *two((0, 0, 0), (2.0, 0, 0), H, H))) == 1
    # 3.3 is bound for Si-Si (threshold 3.45) but not for H-H (2.4)
    assert len(O.bound_pairs(ph, *two((0, 0, 0), (3.3, 0, 0), SI, SI))) == 1
    assert len(O.bound_pairs(ph, *two((0, 0, 0), (3.3, 0, 0), H, H))) == 0


def brute_bound(ph, x, y, z, sp):
    """The definition written out: |min_image(a - b)| < k_bind * (sigma_a + sigma_b) / 2, d > eps."""
    L = np.asarray(ph.box)
    sig = np.asarray(ph.sigma)
    cut = ph.bond_k_bind * max(ph.sigma)
    out = set()
    n = len(x)
    for i in range(n):
        d = np.stack([x[i] - x, y[i] - y, z[i] - z], 1)
        r = d / L
        d = d - L * np.where(r >= 0, np.floor(r + 0.5), np.ceil(r - 0.5))
        d2 = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]
        dd = np.sqrt(d2)
        thr = ph.bond_k_bind * (0.5 * (sig[sp[i]] + sig[sp]))
        for j in np.nonzero((d2 < cut * cut) & (dd > np.finfo(float).eps) & (dd < thr))[0]:
            if j > i:
                out.add((i, int(j)))
    return out


@pytest.mark.parametrize("L,semantics", [(30.0, 0), (7.5, 0), (7.5, GRID_MINIMAGE), (4.0, 0)])
def test_bound_pairs_equal_brute_force(L, semantics):
    # includes boxes with 2 cells per axis (the grid reports those pairs
    # twice; the tracker's set makes them once) and 1 cell
    n = 300 if L > 10 else 40
    st = random_gas(seed=int(L * 10) + semantics, n=n, box=(L, L, L), species_pool=(H, O_, NA, SI))
    ph = Physics(box=(L, L, L), grid_semantics=semantics)
    got = {tuple(map(int, r)) for r in O.bound_pairs(ph, st.x, st.y, st.z, st.species)}
    assert got == brute_bound(ph, st.x, st.y, st.z, st.species)
    assert len(got) > 0


def test_persistence_threshold_closed_form():
    ph = Physics()
    for a, b in ((H, H), (NA, O_), (SI, H)):
        mu = ph.mass[a] * ph.mass[b] / (ph.mass[a] + ph.mass[b])
        eps = ph.thermal_constant * math.sqrt(ph.epsilon_k[a] * ph.epsilon_k[b])
        sig = 0.5 * (ph.sigma[a] + ph.sigma[b])
        k = 72.0 * 2.0 ** (-1.0 / 3.0) * eps / (sig * sig)
        want = math.ceil(ph.bond_min_periods * 2.0 * math.pi * math.sqrt(mu / k) / ph.dt)
        assert O.persistence_ticks(ph, a, b) == want


# ---- BondObservationSystem: src/systems/bonds.rs:263-343 ---------------------

def test_lifetime_and_species_matrix_are_measured():
    ph = Physics()
    thr = O.persistence_ticks(ph, H, H)
    bound = ((0.0, 1.0), (0.0, 0.0), (0.0, 0.0))
    apart = ((0.0, 6.0), (0.0, 0.0), (0.0, 0.0))
    frames = [tuple(np.array(c) for c in bound)] * (thr + 500) + [tuple(np.array(c) for c in apart)] * (DEBOUNCE + 2)
    bs = O.observe_frames(ph, frames, np.array([H, H], dtype=np.uint8))
    s = bs[thr + 499]
    assert s[0] == thr + 500
    assert s[1] == 1 and s[3] == 0, "bound pair persistent, not yet formed"
    assert s[6 + H * 10 + H] == 2, "H-H counted symmetrically"
    assert bs[thr - 2][1] == 0 and bs[thr - 1][1] == 1, "persistent from the threshold tick on"
    e = bs[-1]
    assert e[1] == 0 and e[3] == 1
    # persistent from tick thr; the debounce closes it at tick thr + 502
    assert e[4] == 502.0 >= 400.0
    assert bs[thr + 500][1] == 1, "one missing tick keeps the bond"


def test_observed_bonds_feed_the_next_force_pass():
    # A cold Na-O dimer at its LJ minimum stays bound: after the persistence
    # threshold the Bonds adjacency holds the pair and, with the interaction
    # on, the next ticks carry a bond energy.  The final state's force pass
    # with that adjacency reproduces the last tick's accelerations.
    sig = 0.5 * (Physics().sigma[NA] + Physics().sigma[O_])
    r0 = sig * 2.0 ** (1.0 / 6.0) * 1.02
    ph = Physics(box=(30.0, 30.0, 30.0), dt=0.002, enable_bond_observation=True, enable_bond_interaction=True,
                 bond_min_periods=0.5)
    thr = O.persistence_ticks(ph, NA, O_)
    st = State(np.array([0.0, r0]), np.zeros(2), np.zeros(2), np.zeros(2), np.zeros(2), np.zeros(2),
               np.array([NA, O_], dtype=np.uint8), np.zeros(2))
    n = thr + 20
    fin, obs, bs, (ptr, idx) = O.run_observed(ph, st, n)
    assert bs[thr - 2][1] == 0 and bs[thr - 1][1] == 1
    assert list(ptr) == [0, 1, 2] and list(idx) == [1, 0]
    assert np.all(obs[: thr, 2] == 0.0) and np.all(obs[thr + 1:, 2] != 0.0)
    f = O.forces(ph, fin, bonds_csr=(ptr, idx))
    assert np.array_equal(f["ax"], fin.ax)


def test_bond_term_is_minus_gradient_of_total_potential():
    # With a fixed Bonds adjacency the force pass is conservative in
    # V + V_bond: central differences of the total against the accelerations.
    ph = Physics(box=(20.0, 20.0, 20.0), enable_bond_interaction=True)
    rng = np.random.default_rng(5)
    n = 6
    # a line of atoms 2.0 apart (outside every hardened core, where F is not
    # -dV/dr), bonds 0-1 and 4-5 inside their binding radius, 0-3 beyond it
    x = 2.0 * np.arange(n) + rng.uniform(-0.1, 0.1, n)
    y = rng.uniform(-0.2, 0.2, n)
    z = rng.uniform(-0.2, 0.2, n)
    sp = np.array([NA, O_] * 3, dtype=np.uint8)
    st = State(x, y, z, np.zeros(n), np.zeros(n), np.zeros(n), sp, np.zeros(n))
    csr = (np.array([0, 2, 3, 3, 4, 5, 6], dtype=np.int32), np.array([1, 3, 0, 0, 5, 4], dtype=np.int32))
    f = O.forces(ph, st, bonds_csr=csr)
    assert f["Vb"] != 0.0
    m = np.asarray(ph.mass)[sp]
    h = 1e-6
    for i in range(n):
        for ax, arr in (("ax", "x"), ("ay", "y"), ("az", "z")):
            s1, s2 = st.copy(), st.copy()
            getattr(s1, arr)[i] += h
            getattr(s2, arr)[i] -= h
            f1, f2 = O.forces(ph, s1, bonds_csr=csr), O.forces(ph, s2, bonds_csr=csr)
            g = ((f1["V"] + f1["Vb"]) - (f2["V"] + f2["Vb"])) / (2 * h)
            assert f[ax][i] * m[i] == pytest.approx(-g, rel=1e-5, abs=1e-6)
