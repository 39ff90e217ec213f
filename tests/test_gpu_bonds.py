"""Observed bonds on the GPU (gs_bond_*, the bond term of the force pass)
against the oracle (oracle/oracle.cpp, pinned by tests/test_oracle_bonds.py).

Bound pairs, tracker counts, persistence, lifetimes, the species matrix,
the Bonds adjacency and its components are integer results: bit-exact.
Forces and energies carry the force-pass tolerance (1e-10, conditioned)."""
import numpy as np
import pytest

from genesis_amd.inputs import State, ionic_melt, random_gas
from genesis_amd.params import DEBOUNCE, GRID_MINIMAGE, Physics, SPECIES_INDEX
from oracle import oracle as O
from tests.helpers import force_tolerance_ok

pytestmark = pytest.mark.gpu

H, NA, O_, SI = SPECIES_INDEX["H"], SPECIES_INDEX["Na"], SPECIES_INDEX["O"], SPECIES_INDEX["Si"]
REL = 1e-10


def say(msg):
    print(f"[gpu-bonds] {msg}", flush=True)


@pytest.fixture(scope="module")
def GU():
    from genesis_amd.universe import GpuUniverse
    return GpuUniverse


def melt(n_side, seed=1105, **kw):
    # bond_min_periods is clamped at 0.1 periods (src/systems/bonds.rs:52):
    # about 116 ticks for Na-O at dt = 0.5 fs
    w = ionic_melt(seed=seed, nx=n_side, ny=n_side, nz=n_side, steps=60)
    ph = Physics(**{**w.physics.__dict__, "bond_min_periods": 0.1, **kw})
    return ph, w.state


def frames_of(ph, st, nframes, jitter, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nframes):
        out.append(tuple(np.asarray(a, dtype=float) + rng.normal(0.0, jitter, st.n) for a in (st.x, st.y, st.z)))
    return out


@pytest.mark.parametrize("L,semantics", [(30.0, 0), (9.0, 0), (9.0, GRID_MINIMAGE), (5.0, 0)])
def test_bound_pairs_bit_exact(GU, L, semantics):
    n = 2000 if L > 10 else 25
    st = random_gas(seed=int(L * 7) + semantics, n=n, box=(L, L, L), species_pool=(H, O_, NA, SI))
    ph = Physics(box=(L, L, L), grid_semantics=semantics)
    say(f"bound pairs n={n} L={L} semantics={semantics}")
    with GU(ph, st) as u:
        got = u.bound_pairs()
    ref = O.bound_pairs(ph, st.x, st.y, st.z, st.species)
    assert len(ref) > 0
    assert np.array_equal(got, ref)


def test_bound_pairs_bit_exact_on_a_large_aux_grid(GU):
    # aux cells of edge 1.5 max sigma in a 100 A box: 28^3 > 16384 cells, so
    # the lens-grid scan runs in several workgroups
    L = 100.0
    st = random_gas(seed=91, n=40000, box=(L, L, L), species_pool=(H, O_, NA, SI))
    ph = Physics(box=(L, L, L))
    say(f"bound pairs on a large aux grid n={st.n} L={L}")
    with GU(ph, st) as u:
        got = u.bound_pairs()
    ref = O.bound_pairs(ph, st.x, st.y, st.z, st.species)
    assert len(ref) > 1000
    assert np.array_equal(got, ref)


def test_lifetime_and_species_matrix_fixture(GU):
    # the reference's own observation test (src/systems/bonds.rs:263-343),
    # frame by frame through upload + observe
    ph = Physics()
    thr = O.persistence_ticks(ph, H, H)
    sp = np.array([H, H], dtype=np.uint8)
    z2 = np.zeros(2)
    bound = State(np.array([0.0, 1.0]), z2, z2, z2, z2, z2, sp, z2)
    apart = State(np.array([0.0, 6.0]), z2, z2, z2, z2, z2, sp, z2)
    say(f"H-H lifetime fixture, threshold {thr} ticks")
    with GU(ph, bound) as u:
        for t in range(1, thr + 501):
            u.bond_observe(t)
        s = u.bond_stats()
        assert s["bonded_pairs"] == 1 and s["bonds_formed"] == 0 and s["species_matrix"][H, H] == 2
        ptr, idx = u.bonds()
        assert list(ptr) == [0, 1, 2] and list(idx) == [1, 0]
        u.upload(apart)
        for t in range(thr + 501, thr + 501 + DEBOUNCE + 2):
            u.bond_observe(t)
        e = u.bond_stats()
    assert e["bonded_pairs"] == 0 and e["bonds_formed"] == 1 and e["lifetime_sum_ticks"] == 502.0


@pytest.mark.parametrize("n_side,jitter", [(8, 0.12), (12, 0.08)])
def test_observation_frames_match_oracle(GU, n_side, jitter):
    # a jittered rock salt: pairs flicker across their thresholds, so episodes
    # open, survive one-tick gaps, close at the debounce, become persistent
    # and break; every tick's summary and the final adjacency are integers
    # dt only enters the thresholds here: 0.05 gives about 11 ticks
    ph, st = melt(n_side, dt=0.05)
    frames = frames_of(ph, st, 40, jitter, seed=n_side)
    ref, (rptr, ridx) = O.observe_frames(ph, frames, st.species, with_bonds=True)
    say(f"observation frames n={st.n}, thresholds {O.persistence_ticks(ph, NA, O_)} ticks")
    with GU(ph, st) as u:
        for f, (x, y, z) in enumerate(frames):
            u.upload(State(x, y, z, st.vx, st.vy, st.vz, st.species, st.charge))
            u.bond_observe(f + 1)
            s = u.bond_stats()
            row = ref[f]
            got = [s["tick"], s["bonded_pairs"], s["bonded_entities"], s["bonds_formed"], s["lifetime_sum_ticks"],
                   s["mean_coordination"]]
            assert got == list(row[:6]), (f, got, row[:6])
            assert np.array_equal(s["species_matrix"].ravel(), row[6:].astype(np.int64))
        ptr, idx = u.bonds()
    assert ref[-1][1] > 0 and ref[-1][3] > 0, "bonds formed and broke during the frames"
    assert np.array_equal(ptr, rptr) and np.array_equal(idx, ridx)


def test_bond_force_term_matches_oracle(GU):
    ph, st = melt(12, enable_bond_interaction=True)
    # bonds: every Na-O pair closer than 2.3 among the first 400 atoms
    pairs = O.bound_pairs(ph, st.x, st.y, st.z, st.species)
    keep = [(a, b) for a, b in pairs if a < 400 and b < 400 and st.species[a] != st.species[b]]
    adj = [[] for _ in range(st.n)]
    for a, b in keep:
        adj[a].append(b)
        adj[b].append(a)
    ptr = np.zeros(st.n + 1, dtype=np.int32)
    ptr[1:] = np.cumsum([len(x) for x in adj])
    idx = np.array([m for x in adj for m in sorted(x)], dtype=np.int32)
    say(f"bond force term: {len(keep)} bonds, n={st.n}")
    ref = O.forces(ph, st, bonds_csr=(ptr, idx))
    with GU(ph, st) as u:
        u.upload_bonds(ptr, idx)
        V, Vb = u.forces()
        got = u.download()
    m = np.asarray(ph.mass)[st.species]
    ok, worst = force_tolerance_ok(got.ax, got.ay, got.az, ref, m, REL)
    assert ok, worst
    assert V == pytest.approx(ref["V"], rel=REL) and Vb == pytest.approx(ref["Vb"], rel=REL) and Vb != 0.0


@pytest.mark.parametrize("n_side,mode,lanes,aux", [(8, 0, 0, False), (12, 0, 0, False), (12, 1, 4, False),
                                                    (12, 1, 2, False), (12, 1, 2, True)])
def test_bonded_trajectory_matches_oracle(GU, n_side, mode, lanes, aux):
    # observation and interaction on, thresholds of ~116 ticks: bonds appear
    # late in the run and their springs act from the next tick on.  The list
    # path at 2 and 4 lanes per atom finds the bound pairs inside the pair
    # pass (the ionic melt's bench setting is 2 lanes); aux=True forces the
    # aux-grid search instead.
    ph, st = melt(n_side)
    nsteps = 150
    ref, obs_ref, bs_ref, (rptr, ridx) = O.run_observed(ph, st, nsteps, nthreads=8)
    say(f"bonded trajectory n={st.n} dims={ph.grid_dims()} mode={mode} lanes={lanes} aux={aux}")
    with GU(ph, st) as u:
        if mode == 1:
            u.set_tuning(1, lanes)
        u.set_bound_search(aux)
        obs = u.step(nsteps)
        got = u.download()
        s = u.bond_stats()
        ptr, idx = u.bonds()
    assert bs_ref[-1][1] > 0 and np.any(obs_ref[:, 2] != 0.0)
    np.testing.assert_allclose(obs[:, 2], obs_ref[:, 2], rtol=1e-8, atol=1e-12)
    np.testing.assert_allclose(obs[:, 5], obs_ref[:, 5], rtol=1e-9)
    np.testing.assert_allclose(got.x, ref.x, rtol=0, atol=1e-8)
    assert [s["tick"], s["bonded_pairs"], s["bonded_entities"], s["bonds_formed"]] == list(bs_ref[-1][:4])
    assert np.array_equal(ptr, rptr) and np.array_equal(idx, ridx)


def test_bond_components_fixtures(GU):
    # src/systems/bond_structure.rs:265-327: two Na-O dimers and a monomer,
    # and the 4-ring; formulas as src/analysis/mod.rs:589-605
    sp = np.array([NA, O_, NA, O_, NA], dtype=np.uint8)
    x = np.array([0.0, 2.0, 5.0, 7.0, 10.0])
    z5 = np.zeros(5)
    st = State(x, z5, z5, z5, z5, z5, sp, z5)
    ph = Physics(box=(40.0, 40.0, 40.0))
    with GU(ph, st) as u:
        u.upload_bonds(np.array([0, 1, 2, 3, 4, 4], dtype=np.int32), np.array([1, 0, 3, 2], dtype=np.int32))
        dimers = u.bond_components(sp)
        u.upload_bonds(np.array([0, 2, 4, 6, 8, 8], dtype=np.int32),
                       np.array([1, 3, 0, 2, 1, 3, 2, 0], dtype=np.int32))
        ring = u.bond_components(sp)
        labels = u.bond_labels()
    assert [(c["size"], c["formula"], c["members"]) for c in dimers] == [(2, "Na-O", [0, 1]), (2, "Na-O", [2, 3])]
    assert [(c["size"], c["formula"], c["members"]) for c in ring] == [(4, "Na2-O2", [0, 1, 2, 3])]
    assert list(labels) == [0, 0, 0, 0, 0xFFFFFFFF]
    ref = O.bond_components([(e, int(sp[e])) for e in range(4)], [(0, 1), (1, 2), (2, 3), (3, 0)])
    assert [(c["size"], c["formula"], c["members"]) for c in ref] == [(4, "Na2-O2", [0, 1, 2, 3])]


def test_bond_components_of_an_observed_melt_match_oracle(GU):
    # a dilute Na/O gas: a few hundred bonded components of many sizes and
    # formulas (ties in size exercise the formula and member ordering)
    st = random_gas(seed=77, n=1200, box=(40.0, 40.0, 40.0), species_pool=(NA, O_))
    ph = Physics(box=(40.0, 40.0, 40.0), dt=0.05, bond_min_periods=0.1)
    frames = frames_of(ph, st, 20, 0.05, seed=3)
    _, (rptr, ridx) = O.observe_frames(ph, frames, st.species, with_bonds=True)
    ents = [(e, int(st.species[e])) for e in range(st.n) if rptr[e + 1] > rptr[e]]
    edges = [(e, int(m)) for e in range(st.n) for m in ridx[rptr[e]:rptr[e + 1]]]
    ref = O.bond_components(ents, edges)
    with GU(ph, st) as u:
        for f, (x, y, z) in enumerate(frames):
            u.upload(State(x, y, z, st.vx, st.vy, st.vz, st.species, st.charge))
            u.bond_observe(f + 1)
        got = u.bond_components(st.species)
    assert len(ref) > 100
    assert [(c["size"], c["formula"], c["members"]) for c in got] == \
        [(c["size"], c["formula"], c["members"]) for c in ref]


def test_replayed_call_restores_the_bond_state(GU):
    # A list skin so large that the first call's lists overflow their
    # capacity and its stencils the build's LDS: the call is replayed from
    # its snapshot (on the cell tiles), and the bond tracker, adjacency,
    # counters and tick return to the call's start with the positions.  The
    # observed bonds then match the oracle's run.
    ph, st = melt(16)
    nsteps = 150
    ref, obs_ref, bs_ref, (rptr, ridx) = O.run_observed(ph, st, nsteps, nthreads=8)
    say(f"replayed call with bonds n={st.n}")
    with GU(ph, st) as u:
        u.set_tuning(1, 4, skin=0.65 * ph.force_rc())
        obs = u.step(nsteps)
        tun = u.tuning()
        got = u.download()
        s = u.bond_stats()
        ptr, idx = u.bonds()
    assert tun["list_replays"] >= 1
    assert bs_ref[-1][1] > 0
    np.testing.assert_allclose(obs[:, 5], obs_ref[:, 5], rtol=1e-9)
    np.testing.assert_allclose(got.x, ref.x, rtol=0, atol=1e-8)
    assert [s["tick"], s["bonded_pairs"], s["bonded_entities"], s["bonds_formed"]] == list(bs_ref[-1][:4])
    assert np.array_equal(ptr, rptr) and np.array_equal(idx, ridx)


@pytest.mark.parametrize("lanes", [2, 4])
def test_fused_bound_search_matches_the_aux_grid_search(GU, lanes):
    # The list pass screens its entries against the bind cutoff and applies
    # the exact test to the hits; the aux-grid search scans the grid.  With
    # the bond interaction on, any difference in the bound pairs would move
    # the trajectory, so everything must agree bit for bit.
    ph, st = melt(12)
    out = []
    for aux in (True, False):
        with GU(ph, st) as u:
            u.set_tuning(1, lanes)
            u.set_bound_search(aux)
            obs = np.concatenate([u.step(70), u.step(80)])
            out.append((obs, u.download(), u.bond_stats(), u.bonds(), u.tuning()))
    (o1, s1, b1, (p1, i1), t1), (o2, s2, b2, (p2, i2), t2) = out
    say(f"fused vs aux-grid bound search, {lanes} lanes: bonded pairs {b1['bonded_pairs']}")
    assert b1["bonded_pairs"] > 0 and t2["list_replays"] == 0
    assert np.array_equal(o1, o2)
    for f in ("x", "y", "z", "vx", "vy", "vz"):
        assert np.array_equal(getattr(s1, f), getattr(s2, f)), f
    for k in ("tick", "bonded_pairs", "bonded_entities", "bonds_formed", "lifetime_sum_ticks"):
        assert b1[k] == b2[k], k
    assert np.array_equal(b1["species_matrix"], b2["species_matrix"])
    assert np.array_equal(p1, p2) and np.array_equal(i1, i2)


def _bound_partner_counts(ph, p, species):
    """Bound partners of every atom by brute force (d < k_bind sigma_ab,
    src/analysis/pairs.rs:132-174): all partners, and those of larger id
    (the ones an atom owns)."""
    sig = np.asarray(ph.sigma)
    d = p[:, None, :] - p[None, :, :]
    d -= np.asarray(ph.box) * np.round(d / np.asarray(ph.box))
    r = np.sqrt((d * d).sum(-1))
    thr = ph.bond_k_bind * 0.5 * (sig[species][:, None] + sig[species][None, :])
    b = (r < thr) & (r > 0)
    return b.sum(1), np.triu(b, 1).sum(1)


def _shell_cluster(ph, n_shell=40, radius=2.3, seed=7):
    # n_shell atoms on a sphere around one centre atom that has the largest id
    # (it owns none of its pairs), plus a dilute gas far away
    rng = np.random.default_rng(seed)
    k = np.arange(n_shell) + 0.5
    th, phi = np.arccos(1.0 - 2.0 * k / n_shell), np.pi * (1.0 + 5.0 ** 0.5) * k
    shell = radius * np.stack([np.sin(th) * np.cos(phi), np.sin(th) * np.sin(phi), np.cos(th)], 1)
    gas = rng.uniform(-20.0, 20.0, (400, 3))
    gas = gas[np.linalg.norm(gas, axis=1) > 8.0]
    p = np.concatenate([shell, gas, np.zeros((1, 3))])
    n = len(p)
    species = np.where(np.arange(n) % 2 == 0, 5, 4).astype(np.uint8)  # Na, O
    species[-1] = 5
    charge = np.where(species == 5, 1.0, -1.0)
    z = np.zeros(n)
    return p, State(p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), z, z.copy(), z.copy(), species, charge)


def test_dense_shell_overflows_the_fused_search_and_replays(GU):
    # An atom with 40 bound partners: with 2 lanes per atom one lane screens
    # more candidates than its 16-entry LDS column holds, so the call is
    # replayed with the aux-grid search.  No atom owns more than 32 partners
    # (the centre has the largest id), so nothing is truncated: the result
    # must equal a run that used the aux grid throughout and the oracle's.
    # At rest and with a tiny dt nothing moves, so the overflow is the only
    # reason for a replay.
    base, _ = melt(8)
    ph = Physics(**{**base.__dict__, "box": (40.0, 40.0, 40.0), "dt": 1e-7})
    p, st = _shell_cluster(ph)
    allp, owned = _bound_partner_counts(ph, p, st.species)
    assert allp.max() >= 40 and owned.max() <= 32
    nsteps = 2
    ref, obs_ref, bs_ref, (rptr, ridx) = O.run_observed(ph, st, nsteps, nthreads=8)
    out = []
    for aux in (True, False):
        with GU(ph, st) as u:
            u.set_tuning(1, 2)
            u.set_bound_search(aux)
            obs = u.step(nsteps)
            out.append((obs, u.download(), u.bond_stats(), u.tuning(), u.bound_pairs()))
    (o1, s1, b1, t1, bp1), (o2, s2, b2, t2, bp2) = out
    say(f"dense shell n={st.n}: replays aux {t1['list_replays']}, fused {t2['list_replays']}")
    assert t2["pair_mode"] == 1 and t1["list_replays"] == 0 and t2["list_replays"] == 1
    assert np.array_equal(o1, o2)
    for f in ("x", "y", "z", "vx", "vy", "vz"):
        assert np.array_equal(getattr(s1, f), getattr(s2, f)), f
    assert b1["tick"] == b2["tick"] and np.array_equal(b1["species_matrix"], b2["species_matrix"])
    # against the oracle
    np.testing.assert_allclose(o2[:, 5], obs_ref[:, 5], rtol=1e-9)
    np.testing.assert_allclose(s2.x, ref.x, rtol=0, atol=1e-12)
    assert [b2["tick"], b2["bonded_pairs"], b2["bonded_entities"], b2["bonds_formed"]] == list(bs_ref[-1][:4])
    refp = O.bound_pairs(ph, s2.x, s2.y, s2.z, st.species)
    assert np.array_equal(bp2, refp) and len(refp) >= 40 + 200


@pytest.mark.parametrize("mode", [0, 1])
def test_more_than_32_owned_partners_is_an_overflow_error(GU, mode):
    # 48 atoms inside one bind cutoff: the smallest ids own more than 32
    # bound partners, past the tracker's per-entity capacity.  The reference
    # has no cap; here gs_step must report GS_E_OVERFLOW rather than drop
    # partners, and on the list path the call returns to its start.
    from genesis_amd._lib import GS_E_OVERFLOW, GsError
    base, _ = melt(8)
    ph = Physics(**{**base.__dict__, "box": (40.0, 40.0, 40.0), "dt": 1e-7})
    rng = np.random.default_rng(7)
    cl = rng.uniform(-1.8, 1.8, (48, 3))
    gas = rng.uniform(-20.0, 20.0, (400, 3))
    gas = gas[np.linalg.norm(gas, axis=1) > 8.0]
    p = np.concatenate([cl, gas])
    n = len(p)
    species = np.where(np.arange(n) % 2 == 0, 5, 4).astype(np.uint8)
    z = np.zeros(n)
    st = State(p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), z, z.copy(), z.copy(), species,
               np.where(species == 5, 1.0, -1.0))
    _, owned = _bound_partner_counts(ph, p, species)
    assert owned.max() > 32
    say(f"owned-partner overflow, mode {mode}, max owned {owned.max()}")
    with GU(ph, st) as u:
        if mode == 1:
            u.set_tuning(1, 2)
        with pytest.raises(GsError) as e:
            u.step(2)
        assert e.value.code == GS_E_OVERFLOW
        got = u.download()
        if mode == 1:
            assert np.array_equal(got.x, st.x) and u.bond_stats()["tick"] == 0
        # the same state fails again; a fresh universe without the cluster
        # steps normally
        with pytest.raises(GsError):
            u.step(1)
        u.bond_reset()
        sub = State(st.x[48:].copy(), st.y[48:].copy(), st.z[48:].copy(), z[48:].copy(), z[48:].copy(),
                    z[48:].copy(), species[48:].copy(), st.charge[48:].copy())
    with GU(ph, sub) as u2:
        if mode == 1:
            u2.set_tuning(1, 2)
        u2.step(2)


def test_bond_force_with_a_bind_cutoff_inside_the_well(GU):
    # k_bind = 1.1 < 2^(1/6): the bind cutoff lies inside r_eq, where
    # smooth_cutoff(d, r_eq, rc) is (1, 0) (src/physics/forces.rs:302-305),
    # so every bonded pair closer than 1.1 sigma_ab feels the full harmonic
    # spring (src/systems/forces.rs:305-317).
    ph, st = melt(12, enable_bond_interaction=True, bond_k_bind=1.1)
    pairs = O.bound_pairs(ph, st.x, st.y, st.z, st.species)
    keep = [(a, b) for a, b in pairs if a < 1500 and b < 1500]
    adj = [[] for _ in range(st.n)]
    for a, b in keep:
        adj[a].append(b)
        adj[b].append(a)
    ptr = np.zeros(st.n + 1, dtype=np.int32)
    ptr[1:] = np.cumsum([len(x) for x in adj])
    idx = np.array([m for x in adj for m in sorted(x)], dtype=np.int32)
    say(f"bond force, k_bind 1.1: {len(keep)} bonds, n={st.n}")
    assert len(keep) >= 20
    ref = O.forces(ph, st, bonds_csr=(ptr, idx))
    with GU(ph, st) as u:
        u.upload_bonds(ptr, idx)
        V, Vb = u.forces()
        got = u.download()
    m = np.asarray(ph.mass)[st.species]
    ok, worst = force_tolerance_ok(got.ax, got.ay, got.az, ref, m, REL)
    assert ok, worst
    assert np.isfinite(Vb) and Vb > 0.0
    assert V == pytest.approx(ref["V"], rel=REL) and Vb == pytest.approx(ref["Vb"], rel=REL)


def test_autotune_with_every_trial_failing_leaves_the_universe_as_it_was(GU):
    # The 48-atom cluster overflows the bond tracker in every candidate
    # setting: autotune must report the overflow and leave the state, the
    # tuning and the timing mode untouched.
    from genesis_amd._lib import GS_E_OVERFLOW, GsError
    base, _ = melt(8)
    ph = Physics(**{**base.__dict__, "box": (40.0, 40.0, 40.0), "dt": 1e-7})
    rng = np.random.default_rng(7)
    cl = rng.uniform(-1.8, 1.8, (48, 3))
    gas = rng.uniform(-20.0, 20.0, (400, 3))
    gas = gas[np.linalg.norm(gas, axis=1) > 8.0]
    p = np.concatenate([cl, gas])
    n = len(p)
    species = np.where(np.arange(n) % 2 == 0, 5, 4).astype(np.uint8)
    z = np.zeros(n)
    st = State(p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), z, z.copy(), z.copy(), species,
               np.where(species == 5, 1.0, -1.0))
    say("autotune with failing trials")
    with GU(ph, st) as u:
        u.set_tuning(1, 8, skin=0.3)
        u.timing("pair", every=3)
        before = u.tuning()
        with pytest.raises(GsError) as e:
            u.autotune(8)
        assert e.value.code == GS_E_OVERFLOW
        after = u.tuning()
        got = u.download()
        for k in ("pair_mode", "lanes_per_atom", "skin"):
            assert after[k] == before[k], k
        assert np.array_equal(got.x, st.x) and np.array_equal(got.vx, st.vx)
        assert u.bond_stats()["tick"] == 0
