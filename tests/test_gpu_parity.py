"""HIP path vs the oracle on identical seeded inputs, through the C ABI.

Tolerances (the bar this path is held to, written here):
  * per-atom acceleration: |a_gpu - a_cpu|_inf <= 1e-10 max(|a_cpu|_inf, S_i/m_i)
    with S_i = sum_j |f_ij| from the oracle;
  * total potential / kinetic energy and per-atom PE: 1e-10 relative;
  * integrator steps, cell membership, g(r) bin counts, cluster labels:
    bit-exact.
"""
import sys

import numpy as np
import pytest

from genesis_amd.inputs import State, mixture, plain_lj, random_gas
from genesis_amd.params import DEFAULT_SIGMA, GRID_MINIMAGE, Physics
from oracle import oracle as O
from tests.helpers import force_tolerance_ok, reference_like

pytestmark = pytest.mark.gpu

REL = 1e-10


def say(msg):
    print(f"[gpu-parity] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def GU():
    from genesis_amd.universe import GpuUniverse
    return GpuUniverse


def check_forces(GU, ph, st, label):
    say(f"{label}: n={st.n} box={tuple(round(b, 3) for b in ph.box)} dims={ph.grid_dims()}")
    ref = O.forces(ph, st)
    with GU(ph, st) as u:
        V, Vb = u.forces()
        out = u.download()
        pe, sa = u.force_extras()
    m = np.asarray(ph.mass)[st.species]
    ok, worst = force_tolerance_ok(out.ax, out.ay, out.az, ref, m, REL)
    say(f"{label}: worst conditioned force error {worst:.3e}, V gpu {V:.17g} cpu {ref['V']:.17g}")
    assert ok, worst
    assert V == pytest.approx(ref["V"], rel=REL)
    scale = np.maximum(np.abs(ref["pe"]), 1e-3 * np.abs(ref["pe"]).mean())
    assert (np.abs(pe - ref["pe"]) <= REL * scale).all()
    np.testing.assert_allclose(sa, ref["sabs"], rtol=REL, atol=0)
    return out


def test_forces_plain_lj_4096(GU):
    w = plain_lj(n_side=16, steps=1)
    check_forces(GU, w.physics, w.state, "plain-lj-4096")


def test_forces_mixture_coulomb_4096(GU):
    w = mixture(n=4096, seed=21, steps=1)
    check_forces(GU, w.physics, w.state, "mix10-4096")


def test_forces_mixture_coulomb_32768(GU):
    w = mixture(n=32768, seed=1002, steps=1)
    check_forces(GU, w.physics, w.state, "mix10-32768")


def test_forces_reference_like_default_table(GU):
    ph = Physics(box=(24.0, 24.0, 24.0), enable_electrostatics=True)
    st = reference_like(8, 24.0, 100.0, seed=42)
    check_forces(GU, ph, st, "reference-like-512")


def test_forces_separate_coulomb_cutoff_and_gravity(GU):
    # Coulomb cutoff 3.2 max sigma (its own switch), gravity on: the unshared path.
    w = mixture(n=4096, seed=5, steps=1)
    ph = Physics(**{**w.physics.__dict__, "coulomb_cutoff": 3.2, "enable_gravity": True, "gravity_constant": 1e-3})
    check_forces(GU, ph, w.state, "mix10-coulomb3.2-gravity")


def test_forces_dense_cells_use_several_lds_chunks(GU):
    # ~1400 candidates per stencil > one LDS tile: exercises the chunked path.
    ph = Physics(box=(18.0, 18.0, 18.0))
    st = random_gas(3, 4000, ph.box, species_pool=(0, 1), min_sep=0.0)
    check_forces(GU, ph, st, "dense-two-species")


@pytest.mark.parametrize("semantics", [0, GRID_MINIMAGE])
def test_forces_small_box_two_cells(GU, semantics):
    # floor(L / rc) = 2 on every axis: the reference's 2-cell multiplicity
    # (semantics 0) or one count per pair (min-image semantics).
    ph = Physics(box=(12.0, 12.0, 12.0), grid_semantics=semantics)
    st = random_gas(8, 300, ph.box, species_pool=(0, 1, 2), min_sep=1.0)
    assert ph.grid_dims() == (2, 2, 2)
    check_forces(GU, ph, st, f"small-box-semantics{semantics}")


def test_forces_single_cell_axis(GU):
    ph = Physics(box=(30.0, 30.0, 7.0))
    st = random_gas(9, 500, ph.box, species_pool=(2, 4), min_sep=1.0)
    assert ph.grid_dims()[2] == 1
    check_forces(GU, ph, st, "one-cell-axis")


def test_cell_membership_bit_exact(GU):
    w = mixture(n=32768, seed=1002, steps=1)
    ph, st = w.physics, w.state
    say(f"cells: n={st.n}")
    ref, dims = O.grid_cells(st.x, st.y, st.z, np.asarray(ph.box), ph.force_rc())
    with GU(ph, st) as u:
        cells, gdims = u.cells()
    assert gdims == dims
    assert np.array_equal(cells, ref)


def test_cell_membership_bit_exact_on_a_grid_beyond_one_scan_pass(GU):
    # 26^3 = 17576 cells: the histogram scan runs in several workgroups
    L = 26 * 5.75 * 1.001
    ph = Physics(box=(L, L, L))
    st = random_gas(seed=31, n=40000, box=ph.box, species_pool=(0, 4, 5))
    ref, dims = O.grid_cells(st.x, st.y, st.z, np.asarray(ph.box), ph.force_rc())
    assert np.prod(dims) > 16384
    say(f"cells beyond one scan pass: n={st.n} dims={dims}")
    with GU(ph, st) as u:
        cells, gdims = u.cells()
    assert gdims == dims
    assert np.array_equal(cells, ref)


def test_integrator_systems_bit_exact(GU):
    w = mixture(n=4096, seed=4, steps=1)
    ph, st = w.physics, w.state
    st.ax, st.ay, st.az = st.vx * 40.0, -st.vy * 30.0, st.vz * 3.0
    st.x = st.x + 0.6 * ph.box[0]  # partly outside the box: wrap has work to do
    say("integrator systems")
    with GU(ph, st) as u:
        u.half_kick()
        u.drift()
        u.wrap()
        got = u.download()
    ref = st.copy()
    O.half_kick(ref, ph.dt)
    O.drift(ref, ph.dt)
    O.wrap(ref, ph.box)
    for a in ("x", "y", "z", "vx", "vy", "vz"):
        assert np.array_equal(getattr(got, a), getattr(ref, a)), a


def test_thermostat_and_kinetic(GU):
    w = mixture(n=4096, seed=6, temperature=60.0, steps=1)
    ph, st = w.physics, w.state
    ph = Physics(**{**ph.__dict__, "thermostat_temperature": 90.0})
    say("thermostat")
    with GU(ph, st) as u:
        k = u.kinetic()
        lam = u.thermostat()
        got = u.download()
    ref = st.copy()
    assert k == pytest.approx(O.kinetic(ref, ph.mass), rel=1e-13)
    lam_ref = O.thermostat(ph, ref)
    assert lam == pytest.approx(lam_ref, rel=1e-14)
    np.testing.assert_allclose(got.vx, ref.vx, rtol=1e-14, atol=0)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_one_tick_matches_oracle(GU, mode):
    # Tick 1 starts with a = 0 (src/universe/mod.rs:272,357): kick, drift and
    # wrap are then bit-exact, forces and the second kick within tolerance.
    w = mixture(n=4096, seed=8, steps=1)
    ph, st = w.physics, w.state
    say(f"one tick, pair_mode {mode}")
    ref, obs_ref = O.run(ph, st, 1)
    with GU(ph, st) as u:
        u.set_tuning(mode)
        obs = u.step(1)
        got = u.download()
    for a in ("x", "y", "z"):
        assert np.array_equal(getattr(got, a), getattr(ref, a)), a
    ref_f = O.forces(ph, State(ref.x, ref.y, ref.z, ref.vx, ref.vy, ref.vz, st.species, st.charge))
    m = np.asarray(ph.mass)[st.species]
    ok, worst = force_tolerance_ok(got.ax, got.ay, got.az, ref_f, m, REL)
    assert ok, worst
    np.testing.assert_allclose(got.vx, ref.vx, rtol=1e-9, atol=1e-12 * np.abs(ref.vx).max())
    assert obs[0, 1] == pytest.approx(obs_ref[0, 1], rel=REL)  # V
    assert obs[0, 0] == pytest.approx(obs_ref[0, 0], rel=1e-12)  # K
    assert obs[0, 4] == pytest.approx(obs_ref[0, 4], rel=1e-13)  # lambda


@pytest.mark.parametrize("mode,lanes", [(0, 1), (1, 1), (1, 4), (1, 8), (2, 0)])
def test_short_trajectory_tracks_oracle(GU, mode, lanes):
    # 20 ticks: trajectories stay within round-off growth of the oracle's
    # (chaos amplifies ~1e-16 slowly over few steps).
    w = plain_lj(n_side=12, steps=20)
    ph, st = w.physics, w.state
    say(f"20 ticks plain LJ n=1728 pair_mode {mode} lanes {lanes}")
    ref, obs_ref = O.run(ph, st, 20)
    with GU(ph, st) as u:
        u.set_tuning(mode, lanes)
        obs = u.step(20)
        got = u.download()
    np.testing.assert_allclose(got.x, ref.x, rtol=0, atol=1e-9)
    np.testing.assert_allclose(obs[:, 5], obs_ref[:, 5], rtol=1e-9)


def test_nve_energy_and_momentum_conservation(GU):
    w = plain_lj(n_side=16, steps=1000)
    ph, st = w.physics, w.state
    say("NVE 1000 ticks plain LJ n=4096")
    with GU(ph, st) as u:
        obs = u.step(1000)
        got = u.download()
    e = obs[:, 5]
    drift = abs(e[-1] - e[100]) / abs(e[100])
    say(f"NVE relative drift {drift:.3e}")
    assert drift < 1e-4
    m = np.asarray(ph.mass)[st.species]
    p = np.array([(m * got.vx).sum(), (m * got.vy).sum(), (m * got.vz).sum()])
    assert np.linalg.norm(p) < 1e-10 * (m * np.abs(got.vx)).sum()


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_step_is_bitwise_deterministic(GU, mode):
    w = mixture(n=32768, seed=1002, steps=50)
    ph, st = w.physics, w.state
    say(f"determinism 2 x 50 ticks n=32768 pair_mode {mode}")
    outs = []
    for _ in range(2):
        with GU(ph, st) as u:
            u.set_tuning(mode)
            obs = u.step(50)
            outs.append((u.download(), obs))
    for a in ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az"):
        assert np.array_equal(getattr(outs[0][0], a), getattr(outs[1][0], a)), a
    assert np.array_equal(outs[0][1], outs[1][1])


@pytest.mark.parametrize("n,rmax_frac,nbins", [(32768, 0.25, 200), (4096, 0.5, 100), (20000, 0.2, 5000)])
def test_radial_distribution_bit_exact(GU, n, rmax_frac, nbins):
    w = mixture(n=n, seed=n + 1, steps=1)
    ph, st = w.physics, w.state
    say(f"g(r) n={n} r_max={rmax_frac}L bins={nbins}")
    r_max = rmax_frac * ph.box[0]
    c_ref, g_ref, rm_ref, dr_ref = O.radial_distribution(st.x, st.y, st.z, np.asarray(ph.box), r_max, nbins)
    with GU(ph, st) as u:
        c, g, rm, dr = u.radial_distribution(r_max, nbins)
    assert (rm, dr) == (rm_ref, dr_ref)
    assert np.array_equal(c, c_ref)
    assert np.array_equal(g, g_ref)


def test_partial_radial_distribution_bit_exact(GU):
    w = mixture(n=8000, seed=77, steps=1)
    ph, st = w.physics, w.state
    say("partial g(r)")
    with GU(ph, st) as u:
        for ta, tb in ((4, 5), (0, 0), (9, 2)):
            c_ref, g_ref, _, _ = O.radial_distribution(st.x, st.y, st.z, np.asarray(ph.box), 6.0, 120,
                                                       species=st.species, ta=ta, tb=tb)
            c, g, _, _ = u.radial_distribution_between(ta, tb, 6.0, 120)
            assert np.array_equal(c, c_ref), (ta, tb)
            assert np.array_equal(g, g_ref)


def test_full_radial_distribution_bit_exact(GU):
    # every atom of the 32,768-atom mixture (the reference's lens would take
    # 8,192 of them), total and partial, against the oracle's full mode
    w = mixture(n=32768, seed=32769, steps=1)
    ph, st = w.physics, w.state
    r_max = 0.25 * ph.box[0]
    say(f"full g(r) n={st.n} r_max={r_max:.3f}")
    with GU(ph, st) as u:
        for ta, tb in ((-1, -1), (4, 5), (0, 0)):
            c_ref, g_ref, rm_ref, dr_ref = O.radial_distribution(st.x, st.y, st.z, np.asarray(ph.box), r_max, 200,
                                                                 species=st.species, ta=ta, tb=tb, full=True)
            c, g, rm, dr = u.radial_distribution_full(r_max, 200, ta, tb)
            assert (rm, dr) == (rm_ref, dr_ref)
            assert np.array_equal(c, c_ref), (ta, tb)
            assert np.array_equal(g, g_ref)
            if ta < 0:
                assert c.sum() > 16 * O.radial_distribution(st.x, st.y, st.z, np.asarray(ph.box), r_max, 200)[0].sum()


@pytest.mark.parametrize("n,density", [(32768, 1.0), (5000, 0.3)])
def test_cluster_labels_bit_exact(GU, n, density):
    w = mixture(n=n, seed=n + 3, steps=1)
    ph, st = w.physics, w.state
    if density != 1.0:
        s = density ** (-1.0 / 3.0)
        ph = Physics(**{**ph.__dict__, "box": tuple(b * s for b in ph.box)})
        st.x, st.y, st.z = st.x * s, st.y * s, st.z * s
    say(f"clusters n={n} density x{density}")
    lab_ref, s_ref = O.clusters(st.x, st.y, st.z, st.species, DEFAULT_SIGMA, np.asarray(ph.box))
    with GU(ph, st) as u:
        lab, s = u.clusters(1.5)
    assert np.array_equal(lab, lab_ref)
    assert s == s_ref


def test_cluster_fixture_dimers_and_monomers(GU):
    pos = [(10, 10, 10), (12, 10, 10), (30, 10, 10), (32, 10, 10), (50, 50, 50), (5, 50, 50)]
    x, y, z = [np.array([p[k] for p in pos], dtype=float) - 32.0 for k in range(3)]
    st = State(x, y, z, np.zeros(6), np.zeros(6), np.zeros(6), np.zeros(6, dtype=np.uint8), np.zeros(6))
    ph = Physics(box=(64.0, 64.0, 64.0))
    with GU(ph, st) as u:
        lab, s = u.clusters(1.5)
    assert lab.tolist() == [0, 0, 2, 2, 4, 5]
    assert (s["aggregates"], s["monomers"], s["largest"], s["bound_pairs"], s["mean_size"]) == (2, 2, 2, 2, 1.5)


def test_empty_universe(GU):
    z = np.zeros(0)
    st = State(z, z, z, z, z, z, np.zeros(0, dtype=np.uint8), z)
    with GU(Physics(box=(10.0, 10.0, 10.0)), st) as u:
        assert u.forces() == (0.0, 0.0)
        assert u.step(3).shape == (3, 6)
        c, g, _, _ = u.radial_distribution(4.0, 10)
        assert c.sum() == 0


@pytest.fixture(scope="module")
def mixture_60_ticks():
    w = mixture(n=32768, seed=1002, steps=60)
    ref, obs_ref = O.run(w.physics, w.state, 60, nthreads=8)
    return w, ref, obs_ref


@pytest.mark.parametrize("mode,lanes", [(1, 1), (1, 2), (1, 4), (1, 8), (2, 0)])
def test_list_mode_matches_oracle_over_a_mixture_trajectory(GU, mode, lanes, mixture_60_ticks):
    # 60 ticks of the 32k mixture with Coulomb and Berendsen: the list paths
    # (lists rebuilt only when an atom moved > skin/2) against the oracle,
    # which rebuilds its grid every tick.  At the default skin a stencil holds
    # about 1200 candidates, so the cell-staged pass stages many cells in two
    # chunks.
    w, ref, obs_ref = mixture_60_ticks
    ph, st = w.physics, w.state
    say(f"list mode {mode} lanes {lanes}: 60 ticks n=32768 vs oracle")
    with GU(ph, st) as u:
        u.set_tuning(mode, lanes, skin=0.12 * ph.force_rc())
        obs = u.step(60)
        got = u.download()
        tun = u.tuning()
    say(f"list builds {tun['list_builds']}, sort every {tun['sort_every']}")
    assert tun["pair_mode"] == mode and tun["list_builds"] >= 1
    np.testing.assert_allclose(obs[:, 1], obs_ref[:, 1], rtol=1e-8)
    np.testing.assert_allclose(obs[:, 0], obs_ref[:, 0], rtol=1e-8)
    np.testing.assert_allclose(got.x, ref.x, rtol=0, atol=1e-8)


@pytest.mark.parametrize("mode", [1, 2])
def test_list_mode_forces_after_many_ticks_match_a_fresh_force_pass(GU, mode):
    # After 300 list-mode ticks the accelerations in the state must equal a
    # fresh per-tick-grid force pass (oracle) on the same positions.
    w = mixture(n=8000, seed=31, steps=300)
    ph, st = w.physics, w.state
    say(f"list mode {mode} 300 ticks then oracle force pass on the final positions")
    with GU(ph, st) as u:
        u.set_tuning(mode, 4)
        u.step(300)
        got = u.download()
        tun = u.tuning()
    say(f"list builds {tun['list_builds']}, sort every {tun['sort_every']}")
    ref = O.forces(ph, State(got.x, got.y, got.z, got.vx, got.vy, got.vz, st.species, st.charge))
    m = np.asarray(ph.mass)[st.species]
    ok, worst = force_tolerance_ok(got.ax, got.ay, got.az, ref, m, REL)
    say(f"worst conditioned error {worst:.3e}")
    assert ok, worst


@pytest.mark.parametrize("mode", [1, 2])
def test_list_rebuilds_on_device_when_atoms_pass_half_skin(GU, mode):
    # A hot gas with a tiny skin: the device must rebuild the lists on many
    # ticks (no host round trip decides it), and the trajectory must match
    # the per-tick cell tiles.  Two calls, and a short re-sort period, cover
    # the lists surviving across calls and the forced re-sorts.
    w = mixture(n=4096, seed=41, temperature=600.0, steps=40, thermostat=False)
    ph, st = w.physics, w.state
    say(f"device-driven list rebuilds, hot gas, skin 0.02, mode {mode}")
    with GU(ph, st) as u:
        u.set_tuning(mode, 2, skin=0.02, sort_every=15)
        obs_l = np.concatenate([u.step(25), u.step(15)])
        got_l = u.download()
        tun = u.tuning()
    with GU(ph, st) as u:
        u.set_tuning(0)
        obs_t = u.step(40)
        got_t = u.download()
    say(f"list builds {tun['list_builds']} in 40 ticks")
    assert tun["pair_mode"] == mode and tun["list_builds"] >= 10
    np.testing.assert_allclose(obs_l[:, 5], obs_t[:, 5], rtol=1e-9)
    np.testing.assert_allclose(got_l.x, got_t.x, rtol=0, atol=1e-9)


@pytest.mark.parametrize("mode", [1, 2])
def test_list_mode_replays_a_rebuild_its_binning_did_not_foresee(GU, mode):
    # A cold gas heated a hundredfold in one tick (Berendsen with tau of one
    # tick): the next ticks' displacements jump far beyond what the previous
    # tick's predicted, so a tick needs a rebuild whose binning was not
    # requested.  The call is replayed with a binning launch every tick and
    # must reproduce the per-tick cell tiles.
    w = mixture(n=4096, seed=47, temperature=30.0, steps=12)
    ph = Physics(**{**w.physics.__dict__, "thermostat_temperature": 3000.0, "thermostat_tau": 1.0,
                    "enable_thermostat": True})
    st = w.state
    say(f"unforeseen rebuild replay, mode {mode}")
    with GU(ph, st) as u:
        u.set_tuning(mode, 2, skin=0.05, sort_every=100)
        obs_l = np.concatenate([u.step(4), u.step(8)])
        got_l = u.download()
        tun = u.tuning()
    with GU(ph, st) as u:
        u.set_tuning(0)
        obs_t = u.step(12)
        got_t = u.download()
    assert tun["pair_mode"] == mode and tun["list_builds"] >= 2
    np.testing.assert_allclose(obs_l[:, 5], obs_t[:, 5], rtol=1e-9)
    np.testing.assert_allclose(got_l.x, got_t.x, rtol=0, atol=1e-9)


@pytest.mark.parametrize("mode", [1, 2])
def test_list_mode_falls_back_to_tiles_when_a_stencil_is_too_dense(GU, mode):
    # A skin so large that a 27-cell stencil holds more candidates than the
    # list build stages in LDS: the call is replayed on the cell tiles and
    # the tuning reports pair_mode 0.
    w = mixture(n=32768, seed=43, steps=10)
    ph, st = w.physics, w.state
    say(f"dense stencil fallback, mode {mode}")
    with GU(ph, st) as u:
        u.set_tuning(mode, 4, skin=0.6 * ph.force_rc())
        obs_l = u.step(10)
        got_l = u.download()
        tun = u.tuning()
    with GU(ph, st) as u:
        u.set_tuning(0)
        obs_t = u.step(10)
        got_t = u.download()
    assert tun["pair_mode"] == 0
    assert np.array_equal(obs_l, obs_t) and np.array_equal(got_l.x, got_t.x)


@pytest.mark.parametrize("mode", [1, 2])
def test_crowded_cell_overflows_and_replays(GU, mode):
    # 400 atoms packed inside one list-grid cell of an otherwise sparse box:
    # the first ticks overflow the cell buckets and the lists (399 neighbours
    # each), so the call is replayed with doubled capacities, and the
    # cell-staged pass meets more home atoms than lanes (two passes over the
    # home atoms, one lane each).  One tick against the oracle.
    rng = np.random.default_rng(17)
    ph0 = plain_lj(n_side=4, steps=1).physics
    # six cells per axis: the wrap's half-box translation maps cell centres
    # to cell centres, so the blob stays inside one cell
    L = 6.0 * 1.12 * ph0.force_rc() * 1.001
    ph = Physics(**{**ph0.__dict__, "box": (L, L, L)})
    dims = int(np.floor(L / (1.12 * ph.force_rc())))
    assert dims == 6
    edge = L / dims
    centre = -0.5 * L + 1.5 * edge
    nb, ns = 400, 200
    blob = centre + rng.uniform(-0.3 * edge, 0.3 * edge, (nb, 3))
    far = rng.uniform(-0.5 * L, 0.5 * L, (ns, 3))
    far = far[np.linalg.norm(far - centre, axis=1) > 2.5 * edge][:150]
    p = np.concatenate([blob, far])
    n = len(p)
    z = np.zeros(n)
    st = State(p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), z, z.copy(), z.copy(), np.zeros(n, dtype=np.uint8), z.copy())
    say(f"crowded cell: {nb} atoms in one cell of edge {edge:.3f}, n={n}, mode {mode}")
    ref, obs_ref = O.run(ph, st, 1)
    with GU(ph, st) as u:
        u.set_tuning(mode, 4, skin=0.12 * ph.force_rc())
        obs = u.step(1)
        got = u.download()
        tun = u.tuning()
    assert tun["pair_mode"] == mode
    for a in ("x", "y", "z"):
        assert np.array_equal(getattr(got, a), getattr(ref, a)), a
    ref_f = O.forces(ph, State(ref.x, ref.y, ref.z, ref.vx, ref.vy, ref.vz, st.species, st.charge))
    m = np.asarray(ph.mass)[st.species]
    ok, worst = force_tolerance_ok(got.ax, got.ay, got.az, ref_f, m, REL)
    assert ok, worst
    assert obs[0, 1] == pytest.approx(obs_ref[0, 1], rel=REL)


def test_autotune_restores_state_and_picks_a_valid_setting(GU):
    w = mixture(n=32768, seed=1002, steps=1)
    ph, st = w.physics, w.state
    say("autotune")
    with GU(ph, st) as u:
        before = u.download()
        best = u.autotune(16)
        after = u.download()
        say(f"autotune picked {best}")
        for a in ("x", "vx"):
            assert np.array_equal(getattr(before, a), getattr(after, a))
        assert best["pair_mode"] in (0, 1, 2) and best["ms_per_tick"] > 0
        obs = u.step(10)
        assert np.isfinite(obs).all()


def test_invalid_tuning_is_rejected(GU):
    from genesis_amd._lib import GsError
    w = mixture(n=512, seed=2, steps=1)  # 2 cells per axis: no list mode
    with GU(w.physics, w.state) as u:
        assert u.tuning()["pair_mode"] == 0
        with pytest.raises(GsError):
            u.set_tuning(1, 4)
        with pytest.raises(GsError):
            u.set_tuning(1, 3)


@pytest.mark.parametrize("nbins,max_speed", [(64, None), (512, None), (7, 0.0)])
def test_velocity_histogram_bit_exact(GU, nbins, max_speed):
    # velocity_histogram (src/stats/mod.rs:269-297) of the resident state
    # after 20 ticks, with speeds planted on bin edges, at max_speed and zero
    w = mixture(n=4096, seed=12, steps=20)
    ph, st = w.physics, w.state
    vmax = float(np.sqrt(st.vx ** 2 + st.vy ** 2 + st.vz ** 2).max()) * 0.8 if max_speed is None else max_speed
    st = st.copy()
    width = max(vmax, 1e-9) / nbins
    st.vx[:4] = [0.0, width, 3 * width, max(vmax, 1e-9)]
    st.vy[:4] = st.vz[:4] = 0.0
    say(f"velocity histogram nbins={nbins} max={vmax:.4g}")
    with GU(ph, st) as u:
        h0 = u.velocity_histogram(vmax, nbins)
        u.step(20)
        got = u.download()
        h1 = u.velocity_histogram(vmax, nbins)
    for h, s in ((h0, st), (h1, got)):
        ref = O.velocity_histogram(s.vx, s.vy, s.vz, vmax, nbins)
        assert h["bin_width"] == ref["bin_width"] and h["samples"] == ref["samples"] == st.n
        assert h["overflow"] == ref["overflow"]
        assert np.array_equal(h["bins"], ref["bins"])


def _one_tick_forces_vs_oracle(GU, ph, st, mode, lanes, label, skin=None):
    # Tick 1 starts with a = 0, so the GPU's positions after it equal the
    # oracle's bit for bit and the forces of the tick compare directly.
    say(f"{label}: one tick, pair_mode {mode} lanes {lanes}, n={st.n} dims={ph.grid_dims()}")
    ref, obs_ref = O.run(ph, st, 1)
    with GU(ph, st) as u:
        u.set_tuning(mode, lanes, skin=skin)
        obs = u.step(1)
        got = u.download()
        tun = u.tuning()
    assert tun["pair_mode"] == mode
    for a in ("x", "y", "z"):
        assert np.array_equal(getattr(got, a), getattr(ref, a)), a
    ref_f = O.forces(ph, State(ref.x, ref.y, ref.z, ref.vx, ref.vy, ref.vz, st.species, st.charge))
    m = np.asarray(ph.mass)[st.species]
    ok, worst = force_tolerance_ok(got.ax, got.ay, got.az, ref_f, m, REL)
    say(f"{label}: worst conditioned force error {worst:.3e}")
    assert ok, worst
    assert obs[0, 1] == pytest.approx(obs_ref[0, 1], rel=REL)


def _close_pairs(ph, st, cut):
    p = np.stack([st.x, st.y, st.z], 1)
    d = p[:, None, :] - p[None, :, :]
    d -= np.asarray(ph.box) * np.round(d / np.asarray(ph.box))
    r = np.sqrt((d * d).sum(-1))
    iu = np.triu_indices(st.n, 1)
    return r[iu], iu


@pytest.fixture(scope="module")
def charged_gas():
    # Na/O/Si/Fe gas with no minimum separation: many charged pairs closer
    # than the Coulomb clamp (0.5) and inside the hardened LJ core
    # (0.5 sigma_ij), with electrostatics on (src/systems/forces.rs:282-289,
    # src/physics/forces.rs:170-173).
    ph = Physics(box=(30.0, 30.0, 30.0), enable_electrostatics=True)
    st = random_gas(61, 3000, ph.box, species_pool=(4, 5, 6, 9), min_sep=0.0)
    r, (ia, ib) = _close_pairs(ph, st, 0.5)
    sig = np.asarray(DEFAULT_SIGMA)
    core = r < 0.5 * 0.5 * (sig[st.species[ia]] + sig[st.species[ib]])
    charged = (st.charge[ia] != 0) & (st.charge[ib] != 0)
    assert ((r < 0.5) & charged).sum() >= 20 and (core & charged).sum() >= 100
    return ph, st


def test_forces_charged_gas_inside_clamp_and_core(GU, charged_gas):
    ph, st = charged_gas
    check_forces(GU, ph, st, "charged-gas-clamp-core")


@pytest.mark.parametrize("mode,lanes", [(0, 1), (1, 2), (1, 4), (2, 0)])
def test_charged_gas_inside_clamp_and_core_on_every_pair_path(GU, charged_gas, mode, lanes):
    ph, st = charged_gas
    _one_tick_forces_vs_oracle(GU, ph, st, mode, lanes, "charged-gas", skin=0.12 * ph.force_rc())


@pytest.mark.parametrize("mode,lanes", [(0, 1), (1, 4), (2, 0)])
def test_coulomb_switch_straddling_the_clamp(GU, mode, lanes):
    # sigma 0.2 for every species and coulomb_cutoff 2.7: the Coulomb switch
    # runs on [0.486, 0.54), across the 0.5 clamp, so pairs in [0.486, 0.5)
    # take the clamped law with the switch at d, and the LJ core (0.1) and
    # cutoff (0.5) sit inside the Coulomb range.
    ph = Physics(box=(6.0, 6.0, 6.0), enable_electrostatics=True, coulomb_cutoff=2.7, sigma=(0.2,) * 10)
    st = random_gas(62, 2000, ph.box, species_pool=(4, 5, 6, 9), min_sep=0.0)
    r, _ = _close_pairs(ph, st, 0.54)
    assert ((r > 0.486) & (r < 0.5)).sum() >= 20 and (r < 0.1).sum() >= 2
    _one_tick_forces_vs_oracle(GU, ph, st, mode, lanes, "switch-across-clamp", skin=0.12 * ph.force_rc())


@pytest.mark.parametrize("mode,lanes", [(0, 1), (1, 4)])
def test_non_finite_energy_raises_from_the_device_fault_word(GU, mode, lanes):
    # one atom at a finite speed of 1e200: 0.5 m v^2 overflows to +inf on the
    # first tick (positions stay finite, the wrap folds them back into the
    # box), so the per-tick kinetic energy is non-finite on every tick
    from genesis_amd import _lib
    L = 30.0
    ph = Physics(box=(L, L, L))
    st = random_gas(71, 3000, ph.box, species_pool=(0, 4))
    bad = st.copy()
    bad.vx[17] = 1e200
    print(f"[gpu-parity] fault word mode={mode} lanes={lanes}", flush=True)
    with GU(ph, bad) as u:
        if mode:
            u.set_tuning(mode, lanes)
        with pytest.raises(_lib.GsError) as e:
            u.step(5)
        assert e.value.code == _lib.GS_E_NONFINITE
        assert "kinetic energy" in str(e.value) and "potential" not in str(e.value)
        assert "ticks 1..5" in str(e.value)
        got = u.download()
        assert np.isfinite(got.x).all() and got.vx[17] == 1e200
        # the word is cleared at every call: a clean state steps without error
        u.upload(st)
        obs = u.step(3)
        assert np.isfinite(obs).all()
