"""The oracle's lenses against the reference's fixtures
(src/analysis/mod.rs:454-680) and against brute force."""
import math

import numpy as np
import pytest

from genesis_amd.inputs import Stream
from genesis_amd.params import DEFAULT_SIGMA, SPECIES_INDEX
from oracle import oracle as O

H, C, N_, O_, NA = (SPECIES_INDEX[s] for s in ("H", "C", "N", "O", "Na"))


def peak_in(g, dr, lo, hi):
    best = None
    for i, v in enumerate(g):
        r = (i + 0.5) * dr
        if r < lo or r > hi:
            continue
        if best is None or v > best[1]:
            best = (r, v)
    return best


def test_g_r_of_a_pair_peaks_at_the_distance():
    size = np.array([32.0] * 3)
    c, g, rm, dr = O.radial_distribution([0.0, 1.7], [0.0, 0.0], [0.0, 0.0], size, 4.0, 400)
    assert abs(dr - 0.01) < 1e-12
    r, gp = peak_in(g, dr, 0.5, 3.0)
    assert abs(r - 1.7) < 0.03 and gp > 1000.0
    assert c.sum() == 1 and c[int(1.7 / 0.01)] == 1


def test_g_r_of_random_gas_is_approximately_one():
    size = np.array([32.0] * 3)
    st = Stream(7, 0)
    x, y, z = [(st.uniform(400) - 0.5) * 32.0 for _ in range(3)]
    _, g, _, _ = O.radial_distribution(x, y, z, size, 8.0, 32)
    assert 0.8 <= g.mean() <= 1.2


def brute_hist(x, y, z, size, r_max, nbins):
    n = len(x)
    p = np.stack([x, y, z], 1)
    dr = r_max / nbins
    counts = np.zeros(nbins, dtype=np.uint64)
    for i in range(n):
        d = p[i] - p[i + 1:]
        d = d - size * np.round(d / size)
        r = np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2])
        r = r[(r < r_max) & (r > 2.220446049250313e-16)]
        np.add.at(counts, (r / dr).astype(np.int64), 1)
    return counts


@pytest.mark.parametrize("n,L,rmax,nb", [(700, 20.0, 5.0, 50), (2000, 30.0, 7.4, 200)])
def test_g_r_counts_equal_brute_force(n, L, rmax, nb):
    size = np.array([L, L, L])
    st = Stream(n, 0)
    x, y, z = [(st.uniform(n) - 0.5) * L for _ in range(3)]
    c, _, _, _ = O.radial_distribution(x, y, z, size, rmax, nb)
    assert np.array_equal(c, brute_hist(x, y, z, size, rmax, nb))


def test_g_r_subsample_above_cap():
    # g_subsample: N = 10000 > 8192 -> step 2, 5000 sampled atoms.
    n, L = 10000, 60.0
    size = np.array([L] * 3)
    st = Stream(3, 0)
    x, y, z = [(st.uniform(n) - 0.5) * L for _ in range(3)]
    c, _, _, _ = O.radial_distribution(x, y, z, size, 4.0, 40)
    idx = np.arange(0, 5000) * 2
    assert np.array_equal(c, brute_hist(x[idx], y[idx], z[idx], size, 4.0, 40))


def test_g_r_full_mode_counts_every_atom():
    # the same 10000 atoms as above without the subsample: brute force over
    # all of them; below the cap the two modes coincide
    n, L = 10000, 60.0
    size = np.array([L] * 3)
    st = Stream(3, 0)
    x, y, z = [(st.uniform(n) - 0.5) * L for _ in range(3)]
    c, g, _, dr = O.radial_distribution(x, y, z, size, 4.0, 40, full=True)
    assert np.array_equal(c, brute_hist(x, y, z, size, 4.0, 40))
    assert 0.8 <= g[10:].mean() <= 1.2
    k = 3000
    a, _, _, _ = O.radial_distribution(x[:k], y[:k], z[:k], size, 4.0, 40, full=True)
    b, _, _, _ = O.radial_distribution(x[:k], y[:k], z[:k], size, 4.0, 40)
    assert np.array_equal(a, b)


def test_partial_g_r_isolates_the_requested_species():
    size = np.array([32.0] * 3)
    st = Stream(11, 0)
    n = 2000
    x, y, z = [(st.uniform(n) - 0.5) * 32.0 for _ in range(3)]
    sp = np.full(n, H, dtype=np.uint8)
    x[0], y[0], z[0] = 0.0, 0.0, 0.0
    x[1], y[1], z[1] = 1.8, 0.0, 0.0
    sp[0] = sp[1] = C
    _, gcc, _, dr = O.radial_distribution(x, y, z, size, 4.0, 400, species=sp, ta=C, tb=C)
    assert peak_in(gcc, dr, 0.5, 3.0)[1] > 500.0
    _, ghh, _, _ = O.radial_distribution(x, y, z, size, 4.0, 400, species=sp, ta=H, tb=H)
    assert 0.9 <= ghh.mean() <= 1.15
    _, gch, _, _ = O.radial_distribution(x, y, z, size, 4.0, 400, species=sp, ta=C, tb=H)
    assert peak_in(gcc, dr, 0.5, 3.0)[1] / peak_in(gch, dr, 0.5, 3.0)[1] > 100.0


def test_partial_g_r_empty_species_is_flat():
    size = np.array([16.0] * 3)
    sp = np.array([H, H], dtype=np.uint8)
    _, g, _, _ = O.radial_distribution([0.0, 2.0], [0.0, 0.0], [0.0, 0.0], size, 4.0, 100, species=sp, ta=NA,
                                       tb=NA)
    assert (g == 0.0).all()


def test_clusters_detect_dimers_and_monomers():
    pos = [(10, 10, 10), (12, 10, 10), (30, 10, 10), (32, 10, 10), (50, 50, 50), (5, 50, 50)]
    # the reference positions are corner-origin in a 64 box; centre them
    x, y, z = [np.array([p[k] for p in pos], dtype=float) - 32.0 for k in range(3)]
    sp = np.full(6, H, dtype=np.uint8)
    labels, s = O.clusters(x, y, z, sp, DEFAULT_SIGMA, np.array([64.0] * 3))
    assert (s["aggregates"], s["monomers"], s["largest"], s["bound_pairs"]) == (2, 2, 2, 2)
    assert s["mean_size"] == 1.5
    assert labels.tolist() == [0, 0, 2, 2, 4, 5]


def test_clusters_empty_and_isolated():
    labels, s = O.clusters([], [], [], np.zeros(0, dtype=np.uint8), DEFAULT_SIGMA, np.array([64.0] * 3))
    assert s["largest"] == 0 and s["mean_size"] == 0.0
    _, s = O.clusters([-22.0, 18.0], [-22.0, 18.0], [-22.0, 18.0], np.array([H, H], dtype=np.uint8), DEFAULT_SIGMA,
                      np.array([64.0] * 3))
    assert (s["monomers"], s["aggregates"], s["bound_pairs"]) == (2, 0, 0)


def test_cluster_labels_equal_brute_force_components():
    n, L = 1500, 25.0
    st = Stream(21, 0)
    x, y, z = [(st.uniform(n) - 0.5) * L for _ in range(3)]
    sp = (Stream(21, 1).raw(n) % np.uint64(10)).astype(np.uint8)
    labels, s = O.clusters(x, y, z, sp, DEFAULT_SIGMA, np.array([L] * 3))
    sig = np.asarray(DEFAULT_SIGMA)[sp]
    p = np.stack([x, y, z], 1)
    parent = list(range(n))

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a

    bound = 0
    for i in range(n):
        d = p[i] - p[i + 1:]
        d -= L * np.round(d / L)
        r = np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2])
        for j in np.nonzero(r < 1.5 * 0.5 * (sig[i] + sig[i + 1:]))[0]:
            bound += 1
            a, b = find(i), find(i + 1 + j)
            if a != b:
                parent[max(a, b)] = min(a, b)
    ref = np.array([min(k for k in range(n) if find(k) == find(i)) for i in range(n)])
    assert np.array_equal(labels, ref)
    assert s["bound_pairs"] == bound


def test_bond_components_fixtures():
    # src/analysis/mod.rs:588-655
    comps = O.bond_components([(0, H), (1, H), (2, NA), (3, O_)], [(0, 1), (2, 3)])
    assert [c["size"] for c in comps] == [2, 2]
    assert [c["formula"] for c in comps] == ["H2", "Na-O"]
    comps = O.bond_components([(0, NA), (1, O_), (2, NA)], [(0, 1), (1, 2)])
    assert len(comps) == 1 and comps[0]["size"] == 3 and comps[0]["formula"] == "Na2-O"
    comps = O.bond_components([(0, H), (1, H)], [(0, 1), (1, 0), (7, 0), (9, 9)])
    assert len(comps) == 1 and comps[0]["formula"] == "H2" and comps[0]["size"] == 2
    comps = O.bond_components([(0, C), (1, C), (2, C)], [(0, 1), (1, 2), (2, 0)])
    assert len(comps) == 1 and comps[0]["size"] == 3 and comps[0]["formula"] == "C3"
    assert O.bond_components([], []) == []


def test_formula_orders_multiplicity_then_symbol():
    # src/analysis/mod.rs:657-680: 6 H + 2 C -> "H6-C2"; Na2-O
    ents = [(i, H) for i in range(6)] + [(6, C), (7, C)]
    edges = [(i, i + 1) for i in range(7)]
    assert O.bond_components(ents, edges)[0]["formula"] == "H6-C2"


def test_velocity_histogram_hand_fixture():
    # speeds on bin edges land in the upper bin ((s / w) as usize), the speed
    # equal to max_speed and above is overflow, a zero speed is bin 0, and a
    # NaN speed (s >= max fails) lands in bin 0 as `NaN as usize` = 0
    from oracle import oracle as O
    w = 0.25
    vx = np.array([0.0, 0.25, 0.5, 0.74, 0.99, 1.0, 3.0, np.nan])
    z = np.zeros_like(vx)
    h = O.velocity_histogram(vx, z, z, 1.0, 4)
    assert h["bin_width"] == w and h["samples"] == 8 and h["overflow"] == 2
    assert list(h["bins"]) == [2, 1, 2, 1]
    # 3-4-5: |(0.3, 0.4, 0)| = 0.5 exactly -> bin 2 of width 0.25
    h = O.velocity_histogram([0.3], [0.4], [0.0], 1.0, 4)
    assert list(h["bins"]) == [0, 0, 1, 0]
    # bins clamp to 1..512; max_speed floors at 1e-9
    assert len(O.velocity_histogram([1.0], [0.0], [0.0], 2.0, 0)["bins"]) == 1
    assert len(O.velocity_histogram([1.0], [0.0], [0.0], 2.0, 10_000)["bins"]) == 512
    assert O.velocity_histogram([0.0], [0.0], [0.0], -1.0, 4)["max_speed"] == 1e-9


def test_velocity_histogram_brute_force():
    from oracle import oracle as O
    rng = np.random.default_rng(11)
    v = rng.normal(0.0, 1.0, (5000, 3))
    h = O.velocity_histogram(v[:, 0], v[:, 1], v[:, 2], 3.0, 37)
    s = [math.sqrt(a * a + b * b + c * c) for a, b, c in v]
    ref = [0] * 37
    over = 0
    for x in s:
        if x >= 3.0:
            over += 1
        else:
            ref[min(int(x / (3.0 / 37)), 36)] += 1
    assert list(h["bins"]) == ref and h["overflow"] == over and sum(ref) + over == 5000
