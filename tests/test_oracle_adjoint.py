"""The oracle's adjoint (oracle/adjoint.cpp) pinned by finite differences:
per-pair derivatives, the Hessian-vector product, the parameter VJP and
the gradient of the whole trajectory loss (central differences with
h = 1e-6 relative, as the adjoint workload specifies)."""
import numpy as np
import pytest

from genesis_amd.inputs import mixture
from oracle import oracle as O


@pytest.mark.parametrize("d", [0.7, 1.6, 2.05, 3.3, 5.3, 5.55, 5.7])
@pytest.mark.parametrize("coul", [False, True])
def test_pair_derivatives_by_finite_differences(d, coul):
    sig, eps, qq, rc = 1.9, 0.8, -0.7, 5.75
    base = O.pair_derivs(d, sig, eps, qq, rc, coul)

    def fd(name, arg, h):
        kw = dict(d=d, sigma=sig, eps=eps, qq=qq, rc=rc)
        kw[arg] += h
        p = O.pair_derivs(kw["d"], kw["sigma"], kw["eps"], kw["qq"], kw["rc"], coul)
        kw[arg] -= 2 * h
        m = O.pair_derivs(kw["d"], kw["sigma"], kw["eps"], kw["qq"], kw["rc"], coul)
        if name == "f":
            return -(p["d1"] - m["d1"]) / (2 * h)
        return (p[name] - m[name]) / (2 * h)

    h = 1e-6
    checks = [("d1", "phi", "d"), ("d2", "d1", "d"), ("phi_s", "phi", "sigma"), ("phi_e", "phi", "eps"),
              ("phi_rc", "phi", "rc"), ("f_s", "f", "sigma"), ("f_e", "f", "eps"), ("f_rc", "f", "rc")]
    if coul:
        checks += [("phi_q", "phi", "qq"), ("f_q", "f", "qq")]
    for out, of, arg in checks:
        num = fd(of, arg, h)
        scale = max(abs(num), abs(base[out]), 1e-3)
        assert abs(base[out] - num) <= 2e-6 * scale, (out, base[out], num)


def small_system(n=343, seed=5):
    w = mixture(seed=seed, n=n, temperature=60.0, thermostat=False, steps=1)
    keep = (np.abs(w.state.charge) > 0) | (np.asarray(w.physics.charge)[w.state.species] == 0)
    return w.physics, w.state, keep.astype(np.uint8)


def test_hessian_vector_product_adjoint_identity():
    ph, st, keep = small_system()
    th = O.theta_of(ph)
    rng = np.random.default_rng(1)
    w = rng.standard_normal((st.n, 3))
    lam = rng.standard_normal((st.n, 3))
    h = 1e-6
    sp, sm = st.copy(), st.copy()
    sp.x, sp.y, sp.z = st.x + h * w[:, 0], st.y + h * w[:, 1], st.z + h * w[:, 2]
    sm.x, sm.y, sm.z = st.x - h * w[:, 0], st.y - h * w[:, 1], st.z - h * w[:, 2]
    Aw = (O.adj_accel(ph, sp, keep, th)[1] - O.adj_accel(ph, sm, keep, th)[1]) / (2 * h)
    ATl, _ = O.adj_pass(ph, st, keep, th, lam)
    lhs, rhs = float((lam * Aw).sum()), float((w * ATl).sum())
    assert lhs == pytest.approx(rhs, rel=1e-6)


def test_parameter_vjp_by_finite_differences():
    ph, st, keep = small_system()
    th = O.theta_of(ph)
    lam = np.random.default_rng(2).standard_normal((st.n, 3))
    _, g = O.adj_pass(ph, st, keep, th, lam)
    for k in [0, 2, 4, 6, 10, 14, 15, 19, 24, 25, 26, 29]:
        h = 1e-6 * max(abs(th[k]), 1.0)
        tp, tm = th.copy(), th.copy()
        tp[k] += h
        tm[k] -= h
        num = float((lam * (O.adj_accel(ph, st, keep, tp)[1] - O.adj_accel(ph, st, keep, tm)[1])).sum()) / (2 * h)
        assert g[k] == pytest.approx(num, rel=1e-5, abs=1e-7 * np.abs(g).max()), k


def test_potential_parameter_gradient():
    ph, st, keep = small_system()
    th = O.theta_of(ph)
    _, g = O.adj_pass(ph, st, keep, th, np.zeros((st.n, 3)), vcoef=1.0)
    for k in [1, 4, 6, 12, 16, 24, 25]:
        h = 1e-6 * max(abs(th[k]), 1.0)
        tp, tm = th.copy(), th.copy()
        tp[k] += h
        tm[k] -= h
        num = (O.adj_accel(ph, st, keep, tp)[0] - O.adj_accel(ph, st, keep, tm)[0]) / (2 * h)
        assert g[k] == pytest.approx(num, rel=1e-6, abs=1e-9 * np.abs(g).max()), k


def test_trajectory_loss_gradient_by_central_differences():
    ph, st, keep = small_system(n=216, seed=7)
    ph.box = tuple(b * 1.15 for b in ph.box)  # 216 atoms: keep L > 2 rc
    st.x, st.y, st.z = st.x * 1.15, st.y * 1.15, st.z * 1.15
    th0 = O.theta_of(ph)
    gr = (40, 0.45 * ph.box[0], 0.05)
    T, win = 12, 5
    _, _, g0 = O.adjoint(ph, st, keep, th0, T, win, gr=gr, grad=False)
    th = th0 * (1.0 + 0.01 * np.cos(np.arange(30)))  # away from theta0 so the g(r) term is active
    th[20:] = th0[20:] * 1.01
    loss, grad, _ = O.adjoint(ph, st, keep, th, T, win, gr=gr, g0=g0)
    assert np.isfinite(grad).all()
    for k in [0, 3, 6, 9, 11, 14, 16, 19, 24, 25, 26, 29]:
        h = 1e-6 * abs(th[k])
        tp, tm = th.copy(), th.copy()
        tp[k] += h
        tm[k] -= h
        num = (O.adjoint(ph, st, keep, tp, T, win, gr=gr, g0=g0, grad=False)[0]
               - O.adjoint(ph, st, keep, tm, T, win, gr=gr, g0=g0, grad=False)[0]) / (2 * h)
        assert grad[k] == pytest.approx(num, rel=2e-4, abs=1e-6 * np.abs(grad).max()), (k, grad[k], num)
