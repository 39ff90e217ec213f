"""GPU adjoint (gs_hvp, gs_adjoint_forward/backward and the torch op) vs the
oracle's FP64 adjoint and vs central finite differences."""
import sys

import numpy as np
import pytest

from genesis_amd.inputs import mixture
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-adjoint] {msg}", flush=True)


def system(n=1000, seed=5):
    w = mixture(seed=seed, n=n, temperature=60.0, thermostat=False, steps=1)
    keep = ((np.abs(w.state.charge) > 0) | (np.asarray(w.physics.charge)[w.state.species] == 0)).astype(np.uint8)
    return w.physics, w.state, keep


def test_hvp_and_parameter_vjp_match_oracle():
    from genesis_amd.universe import GpuUniverse
    ph, st, keep = system()
    th = O.theta_of(ph) * (1.0 + 0.003 * np.sin(np.arange(30)))
    lam = np.random.default_rng(3).standard_normal((st.n, 3))
    say(f"hvp n={st.n}")
    ref_o, ref_g = O.adj_pass(ph, st, keep, th, lam, vcoef=0.37)
    with GpuUniverse(ph, st) as u:
        o, g = u.hvp(th, keep, lam, vcoef=0.37)
    np.testing.assert_allclose(o, ref_o, rtol=1e-10, atol=1e-10 * np.abs(ref_o).max())
    np.testing.assert_allclose(g, ref_g, rtol=1e-10, atol=1e-12 * np.abs(ref_g).max())


@pytest.mark.parametrize("ckpt", [4, 12])
def test_trajectory_gradient_matches_oracle(ckpt):
    from genesis_amd.universe import GpuUniverse
    ph, st, keep = system()
    th0 = O.theta_of(ph)
    gr = (40, 0.45 * ph.box[0], 0.05)
    T, win = 12, 5
    _, _, g0 = O.adjoint(ph, st, keep, th0, T, win, gr=gr, grad=False)
    th = th0 * (1.0 + 0.01 * np.cos(np.arange(30)))
    say(f"trajectory gradient T={T} ckpt={ckpt}")
    loss_ref, grad_ref, g_ref = O.adjoint(ph, st, keep, th, T, win, gr=gr, g0=g0)
    with GpuUniverse(ph, st) as u:
        loss, g_fin = u.adjoint_forward(th, keep, T, ckpt, win, gr=gr, g0=g0)
        grad = u.adjoint_backward()
    assert loss == pytest.approx(loss_ref, rel=1e-9)
    np.testing.assert_allclose(g_fin, g_ref, rtol=1e-8, atol=1e-10)
    np.testing.assert_allclose(grad, grad_ref, rtol=1e-6, atol=1e-8 * np.abs(grad_ref).max())


def test_trajectory_gradient_by_central_differences_on_gpu():
    from genesis_amd.universe import GpuUniverse
    ph, st, keep = system(n=1728, seed=9)
    th0 = O.theta_of(ph)
    gr = (60, 0.4 * ph.box[0], 0.05)
    T, C, win = 40, 10, 20
    say(f"central differences T={T} C={C} n={st.n}")
    with GpuUniverse(ph, st) as u:
        _, g0 = u.adjoint_forward(th0, keep, T, C, win, gr=gr)
        th = th0 * (1.0 + 0.01 * np.cos(np.arange(30)))
        u.upload(st)
        loss, _ = u.adjoint_forward(th, keep, T, C, win, gr=gr, g0=g0)
        grad = u.adjoint_backward()
        for k in [0, 4, 6, 12, 15, 19, 24, 25]:
            h = 1e-6 * abs(th[k])
            tp, tm = th.copy(), th.copy()
            tp[k] += h
            tm[k] -= h
            u.upload(st)
            lp, _ = u.adjoint_forward(tp, keep, T, C, win, gr=gr, g0=g0)
            u.upload(st)
            lm, _ = u.adjoint_forward(tm, keep, T, C, win, gr=gr, g0=g0)
            num = (lp - lm) / (2 * h)
            say(f"theta[{k}]: adjoint {grad[k]:.10e} fd {num:.10e}")
            assert grad[k] == pytest.approx(num, rel=3e-4, abs=1e-6 * np.abs(grad).max())


def test_torch_autograd_function():
    import torch
    from genesis_amd.adjoint import theta_of, trajectory_loss
    from genesis_amd.universe import GpuUniverse
    ph, st, keep = system()
    with GpuUniverse(ph, st) as u:
        th = (theta_of(ph) * 1.01).requires_grad_(True)
        loss = trajectory_loss(th, u, keep, 10, 5, 3)
        loss.backward()
        u.upload(st)
        l2, _ = u.adjoint_forward(th.detach().numpy(), keep, 10, 5, 3)
        g2 = u.adjoint_backward()
    assert loss.item() == l2
    assert np.array_equal(th.grad.numpy(), g2)
