"""Batched replicas (gs_batch_*) against the same replicas run one at a time
(gs_step), which tests/test_gpu_parity.py pins to the oracle."""
import numpy as np
import pytest

from genesis_amd.inputs import mixture, replica_sweep
from genesis_amd.params import GRID_MINIMAGE, Physics

pytestmark = pytest.mark.gpu


def small_sweep():
    sizes = [512, 700, 1500, 4096, 900, 3000]
    out = []
    for r, n in enumerate(sizes):
        w = mixture(seed=3000 + r, n=n, temperature=20.0 + 20.0 * r, steps=30)
        w.physics = Physics(**{**w.physics.__dict__, "grid_semantics": GRID_MINIMAGE})
        out.append(w)
    return out


def test_batch_matches_replicas_run_alone():
    from genesis_amd.batch import GpuBatch
    from genesis_amd.universe import GpuUniverse
    ws = small_sweep()
    print("[gpu-batch] sizes", [w.state.n for w in ws], "dims", [w.physics.grid_dims() for w in ws], flush=True)
    with GpuBatch([w.physics for w in ws], [w.state for w in ws]) as b:
        obs_b = b.step(30, observe=True)
        got = [b.download(r) for r in range(len(ws))]
    for r, w in enumerate(ws):
        with GpuUniverse(w.physics, w.state) as u:
            u.set_tuning(0)
            obs = u.step(30)
            ref = u.download()
        np.testing.assert_allclose(obs_b[:, r, 5], obs[:, 5], rtol=1e-9)
        np.testing.assert_allclose(obs_b[:, r, 4], obs[:, 4], rtol=1e-12)
        np.testing.assert_allclose(got[r].x, ref.x, rtol=0, atol=1e-9)
        np.testing.assert_allclose(got[r].vx, ref.vx, rtol=0, atol=1e-9 * np.abs(ref.vx).max())


def test_graph_replay_is_bitwise_identical_to_plain_launches():
    from genesis_amd.batch import GpuBatch
    ws = small_sweep()
    res = []
    for seg in (0, 10):
        with GpuBatch([w.physics for w in ws], [w.state for w in ws]) as b:
            obs = b.step(40, graph_segment=seg, observe=True)
            b.step(20, graph_segment=seg)
            res.append((obs, [b.download(r) for r in range(len(ws))]))
    assert np.array_equal(res[0][0], res[1][0])
    for r in range(len(ws)):
        for a in ("x", "y", "z", "vx", "ax"):
            assert np.array_equal(getattr(res[0][1][r], a), getattr(res[1][1][r], a))


def test_batch_autotune_restores_state():
    from genesis_amd.batch import GpuBatch
    ws = small_sweep()[:3]
    with GpuBatch([w.physics for w in ws], [w.state for w in ws]) as b:
        before = [b.download(r) for r in range(3)]
        t = b.autotune(6)
        after = [b.download(r) for r in range(3)]
        assert t["cell_scale"] in (1.0, 1.15, 1.3) and t["ms_per_tick"] > 0
        for r in range(3):
            np.testing.assert_array_equal(np.sort(before[r].x), np.sort(after[r].x))
        obs = b.step(10, observe=True)
        assert np.isfinite(obs).all()


def test_temperature_sweep_workload_matches_the_oracle():
    # The 64-replica sweep in one batch: the first tick starts every replica
    # at its own target (20 K .. 120 K), and whole-replica trajectories of
    # the smallest replicas (including one on a 2-cell grid) match the oracle
    # run of that replica alone.
    from genesis_amd.batch import GpuBatch
    from oracle import oracle as O
    ws = replica_sweep(n_rep=64, steps=10)
    with GpuBatch([w.physics for w in ws], [w.state for w in ws]) as b:
        obs = b.step(20, graph_segment=10, observe=True)
        fin = {r: b.download(r) for r in (1, 2)}
    assert obs.shape == (20, 64, 6) and np.isfinite(obs).all()
    targets = np.array([w.physics.thermostat_temperature for w in ws])
    assert np.corrcoef(obs[0, :, 3], targets)[0, 1] > 0.99
    for r in (1, 2):
        ref, obs_ref = O.run(ws[r].physics, ws[r].state, 20, nthreads=8)
        print(f"[gpu-batch] replica {r} n={ws[r].state.n} dims={ws[r].physics.grid_dims()}", flush=True)
        np.testing.assert_allclose(obs[:, r, 5], obs_ref[:, 5], rtol=1e-9)
        np.testing.assert_allclose(fin[r].x, ref.x, rtol=0, atol=1e-8)
