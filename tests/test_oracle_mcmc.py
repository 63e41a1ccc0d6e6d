"""Pins for the oracle's Metropolis-Hastings chain (CPU only): a hand-worked scenario with supplied draws
and the stationary distribution of a known density."""
import numpy as np

from oracle import mcmc as om


def test_hand_worked_chain():
    # two walkers, one electron in 1D-ish use of 3D; log|psi|^2 = -|r|^2
    f = lambda r: -(r ** 2).sum(axis=(1, 2))
    r0 = np.array([[[1.0, 0.0, 0.0]], [[0.0, 0.0, 0.0]]], np.float32)
    st = om.ChainState(r=r0, log_psi_sqr=np.full(2, -1000.0, np.float32), walker_age=np.array([0, 0]), stepsize=0.5)
    # step 1: walker 0 moves toward the origin (always accepted), walker 1 moves out by 2 with u = 0.999
    #         (p = exp(-4) < u: rejected -> age 1)
    # step 2: walker 0 moves out by 4 (p = exp(-16) < u = 0.5: rejected), walker 1 same proposal as step 1:
    #         rejected by ratio but forced because its OLD age 1 >= max_age 1
    normals = np.zeros((2, 2, 1, 3), np.float32)
    normals[0, 0, 0, 0] = -2.0   # 1.0 + 0.5 * -2 = 0.0
    normals[0, 1, 0, 0] = 4.0    # 0.0 + 0.5 * 4 = 2.0
    normals[1, 0, 0, 0] = 8.0    # 0.0 + 0.5 * 8 = 4.0
    normals[1, 1, 0, 0] = 4.0
    uniforms = np.array([[0.5, 0.999], [0.5, 0.999]], np.float32)
    out, accs, masks = om.mcmc_steps(f, st, normals, uniforms, max_age=1, update_interval=2, target=0.5)
    np.testing.assert_array_equal(masks, [[True, False], [False, True]])
    np.testing.assert_allclose(accs, [0.5, 0.5])
    np.testing.assert_array_equal(out.walker_age, [1, 0])
    np.testing.assert_allclose(out.r[:, 0, 0], [0.0, 2.0])
    np.testing.assert_allclose(out.log_psi_sqr, [0.0, -4.0])
    # acc_rate: 0 -> 0.05 -> 0.095; at step 2 (interval 2) the stepsize uses the OLD rate 0.05 < 0.5: /1.05
    assert out.step_nr == 2
    assert abs(out.acc_rate - 0.095) < 1e-6
    assert abs(out.stepsize - 0.5 / 1.05) < 1e-6


def test_stepsize_clip_and_increase():
    assert abs(om.adjust_stepsize(0.99, 0.9, 0.5, 1e-2, 1.0) - 1.0) < 1e-12   # *1.05 then clipped to 1.0
    assert abs(om.adjust_stepsize(0.0101, 0.1, 0.5, 1e-2, 1.0) - 1e-2) < 1e-12  # /1.05 then clipped to 1e-2
    assert abs(om.adjust_stepsize(0.5, 0.5, 0.5, 1e-2, 1.0) - 0.525) < 1e-12   # not below target: *1.05


def test_stationary_distribution_hydrogen_1s():
    # |psi|^2 = exp(-2 r): <r> = 3/2, <r^2> = 3 (hydrogen 1s)
    f = lambda r: -2.0 * np.linalg.norm(r[:, 0, :], axis=-1)
    rng = np.random.default_rng(0)
    B, S = 4000, 400
    st = om.ChainState(r=rng.standard_normal((B, 1, 3)).astype(np.float64), log_psi_sqr=np.full(B, -1000.0),
                       walker_age=np.zeros(B, np.int64), stepsize=0.7)
    normals = rng.standard_normal((S, B, 1, 3))
    uniforms = rng.uniform(size=(S, B))
    out, accs, _ = om.mcmc_steps(f, st, normals, uniforms, max_age=1000, update_interval=10 ** 9, dtype=np.float64)
    rr = np.linalg.norm(out.r[:, 0], axis=-1)
    # walkers are (nearly) independent after 400 steps; 5 sigma bands
    assert abs(rr.mean() - 1.5) < 5 * np.sqrt(0.75 / B)
    assert abs((rr ** 2).mean() - 3.0) < 5 * np.sqrt(((rr ** 4).mean() - 9.0) / B)
