"""Pins for the oracle's wavefunction and local energy (CPU only): exact invariances of the model, a
closed-form He-like wavefunction, finite differences and an independent autodiff route for the Laplacian."""
import math

import numpy as np
import pytest
import torch

from deeperwin_amd.params import EMB, ENV, ModelShape, init_params
from oracle import hamiltonian as oh
from oracle import wavefunction as owf


def _small(n=3, u=2, I=2, D=3, w1=16, w2=8, seed=0):
    shape = ModelShape(n, u, I, D, width1=w1, width2=w2)
    p = init_params(shape, seed, dtype=np.float64)
    # non-zero biases and envelopes so that every code path carries signal
    rng = np.random.default_rng(seed + 100)
    for path, leaves in p.items():
        for k in leaves:
            if k == "b":
                leaves[k] = rng.normal(0, 0.3, leaves[k].shape)
            if k.startswith("alpha") or k.startswith("weights"):
                leaves[k] = rng.uniform(0.5, 1.5, leaves[k].shape)
    return shape, p


def _lp(p, shape, r, R):
    return owf.log_psi_sqr(p, shape.n_up, shape.n_dn, torch.as_tensor(r), torch.as_tensor(R))


def test_same_spin_swap_flips_sign_keeps_density():
    shape, p = _small(n=4, u=2)
    rng = np.random.default_rng(0)
    R = rng.standard_normal((2, 3))
    r = rng.standard_normal((6, 4, 3))
    ph, lp = _lp(p, shape, r, R)
    for a, b in [(0, 1), (2, 3)]:
        rs = r.copy()
        rs[:, [a, b]] = rs[:, [b, a]]
        ph2, lp2 = _lp(p, shape, rs, R)
        np.testing.assert_allclose(lp2.numpy(), lp.numpy(), atol=1e-10)
        assert torch.all(torch.abs(torch.abs(ph2 - ph) - math.pi) < 1e-12)


def test_joint_translation_invariance():
    shape, p = _small()
    rng = np.random.default_rng(1)
    R = rng.standard_normal((2, 3))
    r = rng.standard_normal((5, 3, 3))
    t = np.array([0.3, -1.2, 2.5])
    _, lp = _lp(p, shape, r, R)
    _, lp2 = _lp(p, shape, r + t, R + t)
    np.testing.assert_allclose(lp2.numpy(), lp.numpy(), atol=1e-10)


def test_embedding_is_permutation_equivariant_within_spin():
    shape, p = _small(n=4, u=2)
    rng = np.random.default_rng(2)
    R = torch.tensor(rng.standard_normal((2, 3)))
    r = torch.tensor(rng.standard_normal((4, 3)))
    perm = [1, 0, 3, 2]

    def emb(x):
        _, _, f_ei = owf.el_ion_features(x, R)
        _, _, f_ee = owf.el_el_features(x)
        return owf.ferminet_embedding(p, f_ei.reshape(4, -1), f_ee, 2, 4)

    np.testing.assert_allclose(emb(r[perm]).numpy(), emb(r)[perm].numpy(), atol=1e-12)


def he_like_params(Z_k_alpha=0.7, a=1.3, b=-0.4, c=0.2, d=0.9, w1=16, w2=8):
    """Zero weight matrices with non-zero biases make the embedding a constant vector e; the backflow
    weights are chosen so that e W_up = (a, b) and e W_dn = (c, d)."""
    shape = ModelShape(2, 1, 1, 1, width1=w1, width2=w2)
    p = init_params(shape, 0, dtype=np.float64)
    rng = np.random.default_rng(9)
    for path, leaves in p.items():
        if "linear_0" in path and "bf_" not in path:
            leaves["w"] = np.zeros_like(leaves["w"])
            leaves["b"] = rng.normal(0, 0.5, leaves["b"].shape)
    e = np.tanh(p[f"{EMB}/h_el_3/linear_0"]["b"])
    p[f"{ENV}/bf_up/linear_0"]["w"] = np.outer(e / (e @ e), [a, b])
    p[f"{ENV}/bf_dn/linear_0"]["w"] = np.outer(e / (e @ e), [c, d])
    for k in ("alpha_up", "alpha_dn"):
        p[ENV][k] = np.full((1, 2), Z_k_alpha)
    for k in ("weights_up", "weights_dn"):
        p[ENV][k] = np.full((1, 2), 0.8)
    return shape, p


def test_he_like_closed_form_psi_and_local_energy():
    Zc = 2.0
    alpha = 0.7
    k = math.log1p(math.exp(alpha))
    shape, p = he_like_params(alpha)
    rng = np.random.default_rng(4)
    r = rng.standard_normal((8, 2, 3)) * 0.8
    R = np.zeros((1, 3))
    r1 = np.linalg.norm(r[:, 0], axis=-1)
    r2 = np.linalg.norm(r[:, 1], axis=-1)
    r12 = np.linalg.norm(r[:, 0] - r[:, 1], axis=-1)
    psi = (1.3 * 0.9 - (-0.4) * 0.2) * 0.8 ** 2 * np.exp(-k * (r1 + r2))
    _, lp = _lp(p, shape, r, R)
    # log_psi_sqr = 2 (log(|psi e^-m| + eps) + m) with a single det: 2 log(|psi| + eps e^m) = 2 log(|psi| (1 + eps))
    np.testing.assert_allclose(lp.numpy(), 2 * np.log(np.abs(psi) * (1 + 1e-8)), atol=1e-10)
    e_loc = oh.local_energy(p, 1, 1, r, R, np.array([Zc])).numpy()
    expected = -k ** 2 + (k - Zc) * (1 / r1 + 1 / r2) + 1 / r12
    np.testing.assert_allclose(e_loc, expected, rtol=1e-9, atol=1e-9)


def test_laplacian_central_finite_differences():
    shape, p = _small(n=3, u=2, I=2, D=2)
    rng = np.random.default_rng(5)
    R = rng.standard_normal((2, 3))
    r = rng.standard_normal((2, 3, 3))
    ekin = oh.kinetic_energy(p, 2, 1, torch.tensor(r), torch.tensor(R)).numpy()
    h = 1e-4
    for b in range(2):
        x0 = r[b].reshape(-1)

        def f(x):
            return float(_lp(p, shape, x.reshape(1, 3, 3), R)[1][0])

        f0 = f(x0)
        lap = 0.0
        g2 = 0.0
        for kk in range(9):
            e = np.zeros(9)
            e[kk] = h
            fp, fm = f(x0 + e), f(x0 - e)
            lap += (fp - 2 * f0 + fm) / h ** 2
            g2 += ((fp - fm) / (2 * h)) ** 2
        ek_fd = -0.5 * (0.25 * g2 + 0.5 * lap)
        assert abs(ek_fd - ekin[b]) < 2e-5 * max(1.0, abs(ekin[b]))


def test_laplacian_matches_reverse_over_reverse_hessian():
    shape, p = _small(n=3, u=1, I=1, D=2)
    rng = np.random.default_rng(6)
    R = rng.standard_normal((1, 3))
    r = rng.standard_normal((3, 3))

    def f(x):
        return owf.log_psi_sqr(p, 1, 2, x.reshape(3, 3), torch.tensor(R))[1]

    x = torch.tensor(r.reshape(-1))
    H = torch.autograd.functional.hessian(f, x)
    g = torch.autograd.functional.jacobian(f, x)
    ek = -0.5 * (0.25 * (g ** 2).sum() + 0.5 * torch.trace(H))
    ek_hvp = oh.kinetic_energy(p, 1, 2, torch.tensor(r[None]), torch.tensor(R))[0]
    assert abs(float(ek) - float(ek_hvp)) < 1e-10 * max(1.0, abs(float(ek)))


def test_local_energy_chunking_is_invisible():
    shape, p = _small()
    rng = np.random.default_rng(7)
    R = rng.standard_normal((2, 3))
    r = rng.standard_normal((5, 3, 3))
    Z = np.array([1, 2])
    a = oh.local_energy(p, 2, 1, r, R, Z, max_batch_size=64).numpy()
    b = oh.local_energy(p, 2, 1, r, R, Z, max_batch_size=2).numpy()
    np.testing.assert_allclose(a, b, atol=1e-12)


def test_mean_and_var_fsum():
    e = np.array([1.0, 2.0, np.nan, 4.0])
    m, v = oh.mean_and_var(e)
    assert abs(m - math.fsum([1.0, 2.0, 4.0]) / 3) < 1e-15
    assert abs(v - math.fsum([(x - m) ** 2 for x in (1.0, 2.0, 4.0)]) / 3) < 1e-15
    # nanmean drops NaN only: an infinite local energy makes the mean infinite (evaluation.py:148-150)
    m, v = oh.mean_and_var(np.array([1.0, np.inf, np.nan]))
    assert m == np.inf


def test_pinning_table_names_existing_tests():
    import ast
    import os

    import oracle

    here = os.path.dirname(os.path.abspath(__file__))
    for fn, pins in oracle.PINNING.items():
        for pin in pins:
            if "::" not in pin:
                assert pin.startswith("not a reference function"), (fn, pin)
                continue
            fname, rest = pin.split("::", 1)
            test = rest.split(" ")[0]
            tree = ast.parse(open(os.path.join(here, fname)).read())
            names = {n.name for n in ast.walk(tree) if isinstance(n, ast.FunctionDef)}
            assert test in names, (fn, pin)


def test_vmapped_kinetic_energy_equals_walker_loop():
    """The chunk-batched kinetic energy (vmap over walkers, the reference's batched_vmap) equals the walker
    loop; one determinant falls back to the loop."""
    from workloads import system, walkers

    for name, D in (("lih", 3), ("he", 1)):
        shape, params, R, Z, mapping = system(name, n_det=D, seed=2)
        r = torch.tensor(walkers(R, mapping, 5, seed=4).astype(np.float64))
        a = oh.kinetic_energy(params, shape.n_up, shape.n_dn, r, torch.tensor(R))
        b = oh.kinetic_energy_vmapped(params, shape.n_up, shape.n_dn, r, torch.tensor(R))
        np.testing.assert_allclose(b.numpy(), a.numpy(), rtol=1e-11, atol=1e-11)
