"""Pins for the oracle's dpe4 model (oracle/wavefunction.py:dpe4_embedding; configuration.py:875-891,
ferminet_embedding.py:97-188, 207-267), CPU only:

* log|psi|^2 and phase equal the scalar restatement in tests/bruteforce.py (the reference's same/diff list
  split and block reshapes followed literally, Leibniz determinants) on tiny systems with non-zero biases;
* planted mistakes in the oracle (same/diff maps swapped, the ion table read without the Z_min offset,
  the self pair dropped from the convolution) fail that comparison;
* the oracle's kinetic energy of the dpe4 model equals central finite differences of its log|psi|^2;
* its parameter gradient (autograd) equals central finite differences on entries of every kind of dpe4 leaf.
"""
import numpy as np
import pytest
import torch

import bruteforce as bf
from deeperwin_amd.params import ModelShape, init_params
from oracle import hamiltonian as oh
from oracle import loss as ol
from oracle import wavefunction as owf


def _system(n=4, u=2, seed=0, W1=16, W2=8):
    Z = np.array([3, 1])
    R = np.array([[0.0, 0.0, 0.0], [1.6, 0.3, -0.4]])
    shape = ModelShape(n, u, 2, 2, W1, W2, 3, "dpe4", 1, 3)
    p = init_params(shape, seed, np.float64)
    rng = np.random.default_rng(seed + 100)
    for leaves in p.values():  # non-zero biases and envelopes, so that every term is exercised
        if "b" in leaves:
            leaves["b"] = rng.uniform(-0.5, 0.5, leaves["b"].shape)
        for k in ("alpha_up", "alpha_dn", "weights_up", "weights_dn"):
            if k in leaves:
                leaves[k] = rng.uniform(0.5, 1.5, leaves[k].shape)
    r = rng.standard_normal((3, n, 3)) + R[rng.integers(0, 2, (3, n))]
    return shape, p, R, Z, r


def _oracle(p, shape, r, R, Z):
    ph, lp = owf.log_psi_sqr(p, shape.n_up, shape.n_dn, torch.tensor(r), R, Z)
    return ph.numpy(), lp.numpy()


def _brute(p, shape, r, R, Z):
    out = []
    for b in range(r.shape[0]):
        rb = [list(map(float, x)) for x in r[b]]
        emb = bf.embedding_dpe4(p, rb, R.tolist(), Z.tolist(), 1, shape.n_up)
        out.append(bf.log_psi_sqr(p, rb, R.tolist(), shape.n_up, shape.n_det, emb=emb))
    return np.array([o[0] for o in out]), np.array([o[1] for o in out])


@pytest.mark.parametrize("n,u", [(3, 2), (4, 2), (5, 3)])
def test_dpe4_matches_scalar_restatement(n, u):
    shape, p, R, Z, r = _system(n, u, seed=n)
    ph, lp = _oracle(p, shape, r, R, Z)
    ph_b, lp_b = _brute(p, shape, r, R, Z)
    np.testing.assert_array_equal(ph, ph_b)
    np.testing.assert_allclose(lp, lp_b, rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize("mistake", ["same_diff_swapped", "no_z_min_offset", "self_pair_dropped"])
def test_planted_mistakes_fail(mistake, monkeypatch):
    shape, p, R, Z, r = _system(4, 2, seed=7)
    if mistake == "same_diff_swapped":
        real = owf._dense

        def dense(params, module, x):
            if module.endswith("w_same"):
                module = module[:-4] + "diff"
            elif module.endswith("w_diff"):
                module = module[:-4] + "same"
            return real(params, module, x)

        monkeypatch.setattr(owf, "_dense", dense)
    elif mistake == "no_z_min_offset":
        real_emb = owf.ion_embedding
        monkeypatch.setattr(owf, "ion_embedding", lambda params, Z_, z_min, dt: real_emb(params, Z_, 0, dt))
        Z = np.array([2, 1])  # the mutant reads rows 2 and 1 instead of 1 and 0, inside the 3-row table
    else:
        real_el_el = owf.el_el_features

        def el_el(r_):
            diff, dist, f = real_el_el(r_)
            n_ = r_.shape[-2]
            return diff, dist + 1e3 * torch.eye(n_, dtype=r_.dtype), f  # the self pair's map differs

        monkeypatch.setattr(owf, "el_el_features", el_el)
    _, lp = _oracle(p, shape, r, R, Z)
    _, lp_b = _brute(p, shape, r, R, Z)  # with Z_min = 1 the brute force reads the rows Z - 1
    assert np.abs(lp - lp_b).max() > 1e-6


def test_dpe4_kinetic_energy_finite_differences():
    shape, p, R, Z, r = _system(3, 2, seed=11)
    ekin = oh.kinetic_energy(p, 2, 1, torch.tensor(r[:2]), torch.tensor(R), Z=Z).numpy()
    h = 1e-4
    for b in range(2):
        x0 = r[b].reshape(-1)

        def f(x):
            return float(owf.log_psi_sqr(p, 2, 1, torch.tensor(x.reshape(1, 3, 3)), R, Z)[1][0])

        f0 = f(x0)
        lap, g2 = 0.0, 0.0
        for k in range(9):
            e = np.zeros(9)
            e[k] = h
            fp, fm = f(x0 + e), f(x0 - e)
            lap += (fp - 2 * f0 + fm) / h ** 2
            g2 += ((fp - fm) / (2 * h)) ** 2
        ek_fd = -0.5 * (0.25 * g2 + 0.5 * lap)
        assert abs(ek_fd - ekin[b]) < 2e-5 * max(1.0, abs(ekin[b]))


def test_dpe4_param_grad_finite_differences():
    shape, p, R, Z, r = _system(3, 2, seed=13)
    c = np.array([0.7, -1.3, 0.4])
    g = ol.param_grad(p, 2, 1, r, R, c, Z)

    def loss(params):
        lp = owf.log_psi_sqr(params, 2, 1, torch.tensor(r), R, Z)[1].numpy()
        return float(np.dot(c, lp))

    h = 1e-6
    rng = np.random.default_rng(0)
    checked = 0
    for path, leaves in p.items():
        if "envelope_orbitals" in path:
            continue
        for leaf, v in leaves.items():
            flat = v.reshape(-1)
            for idx in rng.choice(flat.size, size=min(2, flat.size), replace=False):
                q = {pp: {kk: vv.copy() for kk, vv in lv.items()} for pp, lv in p.items()}
                q[path][leaf].reshape(-1)[idx] += h
                fp = loss(q)
                q[path][leaf].reshape(-1)[idx] -= 2 * h
                fm = loss(q)
                fd = (fp - fm) / (2 * h)
                an = g[path][leaf].reshape(-1)[idx]
                assert abs(fd - an) <= 1e-5 * max(1.0, abs(an)), (path, leaf, idx, fd, an)
                checked += 1
    assert checked > 40
