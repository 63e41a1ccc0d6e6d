"""Pins for the oracle's layouts (CPU only): the block order of the symmetric features, the ion-major
el-ion concat of h_one, and the (det, orbital) column order of backflow and envelope, each with n_up != n_dn,
n_ion >= 2 and n_det >= 2 so that a swapped block or a transposed reshape changes the result.

Three independent routes fix them:
* ``tests/bruteforce.py``: a scalar per-index restatement read off the reference lines (loops only);
* a closed-form two-determinant, two-centre wavefunction whose log|psi|^2 and local energy follow from
  the mathematics alone (zero embedding weights, chosen backflow columns, distinct envelopes per column);
* known answers (N2 ion-ion repulsion 49 / 2.07).
The last test plants the layout mistakes in the oracle and checks that these pins catch each of them.
"""
import math

import numpy as np
import pytest
import torch

import bruteforce as bf
from deeperwin_amd.params import EMB, ENV, ModelShape, init_params
from oracle import hamiltonian as oh
from oracle import wavefunction as owf

CASES = [  # (n, n_up, n_ion, n_det, width1, width2, n_iter)
    (3, 2, 2, 2, 5, 3, 3),
    (4, 3, 3, 3, 6, 3, 4),
    (3, 1, 2, 2, 4, 2, 2),
]


def _random_params(shape: ModelShape, seed: int):
    p = init_params(shape, seed, dtype=np.float64)
    rng = np.random.default_rng(seed + 1000)
    for path, leaves in p.items():
        for k in leaves:
            if k == "b":
                leaves[k] = rng.normal(0, 0.4, leaves[k].shape)
            elif k.startswith("alpha"):
                leaves[k] = rng.uniform(-0.5, 1.0, leaves[k].shape)
            elif k.startswith("weights"):
                leaves[k] = rng.uniform(0.3, 1.7, leaves[k].shape)
            else:
                leaves[k] = leaves[k] * rng.uniform(0.8, 1.6)
    return p


def _config(case, seed):
    n, u, I, D, w1, w2, L = case
    shape = ModelShape(n, u, I, D, width1=w1, width2=w2, n_iter=L)
    rng = np.random.default_rng(seed)
    R = rng.normal(0, 1.2, (I, 3))
    r = rng.normal(0, 1.0, (3, n, 3))
    return shape, _random_params(shape, seed), r, R


def test_symmetric_features_brute_force():
    rng = np.random.default_rng(0)
    n, u, F, E = 5, 3, 3, 2
    h = rng.standard_normal((2, n, F))
    hee = rng.standard_normal((2, n, n, E))
    got = owf.symmetric_features(torch.tensor(h), torch.tensor(hee), u).numpy()
    assert got.shape == (2, n, 3 * F + 2 * E)
    for b in range(2):
        hl = h[b].tolist()
        heel = hee[b].tolist()
        for i in range(n):
            np.testing.assert_allclose(got[b, i], bf.symmetric_row(hl, heel, i, u), atol=1e-14)


def test_el_ion_rows_are_ion_major():
    rng = np.random.default_rng(1)
    r = rng.standard_normal((4, 3))
    R = rng.standard_normal((3, 3))
    _, _, f = owf.el_ion_features(torch.tensor(r), torch.tensor(R))
    h0 = f.reshape(4, -1).numpy()          # the reshape log_psi_sqr applies before the embedding
    for i in range(4):
        np.testing.assert_allclose(h0[i], bf.el_ion_row(r[i].tolist(), R.tolist()), atol=1e-14)
        # feature f of ion I sits at column 4 I + f
        assert abs(h0[i, 4 * 2 + 0] - np.linalg.norm(r[i] - R[2])) < 1e-14
        assert abs(h0[i, 4 * 1 + 3] - (r[i, 2] - R[1, 2])) < 1e-14


@pytest.mark.parametrize("case", CASES)
def test_slater_matrices_element_by_element(case):
    shape, p, r, R = _config(case, 3)
    got = owf.slater_matrices(p, shape.n_up, shape.n_dn, r, R).numpy()    # [B, D, n, n]
    for b in range(r.shape[0]):
        ref = bf.slater_matrices(p, r[b].tolist(), R.tolist(), shape.n_up, shape.n_det)
        np.testing.assert_allclose(got[b], np.array(ref), rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("case", CASES)
def test_log_psi_sqr_brute_force(case):
    shape, p, r, R = _config(case, 4)
    ph, lp = owf.log_psi_sqr(p, shape.n_up, shape.n_dn, torch.tensor(r), R)
    for b in range(r.shape[0]):
        ph_ref, lp_ref = bf.log_psi_sqr(p, r[b].tolist(), R.tolist(), shape.n_up, shape.n_det)
        assert abs(float(lp[b]) - lp_ref) < 1e-10 * max(1.0, abs(lp_ref))
        assert float(ph[b]) == ph_ref


def _two_det_params(w1=6, w2=3):
    """n = 2 (1 up, 1 down), 2 ions, 2 dets.  All embedding weights are zero and the biases are not, so the
    embedding is a constant vector e for both electrons.  The backflow weights are outer(e / |e|^2, beta),
    so bf[c] = beta[c] for every column; alpha and weights differ per (ion, column)."""
    shape = ModelShape(2, 1, 2, 2, width1=w1, width2=w2, n_iter=3)
    p = init_params(shape, 0, dtype=np.float64)
    rng = np.random.default_rng(11)
    for path, leaves in p.items():
        if "linear_0" in path and "bf_" not in path:
            leaves["w"] = np.zeros_like(leaves["w"])
            leaves["b"] = rng.normal(0, 0.6, leaves["b"].shape)
    e = np.tanh(p[f"{EMB}/h_el_2/linear_0"]["b"])
    beta = {"up": np.array([1.3, -0.7, 0.4, 1.1]), "dn": np.array([0.9, 0.5, -1.2, 0.6])}
    for s in ("up", "dn"):
        p[f"{ENV}/bf_{s}/linear_0"]["w"] = np.outer(e / (e @ e), beta[s])
        p[ENV][f"alpha_{s}"] = rng.uniform(-0.3, 1.2, (2, 4))
        p[ENV][f"weights_{s}"] = rng.uniform(0.4, 1.5, (2, 4))
    return shape, p, beta


def _phi(p, s, c, x, R):
    """Envelope of column c for spin s at x, its Laplacian: sum_I w e^{-k rho}, sum_I w (k^2 - 2k/rho) e^{-k rho}."""
    val = lap = 0.0
    for I in range(R.shape[0]):
        k = math.log1p(math.exp(p[ENV][f"alpha_{s}"][I, c]))
        w = p[ENV][f"weights_{s}"][I, c]
        rho = float(np.linalg.norm(x - R[I]))
        val += w * math.exp(-k * rho)
        lap += w * (k * k - 2 * k / rho) * math.exp(-k * rho)
    return val, lap


def test_two_det_two_centre_closed_form():
    """psi = sum_d [b_up(2d) b_dn(2d+1) phi_up,2d(r1) phi_dn,2d+1(r2) - b_up(2d+1) b_dn(2d) phi_up,2d+1(r1)
    phi_dn,2d(r2)] with column c = d n + k; E_kin = -1/2 (lap_1 psi + lap_2 psi) / psi."""
    shape, p, beta = _two_det_params()
    R = np.array([[0.0, 0.0, 0.0], [1.6, 0.3, -0.2]])
    Z = np.array([2.0, 1.0])
    rng = np.random.default_rng(12)
    r = rng.normal(0, 1.0, (6, 2, 3))
    psi = np.zeros(6)
    lap = np.zeros(6)
    dets = np.zeros((6, 2))
    for b in range(6):
        for d in range(2):
            for (k1, k2, sgn) in ((0, 1, 1.0), (1, 0, -1.0)):
                c1, c2 = d * 2 + k1, d * 2 + k2
                f1, l1 = _phi(p, "up", c1, r[b, 0], R)
                f2, l2 = _phi(p, "dn", c2, r[b, 1], R)
                amp = sgn * beta["up"][c1] * beta["dn"][c2]
                psi[b] += amp * f1 * f2
                dets[b, d] += amp * f1 * f2
                lap[b] += amp * (l1 * f2 + f1 * l2)
    _, lp = owf.log_psi_sqr(p, 1, 1, torch.tensor(r), R)
    m = np.log(np.abs(dets)).max(axis=1)     # log|psi|^2 = 2 (log(|psi| e^-m + 1e-8) + m), wavefunction.py:63-83
    np.testing.assert_allclose(lp.numpy(), 2 * (np.log(np.abs(psi) * np.exp(-m) + 1e-8) + m), rtol=0, atol=1e-11)
    e_pot = (1 / np.linalg.norm(r[:, 0] - r[:, 1], axis=-1)
             - sum(Z[I] / np.linalg.norm(r[:, i] - R[I], axis=-1) for i in range(2) for I in range(2))
             + Z[0] * Z[1] / np.linalg.norm(R[0] - R[1]))
    expected = -0.5 * lap / psi + e_pot
    got = oh.local_energy(p, 1, 1, r, R, Z).numpy()
    np.testing.assert_allclose(got, expected, rtol=1e-6, atol=1e-6)


def test_n2_ion_ion_repulsion():
    R = np.array([[0.0, 0.0, 0.0], [2.07, 0.0, 0.0]])
    Z = np.array([7, 7])
    r = np.random.default_rng(5).normal(0, 1, (2, 14, 3))
    e = oh.potential_energy(torch.tensor(r), torch.tensor(R), torch.tensor(Z)).numpy()
    el = np.zeros(2)
    for b in range(2):
        for i in range(14):
            for j in range(i + 1, 14):
                el[b] += 1 / np.linalg.norm(r[b, i] - r[b, j])
            for I in range(2):
                el[b] -= 7 / np.linalg.norm(r[b, i] - R[I])
    np.testing.assert_allclose(e - el, 49 / 2.07, rtol=0, atol=1e-11)


# ---- planted mistakes: each must make at least one pin above fail ----

def _swapped_spin_means(h, h_ee, n_up):
    n = h.shape[-2]
    up = h[..., :n_up, :].mean(dim=-2, keepdim=True).expand(*h.shape[:-2], n, h.shape[-1])
    dn = h[..., n_up:, :].mean(dim=-2, keepdim=True).expand(*h.shape[:-2], n, h.shape[-1])
    return torch.cat([h, dn, up, h_ee[..., :n_up, :].mean(-2), h_ee[..., n_up:, :].mean(-2)], dim=-1)


def _swapped_pair_means(h, h_ee, n_up):
    n = h.shape[-2]
    up = h[..., :n_up, :].mean(dim=-2, keepdim=True).expand(*h.shape[:-2], n, h.shape[-1])
    dn = h[..., n_up:, :].mean(dim=-2, keepdim=True).expand(*h.shape[:-2], n, h.shape[-1])
    return torch.cat([h, up, dn, h_ee[..., n_up:, :].mean(-2), h_ee[..., :n_up, :].mean(-2)], dim=-1)


def _orbital_major_columns(params, h, dist_ei, n_up, n_dn, n_det):
    """Envelope orbitals with columns read as (orbital, det) instead of (det, orbital)."""
    dt = h.dtype
    n = n_up + n_dn
    out = []
    for spin, sl in (("up", slice(0, n_up)), ("dn", slice(n - n_dn, n))):
        w = owf._leaf(params, f"bf_{spin}/linear_0", "w", dt)
        alpha = owf._leaf(params, "envelope_orbitals", f"alpha_{spin}", dt)
        weights = owf._leaf(params, "envelope_orbitals", f"weights_{spin}", dt)
        b = (h[..., sl, :] @ w).reshape(*h.shape[:-2], n_dn if spin == "dn" else n_up, n, n_det).movedim(-1, -3)
        d = dist_ei[..., sl, :, None]
        env = (weights * torch.exp(-torch.nn.functional.softplus(alpha) * d)).sum(dim=-2)
        env = env.reshape(*env.shape[:-1], n, n_det).movedim(-1, -3)
        out.append(env * b)
    return out[0], out[1]


def _pins_pass() -> bool:
    try:
        for case in CASES:
            test_slater_matrices_element_by_element(case)
            test_log_psi_sqr_brute_force(case)
        test_two_det_two_centre_closed_form()
    except AssertionError:
        return False
    return True


@pytest.mark.parametrize("target,mutant", [("symmetric_features", _swapped_spin_means),
                                           ("symmetric_features", _swapped_pair_means),
                                           ("envelope_orbitals", _orbital_major_columns)])
def test_planted_layout_mistakes_are_caught(monkeypatch, target, mutant):
    assert _pins_pass()
    monkeypatch.setattr(owf, target, mutant)
    assert not _pins_pass()
