"""Pins for the oracle's features, potential energy and determinant sum (CPU only).

Each check is independent of the oracle's own formulas: brute-force Python loops, the Leibniz
expansion of the determinant, and closed-form values.
"""
import itertools
import math

import numpy as np
import torch

from oracle import hamiltonian as oh
from oracle import wavefunction as owf


def test_el_ion_features_brute_force():
    rng = np.random.default_rng(0)
    r = rng.standard_normal((2, 3, 3))
    R = rng.standard_normal((2, 3))
    diff, dist, feat = owf.el_ion_features(torch.tensor(r), torch.tensor(R))
    for b in range(2):
        for i in range(3):
            for I in range(2):
                d = [r[b, i, c] - R[I, c] for c in range(3)]
                rho = math.sqrt(sum(x * x for x in d))
                assert abs(float(dist[b, i, I]) - rho) < 1e-12
                np.testing.assert_allclose(feat[b, i, I].numpy(), [rho] + d, atol=1e-12)


def test_el_el_features_brute_force_and_zero_diagonal():
    rng = np.random.default_rng(1)
    r = torch.tensor(rng.standard_normal((4, 3)), requires_grad=True)
    diff, dist, feat = owf.el_el_features(r)
    rn = r.detach().numpy()
    for i in range(4):
        for j in range(4):
            d = rn[j] - rn[i]                      # r_j - r_i
            rho = 0.0 if i == j else float(np.sqrt((d ** 2).sum()))
            np.testing.assert_allclose(feat[i, j].detach().numpy(), [rho, *d], atol=1e-12)
    # the diagonal distance has an exactly zero gradient (no 0/0 from the norm)
    g, = torch.autograd.grad(dist.diagonal().sum(), r)
    assert torch.all(g == 0)


def test_potential_energy_brute_force_and_h2():
    rng = np.random.default_rng(2)
    r = rng.standard_normal((3, 4, 3))
    R = np.array([[0.0, 0.0, 0.0], [1.4, 0.0, 0.0], [0.0, 2.0, 0.5]])
    Z = np.array([1, 2, 3])
    e = oh.potential_energy(torch.tensor(r), torch.tensor(R), torch.tensor(Z)).numpy()
    for b in range(3):
        ref = 0.0
        for i in range(4):
            for j in range(i + 1, 4):
                ref += 1 / np.linalg.norm(r[b, i] - r[b, j])
            for I in range(3):
                ref -= Z[I] / np.linalg.norm(r[b, i] - R[I])
        for I in range(3):
            for J in range(I + 1, 3):
                ref += Z[I] * Z[J] / np.linalg.norm(R[I] - R[J])
        assert abs(e[b] - ref) < 1e-10
    # H2 at 1.4 bohr with the electrons at the two nuclei's midplane points (analytic numbers)
    R2 = np.array([[0.0, 0.0, 0.0], [1.4, 0.0, 0.0]])
    r2 = np.array([[[0.7, 1.0, 0.0], [0.7, -1.0, 0.0]]])
    d_ei = math.sqrt(0.49 + 1.0)
    expected = 1 / 2.0 - 4 / d_ei + 1 / 1.4
    assert abs(float(oh.potential_energy(torch.tensor(r2), torch.tensor(R2), torch.tensor([1, 1]))[0]) - expected) < 1e-12


def _leibniz_det(M):
    n = M.shape[0]
    tot = 0.0
    for perm in itertools.permutations(range(n)):
        inv = sum(1 for a in range(n) for b in range(a + 1, n) if perm[a] > perm[b])
        prod = 1.0
        for i in range(n):
            prod *= M[i, perm[i]]
        tot += (-1) ** inv * prod
    return tot


def test_sum_of_determinants_leibniz():
    rng = np.random.default_rng(3)
    for n, D, u in [(2, 3, 1), (3, 2, 2), (4, 4, 2)]:
        mo = rng.standard_normal((5, D, n, n)) * 2.0
        ph, lp = owf.sum_of_determinants(torch.tensor(mo[..., :u, :]), torch.tensor(mo[..., u:, :]))
        for b in range(5):
            psi = sum(_leibniz_det(mo[b, d]) for d in range(D))
            assert abs(float(lp[b]) - 2 * math.log(abs(psi) + 1e-8 * math.exp(
                max(math.log(abs(_leibniz_det(mo[b, d]))) for d in range(D))))) < 1e-9
            assert float(ph[b]) == (math.pi if psi < 0 else 0.0)


def test_log_epsilon_is_inside_the_shifted_log():
    # two determinants that cancel: psi_tilde = 0 -> log_psi_sqr = 2 (log(1e-8) + m)
    m = torch.tensor([[[[2.0]], [[-2.0]]]], dtype=torch.float64)
    ph, lp = owf.sum_of_determinants(m, m[..., :0, :])
    assert abs(float(lp) - 2 * (math.log(1e-8) + math.log(2.0))) < 1e-12
