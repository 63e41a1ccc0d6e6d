"""Pins of the loss oracle (oracle/loss.py) on the CPU.

* Clipping statistics on a hand-worked set E = [-1, 0, 1, 10, NaN] (NaN dropped): mean 2.5, median 0.5,
  std sqrt(19.25), mean absolute deviation 3.75; tanh and hard clips at those values
  (optimization/loss_function.py:19-60).
* The parameter gradient sum_b c_b d log_psi_sqr(r_b) / d theta against central finite differences of the
  float64 log_psi_sqr on a small model, one entry of every kind of leaf (1-el and 2-el weights and biases,
  backflow, envelope alpha and weights): autograd through the restatement and the difference quotient agree.
"""
import math

import numpy as np
import pytest
import torch

from oracle import loss as ol
from oracle import wavefunction as owf
from workloads import system, walkers

E = np.array([-1.0, 0.0, 1.0, 10.0, np.nan])


def test_clipping_center_and_width_known_answers():
    assert ol.clipping_center_and_width(E, "mean", "std", 1.0) == pytest.approx((2.5, math.sqrt(19.25)))
    assert ol.clipping_center_and_width(E, "mean", "mae", 1.0) == pytest.approx((2.5, 3.75))
    c, w = ol.clipping_center_and_width(E, "median", "mae", 5.0)
    assert c == pytest.approx(0.5) and w == pytest.approx(5.0 * (1.5 + 0.5 + 0.5 + 9.5) / 4)


def test_clip_functions_known_answers():
    hard = ol.clip(E, 2.5, 3.0, "hard")
    np.testing.assert_allclose(hard[:4], [-0.5, 0.0, 1.0, 5.5])
    assert np.isnan(hard[4])
    t = ol.clip(np.array([2.5, 2.5 + 3.0]), 2.5, 3.0, "tanh")
    np.testing.assert_allclose(t, [2.5, 2.5 + 3.0 * math.tanh(1.0)])
    # the initial state (0, 1e12) leaves energies unchanged to rounding
    Ec, (c, w) = ol.clip_energies(E[:4], (0.0, 1e12))
    np.testing.assert_allclose(Ec, E[:4], rtol=0, atol=1e-9)
    assert c == pytest.approx(2.5) and w == pytest.approx(5 * math.sqrt(19.25))


def test_param_grad_matches_finite_differences():
    shape, params, R, Z, mapping = system("lih", n_det=2, seed=3)
    r = walkers(R, mapping, 3, seed=5).astype(np.float64)
    c = np.array([0.7, -1.3, 0.4])
    g = ol.param_grad(params, shape.n_up, shape.n_dn, r, R, c)

    def f(p):
        lp = owf.log_psi_sqr(p, shape.n_up, shape.n_dn, torch.tensor(r), R)[1].numpy()
        return float(np.dot(c, lp))

    rng = np.random.default_rng(0)
    checked = 0
    for path, leaves in params.items():
        for leaf, val in leaves.items():
            idx = tuple(int(rng.integers(s)) for s in np.shape(val))
            h = 1e-5
            p1 = {k: {kk: np.array(vv, dtype=np.float64) for kk, vv in v.items()} for k, v in params.items()}
            p2 = {k: {kk: np.array(vv, dtype=np.float64) for kk, vv in v.items()} for k, v in params.items()}
            p1[path][leaf][idx] += h
            p2[path][leaf][idx] -= h
            fd = (f(p1) - f(p2)) / (2 * h)
            assert g[path][leaf][idx] == pytest.approx(fd, rel=1e-5, abs=1e-7), (path, leaf, idx)
            checked += 1
    assert checked >= 16
