"""The oracle against the committed regression fixtures in ``tests/golden`` (CPU only).

The fixtures hold the float64 oracle's outputs on seeded inputs (see ``tests/golden/make_golden.py``); a
change in the oracle, the seeded parameters, the geometries or the walker initialisation shows up here.
"""
import os

import numpy as np
import pytest
import torch

from golden.make_golden import chain_case, param_digest
from oracle import hamiltonian as oh
from oracle import wavefunction as owf
from workloads import system, walkers

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return dict(np.load(os.path.join(HERE, name), allow_pickle=False))


@pytest.mark.parametrize("name,fixture,B,seed", [("lih", "lih_wavefunction.npz", 8, 11),
                                                 ("n2", "n2_wavefunction.npz", 3, 12)])
def test_wavefunction_fixture(name, fixture, B, seed):
    g = _load(fixture)
    shape, params, R, Z, mapping = system(name)
    r = walkers(R, mapping, B, seed=seed)
    np.testing.assert_array_equal(r, g["r"])
    np.testing.assert_array_equal(np.asarray(mapping), g["mapping"])
    np.testing.assert_allclose(param_digest(params), g["param_digest"], rtol=1e-12)
    ph, lp = owf.log_psi_sqr(params, shape.n_up, shape.n_dn, torch.tensor(r, dtype=torch.float64), R)
    np.testing.assert_allclose(lp.numpy(), g["log_psi_sqr"], rtol=1e-12, atol=1e-12)
    np.testing.assert_array_equal(ph.numpy(), g["phase"])
    e = oh.local_energy(params, shape.n_up, shape.n_dn, r.astype(np.float64), R, Z).numpy()
    np.testing.assert_allclose(e, g["e_loc"], rtol=1e-10, atol=1e-10)


def test_chain_fixture_and_its_coverage():
    g = _load("lih_chain.npz")
    c = chain_case()
    for k in ("r0", "normals", "uniforms", "masks", "walker_age"):
        np.testing.assert_array_equal(c[k], g[k])
    for k in ("r", "log_psi_sqr", "accs", "chain"):
        np.testing.assert_allclose(c[k], g[k], rtol=1e-12, atol=1e-12)
    # the fixture exercises the forced accept (an accept at old age >= 2) and two stepsize updates
    ages = np.zeros(g["masks"].shape[1], np.int64)
    forced = 0
    for s in range(g["masks"].shape[0]):
        forced += int(np.sum(g["masks"][s] & (ages >= 2)))
        ages = np.where(g["masks"][s], 0, ages + 1)
    assert forced >= 1
    assert int(g["chain"][2]) == 6 and abs(g["chain"][0] - 0.3) > 1e-3
