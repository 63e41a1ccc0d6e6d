"""The mathematics of the native kernel design against the oracle (CPU only).

``kernel_model.run`` restates the HIP pipeline's algorithm (channel rows, spin-mean split of the symmetric
features, pair-mean tangent rows, tanh rule, envelope product rule, determinant traces, epsilon-literal
combination) in float64 NumPy; agreement with the oracle's autodiff Laplacian to ~1e-10 shows the design
itself is exact, so GPU differences are rounding only.
"""
import numpy as np
import pytest
import torch

import kernel_model as km
from oracle import hamiltonian as oh
from oracle import wavefunction as owf
from workloads import system, walkers


@pytest.mark.parametrize("name,B", [("lih", 4), ("n2", 2), ("he", 3)])
def test_kernel_design_matches_oracle(name, B):
    shape, params, R, Z, mapping = system(name)
    r = walkers(R, mapping, B, seed=11).astype(np.float64)
    out = km.run(params, shape.n_up, r, R, Z, shape.n_det)
    lp = owf.log_psi_sqr(params, shape.n_up, shape.n_dn, torch.tensor(r), R)[1].numpy()
    ek = oh.kinetic_energy(params, shape.n_up, shape.n_dn, torch.tensor(r), torch.tensor(R)).numpy()
    np.testing.assert_allclose(out["log_psi_sqr"], lp, rtol=0, atol=1e-9)
    np.testing.assert_allclose(out["ekin"], ek, rtol=1e-8, atol=1e-8)
