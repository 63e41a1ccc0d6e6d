"""Write the regression fixtures in this folder from the float64 CPU oracle.

The reference's own tests hold no vectors for this path and the reference cannot run here (JAX, haiku and
folx are absent), so these fixtures are the oracle's outputs, kept to detect drift of the oracle and of
the seeded inputs across changes.  They are data only: inputs (seeds, walkers, draws) and outputs.

    python tests/golden/make_golden.py          # rewrites *.npz next to this file
"""
from __future__ import annotations

import os
import sys

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from workloads import system, walkers  # noqa: E402
from oracle import hamiltonian as oh  # noqa: E402
from oracle import mcmc as om  # noqa: E402
from oracle import wavefunction as owf  # noqa: E402


def param_digest(params) -> np.ndarray:
    """Per-leaf (sum, sum of squares) in sorted path order: catches any change of the seeded parameters."""
    out = []
    for path in sorted(params):
        for leaf in sorted(params[path]):
            a = np.asarray(params[path][leaf], np.float64)
            out.append([a.sum(), (a * a).sum()])
    return np.asarray(out)


def wavefunction_case(name: str, B: int, seed: int) -> dict:
    shape, params, R, Z, mapping = system(name)
    r = walkers(R, mapping, B, seed=seed)
    ph, lp = owf.log_psi_sqr(params, shape.n_up, shape.n_dn, torch.tensor(r, dtype=torch.float64), R)
    e = oh.local_energy(params, shape.n_up, shape.n_dn, r.astype(np.float64), R, Z).numpy()
    return dict(r=r, R=R, Z=Z, mapping=np.asarray(mapping), param_digest=param_digest(params),
                log_psi_sqr=lp.numpy(), phase=ph.numpy(), e_loc=e)


def chain_case() -> dict:
    """LiH, 8 walkers, 6 steps with max_age 2 and a stepsize update every 3 steps."""
    shape, params, R, Z, mapping = system("lih")
    B, S = 8, 6
    r0 = walkers(R, mapping, B, seed=7)
    rng = np.random.default_rng(8)
    normals = rng.standard_normal((S, B, shape.n_el, 3)).astype(np.float32)
    uniforms = rng.uniform(size=(S, B)).astype(np.float32)

    def f(x):
        return owf.log_psi_sqr(params, shape.n_up, shape.n_dn, torch.tensor(x, dtype=torch.float64), R)[1].numpy()

    st = om.ChainState(r=r0, log_psi_sqr=np.full(B, -1000.0), walker_age=np.zeros(B, np.int64), stepsize=0.3)
    out, accs, masks = om.mcmc_steps(f, st, normals, uniforms, max_age=2, update_interval=3, dtype=np.float64)
    return dict(r0=r0, normals=normals, uniforms=uniforms, r=out.r, log_psi_sqr=out.log_psi_sqr,
                walker_age=out.walker_age, accs=accs, masks=masks,
                chain=np.array([out.stepsize, out.acc_rate, out.step_nr]))


def main() -> None:
    np.savez_compressed(os.path.join(HERE, "lih_wavefunction.npz"), **wavefunction_case("lih", 8, 11))
    np.savez_compressed(os.path.join(HERE, "n2_wavefunction.npz"), **wavefunction_case("n2", 3, 12))
    np.savez_compressed(os.path.join(HERE, "lih_chain.npz"), **chain_case())
    print("wrote", sorted(f for f in os.listdir(HERE) if f.endswith(".npz")))


if __name__ == "__main__":
    main()
