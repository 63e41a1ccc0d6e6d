"""CPU oracle for the walker-parallel VMC inner loop — TEST INFRASTRUCTURE ONLY.

This package restates, function by function, the reference's MCMC step, log|psi| forward and local
energy on the CPU (PyTorch CPU tensors, float64 by default; the reference's JAX/haiku/folx stack is not
installable here).  Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker or as the timed CPU baseline.  The product path
(``deeperwin_amd``) never imports it and fails loudly when its HIP library is missing.

Pinning status of each function (what independent check fixes it) is listed in ``PINNING`` below.
"""

PINNING = {
    "wavefunction.el_ion_features": "brute-force Python loops (tests/test_oracle_features.py)",
    "wavefunction.el_el_features": "brute-force Python loops; zero diagonal with zero gradient",
    "wavefunction.ferminet_embedding": "closed-form constant embedding (zero weights); permutation "
                                       "equivariance; brute-force per-electron loop",
    "wavefunction.envelope_orbitals": "brute-force per-element loop; closed-form He-like case",
    "wavefunction.sum_of_determinants": "Leibniz permutation expansion for n<=4; explicit log-sum-exp",
    "wavefunction.log_psi_sqr": "antisymmetry under same-spin swaps; joint translation invariance; "
                                "closed-form He-like psi",
    "hamiltonian.potential_energy": "brute-force pair loops; analytic H2 / N2 values",
    "hamiltonian.kinetic_energy": "closed-form He-like E_loc (-k^2 + (k-Z)(1/r1+1/r2) + 1/r12); "
                                  "central finite differences; reverse-over-reverse Hessian trace",
    "mcmc.mcmc_steps": "closed form with supplied draws; stationary-distribution test on a 1s density",
    "evaluation.mean_and_var": "math.fsum",
}
