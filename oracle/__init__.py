"""CPU oracle for the walker-parallel VMC inner loop — TEST INFRASTRUCTURE ONLY.

This package restates, function by function, the reference's MCMC step, log|psi| forward and local
energy on the CPU (PyTorch CPU tensors, float64 by default; the reference's JAX/haiku/folx stack is not
installable here).  Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU legs of ``bench.py``
(``cpu_baseline`` and the accuracy sample) may import it, and only as the checker or as the timed CPU
baseline.  The product path (``deeperwin_amd``) never imports it and fails loudly when its HIP library is
missing.

``PINNING`` lists, for each function, the tests that fix it independently of its own tensor code (files
under ``tests/``).  ``tests/golden`` additionally holds the oracle's outputs on seeded inputs as regression
fixtures (``test_golden.py``); those detect drift but pin nothing by themselves.
"""

PINNING = {
    "wavefunction.el_ion_features": [
        "test_oracle_basics.py::test_el_ion_features_brute_force (per-element loops)",
        "test_oracle_features.py::test_el_ion_rows_are_ion_major (feature f of ion I at column 4 I + f)",
    ],
    "wavefunction.el_el_features": [
        "test_oracle_basics.py::test_el_el_features_brute_force_and_zero_diagonal (r_j - r_i, zero diagonal "
        "with zero gradient)",
    ],
    "wavefunction.symmetric_features": [
        "test_oracle_features.py::test_symmetric_features_brute_force (block order, self pair included)",
        "test_oracle_features.py::test_planted_layout_mistakes_are_caught (swapped spin-mean / pair-mean blocks)",
    ],
    "wavefunction.ferminet_embedding": [
        "test_oracle_features.py::test_slater_matrices_element_by_element (scalar restatement, tests/bruteforce.py)",
        "test_oracle_physics.py::test_embedding_is_permutation_equivariant_within_spin",
    ],
    "wavefunction.envelope_orbitals": [
        "test_oracle_features.py::test_slater_matrices_element_by_element (n_det >= 2, n_ion >= 2)",
        "test_oracle_features.py::test_two_det_two_centre_closed_form (column d * n + k = det d, orbital k)",
        "test_oracle_features.py::test_planted_layout_mistakes_are_caught ((orbital, det) column order)",
        "test_oracle_physics.py::test_he_like_closed_form_psi_and_local_energy",
    ],
    "wavefunction.sum_of_determinants": [
        "test_oracle_basics.py::test_sum_of_determinants_leibniz",
        "test_oracle_basics.py::test_log_epsilon_is_inside_the_shifted_log",
    ],
    "wavefunction.log_psi_sqr": [
        "test_oracle_features.py::test_log_psi_sqr_brute_force",
        "test_oracle_features.py::test_two_det_two_centre_closed_form",
        "test_oracle_physics.py::test_same_spin_swap_flips_sign_keeps_density",
        "test_oracle_physics.py::test_joint_translation_invariance",
        "test_oracle_physics.py::test_he_like_closed_form_psi_and_local_energy",
    ],
    "wavefunction.slater_matrices": [
        "test_oracle_features.py::test_slater_matrices_element_by_element",
    ],
    "wavefunction.conditioning": [
        "not a reference function: a float64 sensitivity estimate used only to size test tolerances",
    ],
    "hamiltonian.potential_energy": [
        "test_oracle_basics.py::test_potential_energy_brute_force_and_h2 (pair loops, analytic H2)",
        "test_oracle_features.py::test_n2_ion_ion_repulsion (49 / 2.07)",
    ],
    "hamiltonian.kinetic_energy": [
        "test_oracle_physics.py::test_he_like_closed_form_psi_and_local_energy",
        "test_oracle_features.py::test_two_det_two_centre_closed_form (two dets, two centres)",
        "test_oracle_physics.py::test_laplacian_central_finite_differences",
        "test_oracle_physics.py::test_laplacian_matches_reverse_over_reverse_hessian",
    ],
    "hamiltonian.local_energy": [
        "test_oracle_physics.py::test_local_energy_chunking_is_invisible",
        "test_oracle_features.py::test_two_det_two_centre_closed_form",
    ],
    "hamiltonian.mean_and_var": [
        "test_oracle_physics.py::test_mean_and_var_fsum (math.fsum, NaN dropped, inf kept)",
    ],
    "wavefunction.dpe4_embedding": [
        "test_oracle_dpe4.py::test_dpe4_matches_scalar_restatement (scalar loops over the reference's same/diff "
        "lists and block reshapes, Leibniz determinants)",
        "test_oracle_dpe4.py::test_planted_mistakes_fail (same/diff swapped, no Z_min offset, self pair dropped)",
    ],
    "wavefunction.ion_embedding": [
        "test_oracle_dpe4.py::test_dpe4_matches_scalar_restatement",
    ],
    "hamiltonian.kinetic_energy (dpe4)": [
        "test_oracle_dpe4.py::test_dpe4_kinetic_energy_finite_differences",
    ],
    "mcmc.mcmc_steps": [
        "test_oracle_mcmc.py::test_hand_worked_chain (forced accept at the old age, old-rate stepsize update)",
        "test_oracle_mcmc.py::test_stationary_distribution_hydrogen_1s",
    ],
    "mcmc.adjust_stepsize": [
        "test_oracle_mcmc.py::test_stepsize_clip_and_increase",
    ],
    "mcmc.propose": [
        "test_oracle_proposals.py::test_log_q_is_reverse_over_forward_density (scipy.stats.norm densities)",
        "test_oracle_proposals.py::test_stationary_distribution_hydrogen_1s (local and Langevin chains; fails "
        "without log q)",
        "test_oracle_proposals.py::test_constant_scale_without_drift_is_symmetric",
        "test_oracle_proposals.py::test_one_electron_proposals (one-electron forms, scipy densities)",
        "test_oracle_proposals.py::test_one_electron_chain_samples_two_independent_electrons",
    ],
    "mcmc.log_q_ratio": [
        "test_oracle_proposals.py::test_log_q_is_reverse_over_forward_density (through propose)",
        "test_oracle_proposals.py::test_stationary_distribution_hydrogen_1s",
    ],
    "mcmc.scale_and_drift": [
        "test_oracle_proposals.py::test_drift_is_minus_gradient_of_nuclear_distance_sum (central differences)",
        "test_oracle_proposals.py::test_planted_mistakes_fail_the_density_check",
    ],
    "hamiltonian.kinetic_energy_vmapped": [
        "test_oracle_physics.py::test_vmapped_kinetic_energy_equals_walker_loop (against the pinned walker loop)",
    ],
    "threefry.threefry2x32": [
        "test_oracle_threefry.py::test_threefry_random123_vectors (Random123 known-answer vectors)",
    ],
    "threefry.split": [
        "test_oracle_threefry.py::test_jax_documented_values (split(PRNGKey(0)) as JAX's documentation prints it)",
    ],
    "threefry.random_bits": [
        "test_oracle_threefry.py::test_jax_documented_values (normal(PRNGKey(0), (10,)), normal(PRNGKey(42)))",
    ],
    "threefry.uniform": [
        "test_oracle_threefry.py::test_erfinv32_and_uniform_fill (mantissa fill of hand-picked words)",
        "test_oracle_threefry.py::test_jax_documented_values",
    ],
    "threefry.erfinv32": [
        "test_oracle_threefry.py::test_erfinv32_and_uniform_fill (scipy's float64 erfinv)",
    ],
    "threefry.normal": [
        "test_oracle_threefry.py::test_jax_documented_values",
    ],
    "threefry.mcmc_draws": [
        "test_oracle_threefry.py::test_mcmc_stream_shares_one_subkey_per_step (key chain and shared subkey)",
    ],
    "threefry.init_walkers": [
        "test_oracle_threefry.py::test_jax_documented_values (its split and normal; the composition follows "
        "mcmc.py:42-66, 90)",
    ],
    "loss.clipping_center_and_width": [
        "test_oracle_loss.py::test_clipping_center_and_width_known_answers (hand-worked mean, median, std, mae)",
    ],
    "loss.clip": [
        "test_oracle_loss.py::test_clip_functions_known_answers (hard and tanh clips, NaN kept)",
    ],
    "loss.clip_energies": [
        "test_oracle_loss.py::test_clip_functions_known_answers (initial state (0, 1e12) leaves E unchanged)",
    ],
    "loss.total_energy": [
        "test_oracle_loss.py::test_clip_functions_known_answers (statistics of the clipped energies)",
    ],
    "loss.param_grad": [
        "test_oracle_loss.py::test_param_grad_matches_finite_differences (every leaf kind, central differences)",
    ],
    "loss.loss_gradient": [
        "test_oracle_loss.py::test_clip_functions_known_answers (its statistics: loss.total_energy)",
        "test_oracle_loss.py::test_param_grad_matches_finite_differences (its gradient: loss.param_grad)",
    ],
}
