"""The reference-signature drop-ins on the GPU: build_log_psi_squared -> log_psi_sqr(params, n_up, n_dn, r,
R, Z, fixed_params), MetropolisHastingsMonteCarlo(config).run_mcmc_steps(func, state, params, n_up, n_dn,
fixed_params, n_steps) and build_local_energy(...)(params, spin_state, r, R, Z, fixed_params)
(model/wavefunction.py:262-299, mcmc.py:322-327, hamiltonian.py:272-291), checked against the FP64 oracle
and against the C ABI called directly; parameter updates reach the device."""
import numpy as np
import pytest
import torch

from oracle import hamiltonian as oh
from oracle import wavefunction as owf

pytestmark = pytest.mark.gpu


def _setup():
    from deeperwin_amd.config import ModelConfigFermiNet, OrbitalsConfigFermiNet, PhysicalConfig
    from deeperwin_amd.physics import geometry_lih

    R, Z = geometry_lih()
    phys = PhysicalConfig(R=np.asarray(R, np.float32).astype(np.float64).tolist(), Z=list(map(int, Z)))
    cfg = ModelConfigFermiNet(orbitals=OrbitalsConfigFermiNet(n_determinants=4))
    return phys, cfg


def test_dropin_chain_and_energy_match_oracle(cuda):
    from deeperwin_amd.config import MCMCConfig
    from deeperwin_amd.hamiltonian import build_local_energy
    from deeperwin_amd.mcmc import MCMCState, MetropolisHastingsMonteCarlo, read_chain
    from deeperwin_amd.wavefunction import build_log_psi_squared

    phys, cfg = _setup()
    log_psi_sqr, _, _, params, fixed = build_log_psi_squared(cfg, phys, rng_seed=0)
    n_up, n_dn = phys.n_up, phys.n_dn
    B, S = 64, 6
    state = MCMCState.initialize_around_nuclei(B, phys, rng=3, stepsize=0.02)
    R = np.asarray(phys.R)
    # log|psi|^2 through the reference signature, any leading batch shape
    r0 = state.r.cpu().numpy().astype(np.float64)
    phase, lp = log_psi_sqr(params, n_up, n_dn, state.r.reshape(8, 8, n_up + n_dn, 3), state.R, state.Z, fixed)
    ph_ref, lp_ref = owf.log_psi_sqr(params, n_up, n_dn, torch.tensor(r0), R)
    assert lp.shape == (8, 8)
    kappa = owf.conditioning(params, n_up, n_dn, r0, R)
    tol = 5e-5 * np.maximum(1, np.abs(lp_ref.numpy()) / 10) + 4 * 2.0 ** -24 * kappa
    assert np.all(np.abs(lp.reshape(-1).cpu().numpy() - lp_ref.numpy()) <= tol)
    np.testing.assert_array_equal(phase.reshape(-1).cpu().numpy(), ph_ref.numpy().astype(np.float32))
    # the sampler through the reference signature equals the C ABI called directly on the same draws
    rng = np.random.default_rng(9)
    normals = torch.tensor(rng.standard_normal((S, B, n_up + n_dn, 3)), dtype=torch.float32, device="cuda")
    uniforms = torch.tensor(rng.uniform(size=(S, B)), dtype=torch.float32, device="cuda")
    mc = MetropolisHastingsMonteCarlo(MCMCConfig(max_age=2, stepsize_update_interval=3))
    new = mc.run_mcmc_steps(log_psi_sqr, state, params, n_up, n_dn, fixed, S, draws=(normals, uniforms))
    model = log_psi_sqr.model_for(params, n_up, n_dn)
    from deeperwin_amd.mcmc import make_chain

    r = state.r.clone()
    lp2 = torch.full((B,), -1000.0, device="cuda")
    age = torch.zeros(B, dtype=torch.int32, device="cuda")
    chain = make_chain(0.02)
    model.mcmc_steps(r, state.R, state.Z, lp2, age, normals, uniforms, chain, 2, 3, 0.5, 1e-2, 1.0)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(new.r.cpu().numpy(), r.cpu().numpy())
    np.testing.assert_array_equal(new.walker_age.cpu().numpy(), age.cpu().numpy())
    assert (new.stepsize, new.acc_rate, new.step_nr) == read_chain(chain)
    assert new.step_nr == S
    assert state.r.data_ptr() != new.r.data_ptr()  # a new state; the input state is unchanged
    np.testing.assert_array_equal(state.r.cpu().numpy().astype(np.float64), r0)
    # the local energy through the reference signature
    get_e = build_local_energy(log_psi_sqr)
    e = get_e(params, (n_up, n_dn), new.r, new.R, new.Z, fixed)
    rs = new.r.cpu().numpy().astype(np.float64)[:16]
    e_ref = oh.local_energy(params, n_up, n_dn, rs, R, np.asarray(phys.Z)).numpy()
    kap = owf.conditioning(params, n_up, n_dn, rs, R)
    assert np.all(np.abs(e.cpu().numpy()[:16] - e_ref) <= (1 + np.abs(e_ref)) * (1e-4 + 4 * 2.0 ** -24 * kap))


def test_parameter_updates_reach_the_device(cuda):
    from deeperwin_amd.wavefunction import build_log_psi_squared

    phys, cfg = _setup()
    log_psi_sqr, _, _, params, fixed = build_log_psi_squared(cfg, phys, rng_seed=0)
    n_up, n_dn = phys.n_up, phys.n_dn
    r = torch.tensor(np.random.default_rng(0).standard_normal((16, 4, 3)), dtype=torch.float32, device="cuda")
    R = torch.tensor(phys.R, dtype=torch.float32, device="cuda")
    Z = torch.tensor(phys.Z, dtype=torch.float32, device="cuda")
    lp0 = log_psi_sqr(params, n_up, n_dn, r, R, Z, fixed)[1].clone()
    # in-place update of a numpy leaf: same object, new content
    leaf = params["wf/orbitals/envelope_orbitals"]["alpha_up"]
    leaf += 0.5
    lp1 = log_psi_sqr(params, n_up, n_dn, r, R, Z, fixed)[1].clone()
    ref1 = owf.log_psi_sqr(params, n_up, n_dn, torch.tensor(r.cpu().numpy(), dtype=torch.float64),
                           np.asarray(phys.R))[1].numpy()
    assert not torch.equal(lp0, lp1)
    np.testing.assert_allclose(lp1.cpu().numpy(), ref1, rtol=1e-4, atol=1e-3)
    # a fresh dict of torch tensors, then an in-place update of one of them
    tparams = {p: {k: torch.tensor(v) for k, v in leaves.items()} for p, leaves in params.items()}
    lp2 = log_psi_sqr(tparams, n_up, n_dn, r, R, Z, fixed)[1].clone()
    assert torch.equal(lp1, lp2)
    with torch.no_grad():
        tparams["wf/orbitals/envelope_orbitals"]["alpha_up"].sub_(0.5)
    lp3 = log_psi_sqr(tparams, n_up, n_dn, r, R, Z, fixed)[1].clone()
    assert torch.equal(lp3, lp0)
