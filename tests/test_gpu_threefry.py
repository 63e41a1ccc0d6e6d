"""JAX threefry streams on the device (dw_threefry_*) against the oracle restatement, which
tests/test_oracle_threefry.py pins to the Random123 vectors and JAX's documented values.  Keys and uniforms
must be bit-identical; normals agree to the float32 rounding of erfinv's log1p and polynomial (4e-6
relative + 1e-7).  The sampler fed by the device streams replays the oracle chain driven by the oracle's
streams."""
import numpy as np
import pytest
import torch

from oracle import threefry as tf

pytestmark = pytest.mark.gpu


def _close(a, b):
    np.testing.assert_allclose(a, b, rtol=4e-6, atol=1e-7)


def test_split_and_normal_match_oracle(cuda):
    from deeperwin_amd import threefry as dt

    for seed in (0, 42, 2**33 + 5):
        k = dt.PRNGKey(seed)
        np.testing.assert_array_equal(dt.uint32(k), tf.prng_key(seed))
        for num in (2, 3, 1000):
            np.testing.assert_array_equal(dt.uint32(dt.split(k, num)), tf.split(tf.prng_key(seed), num))
        for shape in ((10,), (7, 3), (2048, 14, 3)):
            _close(dt.normal(k, shape).cpu().numpy(), tf.normal(tf.prng_key(seed), shape))
    np.testing.assert_allclose(dt.normal(dt.PRNGKey(0), (10,)).cpu().numpy(),
                               [-0.3721109, 0.26423115, -0.18252768, -0.7368197, -0.44030377, -0.1521442,
                                -0.67135346, -0.5908641, 0.73168886, 0.5673026], rtol=0, atol=5e-8)


@pytest.mark.parametrize("n_el", [3, 14])
def test_mcmc_draws_match_oracle(cuda, n_el):
    from deeperwin_amd import threefry as dt

    keys = tf.split(tf.prng_key(11), 64)
    nd, ud, kd = dt.mcmc_draws(torch.tensor(keys.view(np.int32), device="cuda"), n_el, 5)
    no, uo, ko = tf.mcmc_draws(keys, n_el, 5)
    np.testing.assert_array_equal(dt.uint32(kd), ko)
    np.testing.assert_array_equal(ud.cpu().numpy(), uo)
    _close(nd.cpu().numpy(), no)


def test_sampler_on_reference_streams(cuda):
    """initialize_around_nuclei from a PRNG key and run_mcmc_steps on the walkers' threefry keys: the walker
    positions, keys and chain equal the oracle's replay of the same streams."""
    from deeperwin_amd import threefry as dt
    from deeperwin_amd.config import MCMCConfig, ModelConfigFermiNet, OrbitalsConfigFermiNet, PhysicalConfig
    from deeperwin_amd.mcmc import MCMCState, MetropolisHastingsMonteCarlo
    from deeperwin_amd.physics import geometry_lih
    from deeperwin_amd.wavefunction import build_log_psi_squared
    from oracle import mcmc as om
    from oracle import wavefunction as owf

    Rg, Zg = geometry_lih()
    phys = PhysicalConfig(R=np.asarray(Rg, np.float32).astype(np.float64).tolist(), Z=list(map(int, Zg)))
    cfg = ModelConfigFermiNet(orbitals=OrbitalsConfigFermiNet(n_determinants=2))
    log_psi_sqr, _, _, params, fixed = build_log_psi_squared(cfg, phys, rng_seed=0, precision="fp64")
    B, S = 64, 6
    state = MCMCState.initialize_around_nuclei(B, phys, rng=dt.PRNGKey(5), stepsize=0.02)
    r0, keys0 = tf.init_walkers(tf.prng_key(5), B, phys.R, phys.el_ion_mapping)
    np.testing.assert_array_equal(dt.uint32(state.rng_state), keys0)
    _close(state.r.cpu().numpy(), r0)
    mc = MetropolisHastingsMonteCarlo(MCMCConfig(max_age=3, stepsize_update_interval=100))
    new = mc.run_mcmc_steps(log_psi_sqr, state, params, phys.n_up, phys.n_dn, fixed, S)
    normals, uniforms, keys = tf.mcmc_draws(keys0, phys.n_electrons, S)
    np.testing.assert_array_equal(dt.uint32(new.rng_state), keys)
    R = np.asarray(phys.R)

    def f(x):
        return owf.log_psi_sqr(params, phys.n_up, phys.n_dn, torch.tensor(np.asarray(x, np.float64)), R)[1].numpy()

    st = om.ChainState(r=state.r.cpu().numpy(), log_psi_sqr=np.full(B, -1000.0), walker_age=np.zeros(B, np.int64),
                       stepsize=0.02)
    out, _, _ = om.mcmc_steps(f, st, normals, uniforms, max_age=3, update_interval=100, dtype=np.float64)
    same = np.all(np.abs(new.r.cpu().numpy() - out.r) < 1e-5, axis=(1, 2))
    assert same.mean() >= 0.95, same.mean()
