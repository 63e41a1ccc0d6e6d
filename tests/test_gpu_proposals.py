"""The local-stepsize and approximate-Langevin proposals on the device (mcmc.py:218-236, 255-284), through the
C ABI:

* dw_mcmc_propose against the FP64 oracle: r' = r + s xi (+ g s^2) within 1e-6 (1 + |r| + |s xi| +
  |g s^2|) of the FP64 move of the same float32 inputs, and log q within 1e-5 (1 + sum_i T_i) of the FP64
  log q of the device's own move r -> r' (T_i: the magnitudes of electron i's four terms), on N2, the
  26-electron C4H2-like molecule and a 60-electron molecule, at stepsizes 0.02 and 0.3;
* MCMC decisions of the FP64 form in lockstep with the FP64 oracle on N2 (max_age 2, stepsize interval 3):
  p = exp(lp' - lp + log q) with the oracle's lp, lp' and log q; decisions must agree except within the
  margin of the float32 lp values and of the log q bound above; accepted walkers hold exactly the
  dw_mcmc_propose move, rejected ones their old position;
* the drop-in sampler with MCMCLangevinProposalConfig / LocalStepsizeProposalConfig equals the C ABI called
  directly, and dw_mcmc_steps_noalloc equals dw_mcmc_steps bit for bit with a proposal selected.
"""
import numpy as np
import pytest
import torch

from oracle import mcmc as om
from oracle import wavefunction as owf
from workloads import system, walkers

pytestmark = pytest.mark.gpu

U32 = 2.0 ** -24
KW = {"local": dict(r_min=0.1, r_max=1.0, langevin_scale=1.0),
      "langevin": dict(r_min=0.2, r_max=2.0, langevin_scale=1.0)}


def _dev(x, dtype=torch.float32):
    return torch.as_tensor(np.asarray(x)).to("cuda", dtype).contiguous()


def _model(shape, params, precision, kind):
    from deeperwin_amd._native import NativeModel

    m = NativeModel(shape, precision)
    m.set_params(params)
    m.set_proposal(kind, **KW[kind])
    return m


def _terms(kind, r, r_new, R, Z, stepsize):
    """Sum over electrons of the magnitudes of the four log q terms (FP64), per walker."""
    kw = KW[kind]
    scale = kw["langevin_scale"] if kind == "langevin" else 0.0
    s, g = om.scale_and_drift(r, R, Z, stepsize, kw["r_min"], kw["r_max"], scale, np.float64)
    s1, g1 = om.scale_and_drift(r_new, R, Z, stepsize, kw["r_min"], kw["r_max"], scale, np.float64)
    d_fwd = np.sum((r_new - r - g * s[..., None] ** 2) ** 2, axis=-1)
    d_rev = np.sum((r - r_new - g1 * s1[..., None] ** 2) ** 2, axis=-1)
    t = 3 * np.abs(np.log(s)) + 3 * np.abs(np.log(s1)) + 0.5 * d_fwd / s ** 2 + 0.5 * d_rev / s1 ** 2
    return t.sum(axis=-1)


def _logq_tol(kind, r, r_new, R, Z, stepsize):
    return 1e-5 * (1.0 + _terms(kind, r, r_new, R, Z, stepsize))


@pytest.mark.parametrize("name", ["n2", "c4h2", "big60"])
@pytest.mark.parametrize("kind", ["local", "langevin"])
@pytest.mark.parametrize("stepsize", [0.02, 0.3])
def test_propose_matches_oracle(cuda, name, kind, stepsize):
    from deeperwin_amd.mcmc import make_chain

    shape, params, R, Z, mapping = system(name, n_det=2, seed=1)
    B = 512
    print(f"propose {kind} on {name} (n = {shape.n_el}), stepsize {stepsize}", flush=True)
    m = _model(shape, params, "fp32", kind)
    r = walkers(R, mapping, B, seed=5)
    xi = np.random.default_rng(6).standard_normal((B, shape.n_el, 3)).astype(np.float32)
    out_r = torch.empty((B, shape.n_el, 3), device="cuda")
    out_q = torch.full((B,), float("nan"), device="cuda")
    m.propose(_dev(r), _dev(R), _dev(Z), _dev(xi), make_chain(stepsize), out_r, out_q)
    torch.cuda.synchronize()
    r_dev = out_r.cpu().numpy().astype(np.float64)
    q_dev = out_q.cpu().numpy().astype(np.float64)
    step = float(np.float32(stepsize))
    r64 = r.astype(np.float64)
    r_ref, _ = om.propose(kind, r64, R, Z, xi.astype(np.float64), step, dtype=np.float64, **KW[kind])
    # float32 rounding of the sum r + s xi + g s^2: bound by the magnitudes of its parts
    kw = KW[kind]
    s, g = om.scale_and_drift(r64, R, Z, step, kw["r_min"], kw["r_max"],
                              kw["langevin_scale"] if kind == "langevin" else 0.0, np.float64)
    err_r = np.abs(r_dev - r_ref)
    assert np.all(err_r <= 1e-6 * (1 + np.abs(r64) + np.abs(xi) * s[..., None] + np.abs(g) * s[..., None] ** 2)), \
        err_r.max()
    q_ref = om.log_q_ratio(kind, r64, r_dev, R, Z, step, dtype=np.float64, **KW[kind])
    tol = _logq_tol(kind, r64, r_dev, R, Z, step)
    err_q = np.abs(q_dev - q_ref)
    print(f"  max |dr| {err_r.max():.2e}, max |dlog q| {err_q.max():.2e} (max tol {tol.max():.2e}, "
          f"|log q| up to {np.abs(q_ref).max():.3f})", flush=True)
    assert np.all(err_q <= tol), (np.argmax(err_q - tol), err_q.max())
    assert np.abs(q_ref).max() > 1e-3


def test_normal_proposal_log_q_is_zero(cuda):
    from deeperwin_amd.mcmc import make_chain

    shape, params, R, Z, mapping = system("lih", seed=1)
    B = 64
    from deeperwin_amd._native import NativeModel

    m = NativeModel(shape, "fp32")
    m.set_params(params)
    r = walkers(R, mapping, B, seed=1)
    xi = np.random.default_rng(2).standard_normal((B, shape.n_el, 3)).astype(np.float32)
    out_r = torch.empty((B, shape.n_el, 3), device="cuda")
    out_q = torch.full((B,), float("nan"), device="cuda")
    m.propose(_dev(r), _dev(R), _dev(Z), _dev(xi), make_chain(0.25), out_r, out_q)
    torch.cuda.synchronize()
    assert np.all(out_q.cpu().numpy() == 0)
    np.testing.assert_allclose(out_r.cpu().numpy(), r + xi * np.float32(0.25), rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("kind", ["local", "langevin"])
def test_mcmc_decisions_lockstep_n2(cuda, kind):
    from deeperwin_amd.mcmc import make_chain, read_chain

    shape, params, R, Z, mapping = system("n2", seed=1)
    B, S, max_age, interval = 256, 6, 2, 3
    rng = np.random.default_rng(31)
    normals = rng.standard_normal((S, B, shape.n_el, 3)).astype(np.float32)
    uniforms = rng.uniform(size=(S, B)).astype(np.float32)
    m = _model(shape, params, "fp64", kind)
    r = _dev(walkers(R, mapping, B, seed=32))
    lp = torch.full((B,), -1000.0, device="cuda")
    age = torch.zeros(B, dtype=torch.int32, device="cuda")
    chain = make_chain(0.2)
    Rd, Zd = _dev(R), _dev(Z)
    n_differ, n_forced = 0, 0
    for k in range(S):
        r_before = r.cpu().numpy().astype(np.float64)
        age_before = age.cpu().numpy().astype(np.int64)
        step = read_chain(chain)[0]
        prop_r = torch.empty_like(r)
        prop_q = torch.empty(B, device="cuda")
        m.propose(r, Rd, Zd, _dev(normals[k]), chain, prop_r, prop_q)
        m.mcmc_steps(r, Rd, Zd, lp, age, _dev(normals[k:k + 1]), _dev(uniforms[k:k + 1]), chain, max_age, interval,
                     0.5, 1e-2, 1.0)
        torch.cuda.synchronize()
        r_prop = prop_r.cpu().numpy()
        dev_accept = age.cpu().numpy() == 0
        r_after = r.cpu().numpy()
        np.testing.assert_array_equal(r_after[dev_accept], r_prop[dev_accept])
        np.testing.assert_array_equal(r_after[~dev_accept], r_before[~dev_accept].astype(np.float32))
        rp = r_prop.astype(np.float64)
        lp0 = owf.log_psi_sqr(params, shape.n_up, shape.n_dn, torch.tensor(r_before), R)[1].numpy()
        lp1 = owf.log_psi_sqr(params, shape.n_up, shape.n_dn, torch.tensor(rp), R)[1].numpy()
        lq = om.log_q_ratio(kind, r_before, rp, R, Z, step, dtype=np.float64, **KW[kind])
        with np.errstate(over="ignore"):
            p = np.exp(lp1 - lp0 + lq)
        forced = age_before >= max_age
        ref_accept = (p > uniforms[k]) | forced
        margin = 1e-6 + np.minimum(p, 1e30) * ((np.abs(lp0) + np.abs(lp1)) * U32 +
                                               _logq_tol(kind, r_before, rp, R, Z, step))
        near = (np.abs(p - uniforms[k]) <= margin) & ~forced
        bad = (dev_accept != ref_accept) & ~near
        differ = dev_accept != ref_accept
        print(f"{kind} step {k}: stepsize {step:.4f}, accepted {dev_accept.sum()}/{B}, forced {forced.sum()}, "
              f"|log q| up to {np.abs(lq).max():.3f}, differing {differ.sum()} (outside margin {bad.sum()})",
              flush=True)
        assert not bad.any(), np.nonzero(bad)[0][:8]
        n_differ += int(differ.sum())
        n_forced += int(forced.sum())
    assert n_forced > 0
    assert n_differ <= 0.01 * B * S


@pytest.mark.parametrize("kind", ["local", "langevin"])
def test_dropin_and_noalloc_with_proposal(cuda, kind):
    from deeperwin_amd.config import (LocalStepsizeProposalConfig, MCMCConfig, MCMCLangevinProposalConfig,
                                      ModelConfigFermiNet, OrbitalsConfigFermiNet, PhysicalConfig)
    from deeperwin_amd.mcmc import MCMCState, MetropolisHastingsMonteCarlo, make_chain, read_chain
    from deeperwin_amd.physics import geometry_lih
    from deeperwin_amd.wavefunction import build_log_psi_squared

    R, Z = geometry_lih()
    phys = PhysicalConfig(R=np.asarray(R, np.float32).astype(np.float64).tolist(), Z=list(map(int, Z)))
    cfg = ModelConfigFermiNet(orbitals=OrbitalsConfigFermiNet(n_determinants=4))
    log_psi_sqr, _, _, params, fixed = build_log_psi_squared(cfg, phys, rng_seed=0)
    n_up, n_dn = phys.n_up, phys.n_dn
    B, S = 128, 8
    state = MCMCState.initialize_around_nuclei(B, phys, rng=4, stepsize=0.1)
    rng = np.random.default_rng(12)
    normals = _dev(rng.standard_normal((S, B, n_up + n_dn, 3)))
    uniforms = _dev(rng.uniform(size=(S, B)))
    prop = LocalStepsizeProposalConfig() if kind == "local" else MCMCLangevinProposalConfig()
    mc = MetropolisHastingsMonteCarlo(MCMCConfig(max_age=3, stepsize_update_interval=4, proposal=prop))
    new = mc.run_mcmc_steps(log_psi_sqr, state, params, n_up, n_dn, fixed, S, draws=(normals, uniforms))
    model = log_psi_sqr.model_for(params, n_up, n_dn)
    outs = []
    for noalloc in (False, True):
        model.set_proposal(kind, prop.r_min, prop.r_max, getattr(prop, "langevin_scale", 1.0))
        r = state.r.clone()
        lp = torch.full((B,), -1000.0, device="cuda")
        age = torch.zeros(B, dtype=torch.int32, device="cuda")
        chain = make_chain(0.1)
        model.mcmc_steps(r, state.R, state.Z, lp, age, normals, uniforms, chain, 3, 4, 0.5, 1e-2, 1.0,
                         noalloc=noalloc)
        torch.cuda.synchronize()
        outs.append((r.cpu().numpy(), lp.cpu().numpy(), age.cpu().numpy(), read_chain(chain)))
    for o in outs:
        np.testing.assert_array_equal(new.r.cpu().numpy(), o[0])
        np.testing.assert_array_equal(new.log_psi_sqr.cpu().numpy(), o[1])
        np.testing.assert_array_equal(new.walker_age.cpu().numpy(), o[2])
        assert read_chain(new.chain) == o[3]
    # the proposal changed the chain: the Gaussian rule on the same draws ends elsewhere
    model.set_proposal("normal")
    r = state.r.clone()
    model.mcmc_steps(r, state.R, state.Z, torch.full((B,), -1000.0, device="cuda"),
                     torch.zeros(B, dtype=torch.int32, device="cuda"), normals, uniforms, make_chain(0.1), 3, 4,
                     0.5, 1e-2, 1.0)
    assert not np.array_equal(r.cpu().numpy(), outs[0][0])


@pytest.mark.parametrize("name", ["n2", "big60"])
@pytest.mark.parametrize("kind", ["normal_one_el", "local_one_el"])
def test_one_electron_propose_matches_oracle(cuda, name, kind):
    """Only electron step_nr mod n moves (mcmc.py:183-192, 239-252): the other electrons are copied bit for
    bit, the moved one and log q match the FP64 oracle (log q of the device's own move)."""
    from deeperwin_amd.mcmc import make_chain

    shape, params, R, Z, mapping = system(name, n_det=2, seed=1)
    B, step_nr = 256, 17
    k = step_nr % shape.n_el
    m = _model(shape, params, "fp32", "local")
    m.set_proposal(kind, 0.1, 1.0)
    r = walkers(R, mapping, B, seed=7)
    xi = np.random.default_rng(8).standard_normal((B, 3)).astype(np.float32)
    out_r = torch.empty((B, shape.n_el, 3), device="cuda")
    out_q = torch.full((B,), float("nan"), device="cuda")
    m.propose(_dev(r), _dev(R), _dev(Z), _dev(xi), make_chain(0.3, 0.0, step_nr), out_r, out_q)
    torch.cuda.synchronize()
    r_dev = out_r.cpu().numpy()
    np.testing.assert_array_equal(np.delete(r_dev, k, axis=1), np.delete(r, k, axis=1))
    step = float(np.float32(0.3))
    r_ref, _ = om.propose(kind, r.astype(np.float64), R, Z, xi.astype(np.float64), step, 0.1, 1.0,
                          dtype=np.float64, step_nr=step_nr)
    np.testing.assert_allclose(r_dev[:, k], r_ref[:, k], rtol=0, atol=2e-6 * (1 + np.abs(r[:, k]).max()))
    q_dev = out_q.cpu().numpy().astype(np.float64)
    if kind == "normal_one_el":
        assert np.all(q_dev == 0)
    else:
        rd = r_dev.astype(np.float64)
        s0, _ = om.scale_and_drift(r.astype(np.float64)[:, k:k + 1], R, Z, step, 0.1, 1.0, 0.0, np.float64)
        s1, _ = om.scale_and_drift(rd[:, k:k + 1], R, Z, step, 0.1, 1.0, 0.0, np.float64)
        d2 = np.sum((rd[:, k] - r[:, k]) ** 2, axis=-1)
        q_ref = 3 * (np.log(s0[:, 0]) - np.log(s1[:, 0])) + 0.5 * d2 * (1 / s0[:, 0] ** 2 - 1 / s1[:, 0] ** 2)
        tol = 1e-5 * (1 + 3 * np.abs(np.log(s0[:, 0])) + 3 * np.abs(np.log(s1[:, 0])) + 0.5 * d2 / s0[:, 0] ** 2 +
                      0.5 * d2 / s1[:, 0] ** 2)
        assert np.all(np.abs(q_dev - q_ref) <= tol), np.abs(q_dev - q_ref).max()
        assert np.abs(q_ref).max() > 1e-3


def test_one_electron_chain_lockstep_n2(cuda):
    """local_one_el MCMC of the FP64 form in lockstep with the FP64 oracle: step k moves electron k mod n of
    every walker; decisions agree outside the float32 rounding margin of the threshold."""
    from deeperwin_amd.mcmc import make_chain, read_chain

    shape, params, R, Z, mapping = system("n2", seed=1)
    B, S = 128, 6
    rng = np.random.default_rng(41)
    normals = rng.standard_normal((S, B, 3)).astype(np.float32)
    uniforms = rng.uniform(size=(S, B)).astype(np.float32)
    m = _model(shape, params, "fp64", "local")
    m.set_proposal("local_one_el", 0.1, 1.0)
    r = _dev(walkers(R, mapping, B, seed=42))
    lp = torch.full((B,), -1000.0, device="cuda")
    age = torch.zeros(B, dtype=torch.int32, device="cuda")
    chain = make_chain(0.5)
    Rd, Zd = _dev(R), _dev(Z)
    moved_total = 0
    for k in range(S):
        r_before = r.cpu().numpy()
        step, _, step_nr = read_chain(chain)
        assert step_nr == k
        m.mcmc_steps(r, Rd, Zd, lp, age, _dev(normals[k:k + 1]), _dev(uniforms[k:k + 1]), chain, 100, 1000, 0.5,
                     1e-2, 1.0)
        torch.cuda.synchronize()
        r_after = r.cpu().numpy()
        dev_accept = age.cpu().numpy() == 0
        e = k % shape.n_el
        np.testing.assert_array_equal(np.delete(r_after, e, axis=1), np.delete(r_before, e, axis=1))
        rp, lq = om.propose("local_one_el", r_before.astype(np.float64), R, Z, normals[k].astype(np.float64), step,
                            0.1, 1.0, dtype=np.float64, step_nr=k)
        lp0 = owf.log_psi_sqr(params, shape.n_up, shape.n_dn, torch.tensor(r_before.astype(np.float64)), R)[1].numpy()
        lp1 = owf.log_psi_sqr(params, shape.n_up, shape.n_dn, torch.tensor(rp), R)[1].numpy()
        p = np.exp(np.minimum(lp1 - lp0 + lq, 80.0))
        ref_accept = p > uniforms[k]
        margin = 1e-6 + p * ((np.abs(lp0) + np.abs(lp1)) * U32 + 1e-5 * (1 + np.abs(lq)))
        near = np.abs(p - uniforms[k]) <= margin
        assert np.all((dev_accept == ref_accept) | near)
        moved_total += int(dev_accept.sum())
    assert moved_total > 0


def test_cauchy_and_one_electron_dropin(cuda):
    """The drop-in sampler with 'cauchy' (the normal kernel on Cauchy draws) and 'normal_one_el' equals the
    ABI on the same draws; one-electron draws from the walkers' threefry keys are normal(subkey, (3,)) of the
    oracle's key chain, bit for bit on the uniforms."""
    from deeperwin_amd import threefry
    from deeperwin_amd.config import (MCMCConfig, MCMCSimpleProposalConfig, ModelConfigFermiNet,
                                      OrbitalsConfigFermiNet, PhysicalConfig)
    from deeperwin_amd.mcmc import MCMCState, MetropolisHastingsMonteCarlo, make_chain
    from deeperwin_amd.physics import geometry_lih
    from deeperwin_amd.wavefunction import build_log_psi_squared
    from oracle import threefry as otf

    R, Z = geometry_lih()
    phys = PhysicalConfig(R=np.asarray(R, np.float32).astype(np.float64).tolist(), Z=list(map(int, Z)))
    log_psi_sqr, _, _, params, fixed = build_log_psi_squared(
        ModelConfigFermiNet(orbitals=OrbitalsConfigFermiNet(n_determinants=4)), phys, rng_seed=0)
    n_up, n_dn = phys.n_up, phys.n_dn
    n = n_up + n_dn
    B, S = 64, 6
    rng = np.random.default_rng(13)
    for name, shape in (("cauchy", (S, B, n, 3)), ("normal_one_el", (S, B, 3))):
        state = MCMCState.initialize_around_nuclei(B, phys, rng=5, stepsize=0.05)
        draws = (_dev(rng.standard_cauchy(shape) if name == "cauchy" else rng.standard_normal(shape)),
                 _dev(rng.uniform(size=(S, B))))
        mc = MetropolisHastingsMonteCarlo(MCMCConfig(max_age=3, stepsize_update_interval=4,
                                                     proposal=MCMCSimpleProposalConfig(name)))
        new = mc.run_mcmc_steps(log_psi_sqr, state, params, n_up, n_dn, fixed, S, draws=draws)
        model = log_psi_sqr.model_for(params, n_up, n_dn)
        model.set_proposal(name)
        r = state.r.clone()
        lp = torch.full((B,), -1000.0, device="cuda")
        age = torch.zeros(B, dtype=torch.int32, device="cuda")
        model.mcmc_steps(r, state.R, state.Z, lp, age, draws[0], draws[1], make_chain(0.05), 3, 4, 0.5, 1e-2, 1.0)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(new.r.cpu().numpy(), r.cpu().numpy())
        np.testing.assert_array_equal(new.walker_age.cpu().numpy(), age.cpu().numpy())
    # threefry one-electron draws
    key = threefry.PRNGKey(11)
    state = MCMCState.initialize_around_nuclei(B, phys, rng=key, stepsize=0.05)
    mc = MetropolisHastingsMonteCarlo(MCMCConfig(proposal=MCMCSimpleProposalConfig("normal_one_el")))
    xi, u, keys = mc.draw(state, 3)
    xi_ref, u_ref, keys_ref = otf.mcmc_draws(threefry.uint32(state.rng_state), 1, 3)
    np.testing.assert_array_equal(u.cpu().numpy(), u_ref)
    np.testing.assert_array_equal(threefry.uint32(keys), keys_ref)
    np.testing.assert_allclose(xi.cpu().numpy(), xi_ref.reshape(3, B, 3), rtol=4e-6, atol=1e-6)
