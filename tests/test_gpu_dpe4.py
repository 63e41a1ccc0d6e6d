"""The dpe4 model (the reference's default, configuration.py:875-891; ferminet_embedding.py:97-188,
207-267) on the device, through the C ABI, against the FP64 oracle (pinned by tests/test_oracle_dpe4.py):

* FP64 form: log|psi|^2 and E_loc within the output rounding plus an FP64 conditioning term on
  LiH (4 electrons, two ion types), N2 (14, one type), a 17-electron random molecule and the 26-electron
  C4H2-like molecule, with non-zero biases everywhere;
* FP32 form: the kappa-scaled per-walker bounds and median bounds of test_gpu_parity.py;
* the drop-in build_log_psi_squared(ModelConfigDeepErwin4(), ...) -> log_psi_sqr / build_local_energy /
  run_mcmc_steps equals the ABI called directly, and a dpe4 MCMC step of the FP64 form takes the oracle's
  decisions (outside the float32 rounding margin of the threshold).
"""
import numpy as np
import pytest
import torch

from oracle import hamiltonian as oh
from oracle import wavefunction as owf
from workloads import system_dpe4, walkers

pytestmark = pytest.mark.gpu

U32 = 2.0 ** -24
U64 = 2.0 ** -53


def _native(shape, params, precision):
    from deeperwin_amd._native import NativeModel

    m = NativeModel(shape, precision)
    m.set_params(params)
    return m


def _dev(x, dtype=torch.float32):
    return torch.as_tensor(np.asarray(x)).to("cuda", dtype).contiguous()


def _run(m, shape, r, R, Z):
    B = r.shape[0]
    e = torch.empty(B, device="cuda")
    lp = torch.empty(B, device="cuda")
    lp2 = torch.empty(B, device="cuda")
    fl = torch.zeros(B, device="cuda", dtype=torch.int32)
    m.local_energy(_dev(r), _dev(R), _dev(Z), e, lp, fl)
    m.logpsi(_dev(r), _dev(R), _dev(Z), lp2)
    torch.cuda.synchronize()
    return e.cpu().numpy(), lp.cpu().numpy(), lp2.cpu().numpy(), fl.cpu().numpy()


def _ref(params, shape, r, R, Z):
    r64 = r.astype(np.float64)
    e_ref = oh.local_energy(params, shape.n_up, shape.n_dn, r64, R, Z).numpy()
    lp_ref = owf.log_psi_sqr(params, shape.n_up, shape.n_dn, torch.tensor(r64), R, Z)[1].numpy()
    kappa = owf.conditioning(params, shape.n_up, shape.n_dn, r64, R, Z)
    return e_ref, lp_ref, kappa


@pytest.mark.parametrize("name,B", [("lih", 8), ("n2", 6), ("odd17", 3), ("c4h2", 2)])
def test_dpe4_fp64_form_matches_oracle(cuda, name, B):
    shape, params, R, Z, mapping = system_dpe4(name, seed=2)
    print(f"dpe4 fp64 {name}: n = {shape.n_el}, Z_min {shape.z_min}, {shape.n_ion_types} ion types", flush=True)
    r = walkers(R, mapping, B, seed=3)
    m = _native(shape, params, "fp64")
    e, lp, lp2, fl = _run(m, shape, r, R, Z)
    e_ref, lp_ref, kappa = _ref(params, shape, r, R, Z)
    de, dl, dl2 = np.abs(e - e_ref), np.abs(lp - lp_ref), np.abs(lp2 - lp_ref)
    print(f"  E_loc ref {e_ref[:3]}, max |dE| {de.max():.2e}, max |dlp| {max(dl.max(), dl2.max()):.2e}, "
          f"kappa up to {kappa.max():.1e}", flush=True)
    # E_loc of both FP64 evaluations moves by ~u64 kappa times the size of the kinetic terms, which
    # exceed |E_loc| by up to ~1e3 on these walkers: 1e-13 kappa (about 1000 u64 kappa)
    assert np.all(de <= (1 + np.abs(e_ref)) * (2e-7 + 1e-13 * kappa))
    tl = 2e-5 * np.maximum(1, np.abs(lp_ref) / 10) + 64 * U64 * kappa
    assert np.all(dl <= tl) and np.all(dl2 <= tl)
    assert int(fl.max()) == 0


@pytest.mark.parametrize("name,B", [("lih", 16), ("n2", 8), ("odd17", 4), ("c4h2", 4)])
def test_dpe4_fp32_form_matches_oracle(cuda, name, B):
    shape, params, R, Z, mapping = system_dpe4(name, seed=4)
    print(f"dpe4 fp32 {name}: n = {shape.n_el}", flush=True)
    r = walkers(R, mapping, B, seed=5)
    m = _native(shape, params, "fp32")
    e, lp, lp2, fl = _run(m, shape, r, R, Z)
    e_ref, lp_ref, kappa = _ref(params, shape, r, R, Z)
    tol_e = (1.0 + np.abs(e_ref)) * (1e-4 + 4 * U32 * kappa)
    tol_l = 5e-5 * np.maximum(1.0, np.abs(lp_ref) / 10) + 4 * U32 * kappa
    de, dl = np.abs(e - e_ref), np.abs(lp2 - lp_ref)
    print(f"  max |dE| {de.max():.2e} (max err/tol {np.max(de / tol_e):.3f}), max |dlp| {dl.max():.2e}, "
          f"kappa up to {kappa.max():.1e}", flush=True)
    assert np.all(de <= tol_e) and np.all(dl <= tol_l) and np.all(np.abs(lp - lp_ref) <= tol_l)
    assert np.median(de) <= 1e-4 * (1 + np.median(np.abs(e_ref))) or np.median(de / tol_e) <= 0.25
    assert np.median(dl) <= 5e-4 or np.median(dl / tol_l) <= 0.25


def test_dpe4_dropin_and_mcmc_step(cuda):
    from deeperwin_amd.config import MCMCConfig, ModelConfigDeepErwin4, OrbitalsConfigFermiNet, PhysicalConfig
    from deeperwin_amd.hamiltonian import build_local_energy
    from deeperwin_amd.mcmc import MCMCState, MetropolisHastingsMonteCarlo, make_chain
    from deeperwin_amd.physics import geometry_lih
    from deeperwin_amd.wavefunction import build_log_psi_squared

    R, Z = geometry_lih()
    phys = PhysicalConfig(R=np.asarray(R, np.float32).astype(np.float64).tolist(), Z=list(map(int, Z)))
    cfg = ModelConfigDeepErwin4(orbitals=OrbitalsConfigFermiNet(n_determinants=4))
    log_psi_sqr, _, _, params, fixed = build_log_psi_squared(cfg, phys, rng_seed=5, precision="fp64")
    n_up, n_dn = phys.n_up, phys.n_dn
    B, S = 64, 3
    state = MCMCState.initialize_around_nuclei(B, phys, rng=6, stepsize=0.1)
    R64 = np.asarray(phys.R)
    r0 = state.r.cpu().numpy().astype(np.float64)
    _, lp = log_psi_sqr(params, n_up, n_dn, state.r, state.R, state.Z, fixed)
    lp_ref = owf.log_psi_sqr(params, n_up, n_dn, torch.tensor(r0), R64, phys.Z)[1].numpy()
    assert np.all(np.abs(lp.cpu().numpy() - lp_ref) <= 2e-5 * np.maximum(1, np.abs(lp_ref) / 10))
    get_e = build_local_energy(log_psi_sqr, False, False, None)
    e = get_e(params, (n_up, n_dn), state.r[:8].contiguous(), state.R, state.Z, fixed)
    e_ref = oh.local_energy(params, n_up, n_dn, r0[:8], R64, np.asarray(phys.Z)).numpy()
    assert np.all(np.abs(e.cpu().numpy() - e_ref) <= (1 + np.abs(e_ref)) * 1e-6)
    # MCMC through the drop-in with supplied draws: the FP64 form's decisions equal the oracle's
    rng = np.random.default_rng(8)
    normals = rng.standard_normal((S, B, n_up + n_dn, 3)).astype(np.float32)
    uniforms = rng.uniform(size=(S, B)).astype(np.float32)
    mc = MetropolisHastingsMonteCarlo(MCMCConfig(max_age=100, stepsize_update_interval=1000))
    new = state
    n_near = 0
    for k in range(S):
        r_before = new.r.cpu().numpy().astype(np.float64)
        new = mc.run_mcmc_steps(log_psi_sqr, new, params, n_up, n_dn, fixed, 1,
                                draws=(_dev(normals[k:k + 1]), _dev(uniforms[k:k + 1])))
        r_prop = (r_before + normals[k].astype(np.float64) * np.float64(np.float32(0.1))).astype(np.float32)
        lp0 = owf.log_psi_sqr(params, n_up, n_dn, torch.tensor(r_before), R64, phys.Z)[1].numpy()
        lp1 = owf.log_psi_sqr(params, n_up, n_dn, torch.tensor(r_prop.astype(np.float64)), R64, phys.Z)[1].numpy()
        p = np.exp(np.minimum(lp1 - lp0, 80.0))
        ref_accept = p > uniforms[k]
        dev_accept = new.walker_age.cpu().numpy() == 0
        near = np.abs(p - uniforms[k]) <= 1e-6 + p * (np.abs(lp0) + np.abs(lp1)) * U32
        n_near += int(near.sum())
        assert np.all((dev_accept == ref_accept) | near)
    print(f"dpe4 drop-in chain: {S} steps of {B} walkers, {n_near} decisions within the rounding margin")


def _compare_grads(grads, ref, rel):
    worst = 0.0
    for path, leaves in ref.items():
        for leaf, g_ref in leaves.items():
            g = grads[path][leaf].double().cpu().numpy()
            assert g.shape == g_ref.shape, (path, leaf)
            scale = np.linalg.norm(g_ref)
            err = np.linalg.norm(g - g_ref) / max(scale, 1e-30)
            if scale < 1e-12:  # e.g. rows of the ion table that no nucleus of the molecule uses
                assert np.abs(g).max() <= 1e-9, (path, leaf)
                continue
            worst = max(worst, err)
            assert err <= rel, (path, leaf, err)
    return worst


@pytest.mark.parametrize("name,B", [("lih", 4), ("n2", 3), ("odd17", 2)])
def test_dpe4_param_grad_fp64_form_matches_oracle(cuda, name, B):
    """dw_param_grad of the dpe4 model (every leaf: conv maps, same/diff and el-ion streams, ion maps, ion
    table, 1-el layers, orbitals) against the oracle's autograd gradient, which tests/test_oracle_dpe4.py pins
    to finite differences."""
    from deeperwin_amd._native import NativeModel
    from oracle import loss as ol

    shape, params, R, Z, mapping = system_dpe4(name, n_det=2, seed=3)
    print(f"dpe4 gradient {name}: n = {shape.n_el}, {B} walkers", flush=True)
    r = walkers(R, mapping, B, seed=6)
    c = np.random.default_rng(3).standard_normal(B)
    m = NativeModel(shape, "fp64")
    m.set_params(params)
    lp = torch.empty(B, device="cuda")
    g = m.param_grad(_dev(r), _dev(R), _dev(Z), _dev(c), out_lp=lp)
    torch.cuda.synchronize()
    ref = ol.param_grad(params, shape.n_up, shape.n_dn, r.astype(np.float64), R, c, Z)
    worst = _compare_grads(m.unpack_grad(g, params), ref, 1e-6)
    print(f"  worst relative leaf error {worst:.2e}", flush=True)


def test_dpe4_value_and_grad_dropin_matches_oracle(cuda):
    """build_value_and_grad_func over the dpe4 drop-in wavefunction: loss, clipping state and every gradient
    leaf against the oracle's total_energy and autograd gradient."""
    from deeperwin_amd.config import ClippingConfig, ModelConfigDeepErwin4, OrbitalsConfigFermiNet, PhysicalConfig
    from deeperwin_amd.hamiltonian import build_local_energy
    from deeperwin_amd.loss import build_value_and_grad_func, init_clipping_state
    from deeperwin_amd.physics import geometry_lih
    from deeperwin_amd.wavefunction import build_log_psi_squared
    from oracle import loss as ol

    Rg, Zg = geometry_lih()
    phys = PhysicalConfig(R=np.asarray(Rg, np.float32).astype(np.float64).tolist(), Z=list(map(int, Zg)))
    cfg = ModelConfigDeepErwin4(orbitals=OrbitalsConfigFermiNet(n_determinants=2))
    log_psi_sqr, _, _, params, fixed = build_log_psi_squared(cfg, phys, rng_seed=1, precision="fp64")
    get_e = build_local_energy(log_psi_sqr)
    vg = build_value_and_grad_func(log_psi_sqr, get_e, ClippingConfig())
    R, Z = np.asarray(phys.R), np.asarray(phys.Z)
    r = walkers(R, [0, 0, 0, 1], 6, seed=9).astype(np.float32)
    batch = (_dev(r), _dev(R), _dev(Z.astype(np.float32)), fixed)
    (loss, (state, aux)), grads = vg(params, init_clipping_state(), (phys.n_up, phys.n_dn), batch)
    e_ref = oh.local_energy(params, phys.n_up, phys.n_dn, r.astype(np.float64), R, Z).numpy()
    (loss_ref, (state_ref, _)), g_ref = ol.loss_gradient(params, phys.n_up, phys.n_dn, r.astype(np.float64), R,
                                                         e_ref, (0.0, 1e12), Z=Z)
    assert float(loss) == pytest.approx(loss_ref, rel=1e-6)
    np.testing.assert_allclose(state.cpu().numpy(), state_ref, rtol=1e-5)
    worst = _compare_grads(grads, g_ref, 1e-5)
    print(f"dpe4 value_and_grad drop-in: loss {float(loss):.6f}, worst relative leaf error {worst:.2e}", flush=True)
