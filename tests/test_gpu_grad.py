"""Parameter gradient on the GPU (dw_param_grad) and the loss drop-in, against the float64 oracle
(oracle/loss.py: autograd through the restated wavefunction, pinned to finite differences on the CPU).

Tolerances, per parameter leaf, on the Frobenius norm of the difference relative to the leaf's norm:
* FP64 form: 1e-6 (only the float32 inputs and outputs round);
* FP32 form: 1e-4 on well-conditioned walkers (conditioning < 1e3: the FP32 network moves log|psi|^2 by
  ~u kappa and the weight gradients sum ~B n FP32 products).
"""
import numpy as np
import pytest
import torch

from oracle import loss as ol
from oracle import wavefunction as owf
from workloads import system, walkers

pytestmark = pytest.mark.gpu


def _dev(x, dtype=torch.float32):
    return torch.as_tensor(np.asarray(x)).to("cuda", dtype).contiguous()


def _compare(grads, ref, rel):
    worst = 0.0
    for path, leaves in ref.items():
        for leaf, g_ref in leaves.items():
            g = grads[path][leaf].double().cpu().numpy()
            assert g.shape == g_ref.shape, (path, leaf)
            err = np.linalg.norm(g - g_ref) / max(np.linalg.norm(g_ref), 1e-30)
            worst = max(worst, err)
            assert err <= rel, (path, leaf, err)
    return worst


@pytest.mark.parametrize("name,B,D", [("lih", 8, 4), ("n2", 4, 3), ("odd17", 2, 2), ("mol32", 2, 2)])
def test_param_grad_fp64_form_matches_oracle(cuda, name, B, D):
    from deeperwin_amd._native import NativeModel

    shape, params, R, Z, mapping = system(name, n_det=D, seed=2)
    r = walkers(R, mapping, B, seed=6)
    c = np.random.default_rng(3).standard_normal(B)
    m = NativeModel(shape, "fp64")
    m.set_params(params)
    lp = torch.empty(B, device="cuda")
    g = m.param_grad(_dev(r), _dev(R), _dev(Z), _dev(c), out_lp=lp)
    torch.cuda.synchronize()
    grads = m.unpack_grad(g, params)
    ref = ol.param_grad(params, shape.n_up, shape.n_dn, r.astype(np.float64), R, c)
    worst = _compare(grads, ref, 1e-6)
    lp_ref = owf.log_psi_sqr(params, shape.n_up, shape.n_dn, torch.tensor(r, dtype=torch.float64), R)[1].numpy()
    np.testing.assert_allclose(lp.cpu().numpy(), lp_ref, rtol=1e-6, atol=1e-5)
    print(f"{name}: fp64 form, worst relative leaf error {worst:.2e}")


def test_param_grad_fp32_form_well_conditioned(cuda):
    from deeperwin_amd._native import NativeModel

    shape, params, R, Z, mapping = system("lih", seed=1)
    pool = walkers(R, mapping, 1024, seed=31)
    kap = owf.conditioning(params, shape.n_up, shape.n_dn, pool, R)
    r = pool[np.argsort(kap)[:64]]
    assert kap[np.argsort(kap)[:64]].max() < 1e3
    c = np.random.default_rng(4).standard_normal(len(r)) / len(r)
    m = NativeModel(shape, "fp32")
    m.set_params(params)
    g = m.param_grad(_dev(r), _dev(R), _dev(Z), _dev(c))
    torch.cuda.synchronize()
    worst = _compare(m.unpack_grad(g, params), ol.param_grad(params, shape.n_up, shape.n_dn, r.astype(np.float64),
                                                            R, c), 1e-4)
    print(f"lih fp32 form, 64 walkers: worst relative leaf error {worst:.2e}")


@pytest.mark.parametrize("center,width,name", [("mean", "std", "tanh"), ("median", "mae", "hard")])
def test_clip_energies_match_oracle(cuda, center, width, name):
    from deeperwin_amd._native import clip_energies

    rng = np.random.default_rng(7)
    e = (-40.0 + rng.standard_normal(1001) * 0.5).astype(np.float32)
    e[[3, 500]] = [-80.0, 10.0]  # outliers that the clipping moves
    e[17] = np.nan
    for state in (None, (float(np.nanmean(e)) + 0.1, 1.5)):
        st = None if state is None else torch.tensor(state, dtype=torch.float64, device="cuda")
        w, ec, stats = clip_energies(_dev(e), st, name, center, width, 5.0, True)
        (loss, (new_state, aux)) = ol.total_energy(e.astype(np.float64), state, name=name, center=center,
                                                   width_metric=width, clip_by=5.0)
        s = stats.cpu().numpy()
        np.testing.assert_allclose(s[[0, 1, 2, 3]], [loss, aux["E_mean"], aux["E_var"], aux["E_var_clipped"]],
                                   rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(s[[4, 5]], new_state, rtol=1e-12)
        ref_ec = aux["E_loc_clipped"]
        np.testing.assert_allclose(ec.cpu().numpy(), ref_ec.astype(np.float32), rtol=1e-6, equal_nan=True)
        ref_w = (ref_ec - aux["E_mean_clipped"]) / len(e)
        np.testing.assert_allclose(w.cpu().numpy(), ref_w.astype(np.float32), rtol=1e-5, atol=1e-9,
                                   equal_nan=True)


def test_value_and_grad_dropin_matches_oracle(cuda):
    from deeperwin_amd.config import ClippingConfig, ModelConfigFermiNet, OrbitalsConfigFermiNet, PhysicalConfig
    from deeperwin_amd.hamiltonian import build_local_energy
    from deeperwin_amd.loss import build_value_and_grad_func, init_clipping_state
    from deeperwin_amd.physics import geometry_lih
    from deeperwin_amd.wavefunction import build_log_psi_squared
    from oracle import hamiltonian as oh

    Rg, Zg = geometry_lih()
    phys = PhysicalConfig(R=np.asarray(Rg, np.float32).astype(np.float64).tolist(), Z=list(map(int, Zg)))
    cfg = ModelConfigFermiNet(orbitals=OrbitalsConfigFermiNet(n_determinants=2))
    log_psi_sqr, _, _, params, fixed = build_log_psi_squared(cfg, phys, rng_seed=0, precision="fp64")
    get_e = build_local_energy(log_psi_sqr)
    vg = build_value_and_grad_func(log_psi_sqr, get_e, ClippingConfig())
    R = np.asarray(phys.R)
    r = walkers(R, [0, 0, 0, 1], 6, seed=8).astype(np.float32)
    batch = (_dev(r), _dev(R), _dev(np.asarray(phys.Z, np.float32)), fixed)
    (loss, (state, aux)), grads = vg(params, init_clipping_state(), (phys.n_up, phys.n_dn), batch)
    e_ref = oh.local_energy(params, phys.n_up, phys.n_dn, r.astype(np.float64), R, np.asarray(phys.Z)).numpy()
    (loss_ref, (state_ref, aux_ref)), g_ref = ol.loss_gradient(params, phys.n_up, phys.n_dn, r.astype(np.float64),
                                                               R, e_ref, (0.0, 1e12))
    assert float(loss) == pytest.approx(loss_ref, rel=1e-6)
    np.testing.assert_allclose(state.cpu().numpy(), state_ref, rtol=1e-5)
    _compare(grads, g_ref, 1e-5)
