"""Edge cases of the walker batch on the device (the reference's behaviour is per walker: a walker with an
electron exactly on a nucleus has E_pot = -inf and a NaN kinetic energy through the norm's derivative at 0,
hamiltonian.py:17-39, utils.py:262-273):

* such a walker gets a non-finite E_loc and DW_FLAG_NONFINITE, while every other walker of the batch is
  bit-identical to the same batch without it (no contamination through the batched kernels), for the
  FermiNet and dpe4 models;
* its log|psi|^2 stays finite (the value pass needs no derivative of the distance), as in the reference;
* energy moments of such a batch follow nanmean: the NaN walker is dropped.
"""
import numpy as np
import pytest
import torch

from workloads import system, system_dpe4, walkers

pytestmark = pytest.mark.gpu


def _dev(x, dtype=torch.float32):
    return torch.as_tensor(np.asarray(x)).to("cuda", dtype).contiguous()


@pytest.mark.parametrize("model", ["ferminet", "dpe4"])
def test_electron_on_a_nucleus(cuda, model):
    from deeperwin_amd._native import FLAG_NONFINITE, NativeModel, energy_moments

    shape, params, R, Z, mapping = (system("lih", seed=1) if model == "ferminet" else system_dpe4("lih", seed=1))
    B, hit = 64, 13
    r = walkers(R, mapping, B, seed=9)
    r_hit = r.copy()
    r_hit[hit, 0] = np.asarray(R[0], np.float32)  # electron 0 of walker 13 exactly on nucleus 0
    m = NativeModel(shape, "fp32")
    m.set_params(params)
    outs = []
    for rr in (r, r_hit):
        e = torch.empty(B, device="cuda")
        lp = torch.empty(B, device="cuda")
        fl = torch.zeros(B, dtype=torch.int32, device="cuda")
        lp2 = torch.empty(B, device="cuda")
        m.local_energy(_dev(rr), _dev(R), _dev(Z), e, lp, fl)
        m.logpsi(_dev(rr), _dev(R), _dev(Z), lp2)
        torch.cuda.synchronize()
        outs.append((e.cpu().numpy(), lp.cpu().numpy(), fl.cpu().numpy(), lp2.cpu().numpy()))
    (e0, lp0, f0, l20), (e1, lp1, f1, l21) = outs
    print(f"{model}: walker on a nucleus E_loc {e1[hit]}, log|psi|^2 {l21[hit]}, flags {f1[hit]}")
    assert not np.isfinite(e1[hit])
    assert f1[hit] & FLAG_NONFINITE
    assert np.isfinite(l21[hit])
    others = np.arange(B) != hit
    np.testing.assert_array_equal(e1[others], e0[others])
    np.testing.assert_array_equal(l21[others], l20[others])
    assert np.all(f1[others] == f0[others])
    mom = energy_moments(torch.tensor(e1, device="cuda")).cpu().numpy()
    if np.isnan(e1[hit]):  # nanmean drops the walker
        ref = e1[others].astype(np.float64)
        assert abs(mom[0] - ref.mean()) <= 1e-9 * max(1.0, abs(ref.mean()))
