"""dw_energy_moments (E_mean, E_var of get_observables, optimization/evaluation.py:143-151) against the
oracle's float64 nanmean restatement, which tests/test_oracle_physics.py pins to math.fsum: NaN walkers
dropped, +-inf kept (jnp.nanmean semantics), deterministic across calls."""
import numpy as np
import pytest
import torch

from oracle import hamiltonian as oh

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", ["plain", "large", "nan", "inf", "all_nan", "one"])
def test_energy_moments_match_oracle(cuda, case):
    from deeperwin_amd._native import energy_moments

    rng = np.random.default_rng(len(case))
    n = {"large": 1 << 20, "one": 1}.get(case, 4099)
    e = (-109.3 + 0.7 * rng.standard_normal(n)).astype(np.float32)
    if case == "nan":
        e[rng.choice(n, 37, replace=False)] = np.nan
    if case == "inf":
        e[5] = np.inf
        e[9] = np.nan
    if case == "all_nan":
        e[:] = np.nan
    out = energy_moments(torch.tensor(e, device="cuda")).cpu().numpy()
    with np.errstate(invalid="ignore"), __import__("warnings").catch_warnings():
        __import__("warnings").simplefilter("ignore")
        ref = oh.mean_and_var(e)
    print(case, out, ref)
    for got, want in zip(out, ref):
        if np.isnan(want) or np.isinf(want):
            assert (np.isnan(got) and np.isnan(want)) or got == want
        else:
            assert abs(got - want) <= 1e-12 * max(1.0, abs(want)) * np.sqrt(n)
    again = energy_moments(torch.tensor(e, device="cuda")).cpu().numpy()
    np.testing.assert_array_equal(out, again)
