"""dw_mcmc_steps_noalloc (every device buffer from the library's fixed scratch array) against dw_mcmc_steps.

Both entry points run the same kernels with the same launch shapes on the same inputs, so walkers,
log|psi|^2, ages, flags, per-step acceptance and the device chain record must be bit-identical.  Two
calls in a row, the second with fewer walkers and steps (the scratch keeps the first call's contents).
max_age = 2 and a stepsize interval of 2 make forced accepts and stepsize updates happen in both.
"""
import numpy as np
import pytest
import torch

from workloads import system, walkers

pytestmark = pytest.mark.gpu


def _dev(x, dtype=torch.float32):
    return torch.as_tensor(np.asarray(x)).to("cuda", dtype).contiguous()


def _run(m, R, Z, r0, normals, uniforms, noalloc):
    from deeperwin_amd.mcmc import make_chain

    B, S = r0.shape[0], normals.shape[0]
    r = _dev(r0)
    lp = torch.full((B,), -1000.0, device="cuda")
    age = torch.zeros(B, dtype=torch.int32, device="cuda")
    chain = make_chain(0.02, 0.3, 5)
    acc = torch.empty(S, device="cuda")
    flags = torch.zeros(B, dtype=torch.int32, device="cuda")
    m.mcmc_steps(r, _dev(R), _dev(Z), lp, age, _dev(normals), _dev(uniforms), chain, 2, 2, 0.5, 1e-2, 1.0,
                 out_acc=acc, out_flags=flags, noalloc=noalloc)
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in (r, lp, age, chain, acc, flags)]


def test_noalloc_mcmc_matches_mcmc_steps(cuda):
    from deeperwin_amd._native import DwError, NativeModel

    shape, params, R, Z, mapping = system("n2")
    m = NativeModel(shape, "fp32")
    m.set_params(params)
    rng = np.random.default_rng(11)
    for B, S in ((512, 5), (96, 3)):
        print(f"noalloc vs mcmc_steps: n2, B={B}, S={S}", flush=True)
        r0 = walkers(R, mapping, B, seed=B)
        normals = rng.standard_normal((S, B, shape.n_el, 3)).astype(np.float32)
        uniforms = rng.uniform(size=(S, B)).astype(np.float32)
        ref = _run(m, R, Z, r0, normals, uniforms, noalloc=False)
        out = _run(m, R, Z, r0, normals, uniforms, noalloc=True)
        for name, a, b in zip(("r", "log_psi_sqr", "age", "chain", "acc", "flags"), ref, out):
            np.testing.assert_array_equal(a, b, err_msg=name)
        assert (ref[5] & 4).any(), "max_age=2 should force some accepts"
    # more walkers than the scratch holds: a clear error before any launch
    big = 1 << 16
    r = torch.zeros((big, shape.n_el, 3), device="cuda")
    lp = torch.zeros(big, device="cuda")
    age = torch.zeros(big, dtype=torch.int32, device="cuda")
    from deeperwin_amd.mcmc import make_chain

    with pytest.raises(DwError, match="walkers"):
        m.mcmc_steps(r, _dev(R), _dev(Z), lp, age, None, None, make_chain(0.02), 2, 2, 0.5, 1e-2, 1.0, noalloc=True)
