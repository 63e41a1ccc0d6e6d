"""Workspace sizing and per-walker records across chunks, through the C ABI:

* a caller-owned workspace of exactly dw_workspace_size(B, 1) carries dw_local_energy even when the FP64
  recompute is triggered for every walker (its double activations come from the same region), with
  results bit-identical to the library-owned buffers;
* dw_model_conditioning after a chunked dw_local_energy covers every walker: equal to the unchunked run.
"""
import numpy as np
import pytest
import torch

from workloads import system, walkers

pytestmark = pytest.mark.gpu


def _dev(x, dtype=torch.float32):
    return torch.as_tensor(np.asarray(x)).to("cuda", dtype).contiguous()


def test_caller_workspace_with_fp64_recompute_of_every_walker(cuda):
    from deeperwin_amd._native import FLAG_FP64_DET, NativeModel

    shape, params, R, Z, mapping = system("n2", seed=1)
    B = 64
    r = _dev(walkers(R, mapping, B, seed=11))
    Rd, Zd = _dev(R), _dev(Z)
    outs = []
    for caller_ws in (False, True):
        m = NativeModel(shape, "fp32", fp64_cond_threshold=1.0)  # every walker has kappa >= 1: all recomputed
        m.set_params(params)
        if caller_ws:
            ws = torch.empty(m.workspace_bytes(B, 1), dtype=torch.uint8, device="cuda")
            m.set_workspace(ws)
        e = torch.empty(B, device="cuda")
        fl = torch.zeros(B, dtype=torch.int32, device="cuda")
        m.local_energy(r, Rd, Zd, e, None, fl)
        torch.cuda.synchronize()
        outs.append((e.cpu().numpy(), fl.cpu().numpy()))
    assert np.all(outs[1][1] & FLAG_FP64_DET)
    np.testing.assert_array_equal(outs[0][0], outs[1][0])
    np.testing.assert_array_equal(outs[0][1], outs[1][1])


def test_conditioning_covers_every_chunk(cuda):
    from deeperwin_amd._native import NativeModel

    shape, params, R, Z, mapping = system("lih", seed=1)
    B = 96
    r = _dev(walkers(R, mapping, B, seed=12))
    Rd, Zd = _dev(R), _dev(Z)
    kap = []
    for chunk in (0, 32):
        m = NativeModel(shape, "fp32", chunk=chunk)
        m.set_params(params)
        e = torch.empty(B, device="cuda")
        m.local_energy(r, Rd, Zd, e)
        kap.append(m.conditioning(B).cpu().numpy())
    assert np.all(np.isfinite(kap[0])) and np.all(kap[0] >= 1.0)
    np.testing.assert_array_equal(kap[0], kap[1])
