"""The precision, batching and serving forms at their full workload sizes on the GPU.

* Precision form (C4H2-like molecule from geometry_c4h2(3): 6 nuclei, 26 electrons, 16 dets, 4096 Gaussian
  walkers): BF16x3 GEMMs with the FP64 recompute of walkers whose conditioning exceeds 1e6.  Recomputed
  walkers equal the FP64 form; |dE_mean| against the FP64 form over all 4096 walkers <= 0.1 mHa; the FP64
  form matches the float64 oracle on a 3-walker sample.
* Batching form (batch_geometries(seed=4): 64 geometries of 4-42 electrons, 8 dets, 512 walkers each, 20
  MCMC steps + E_loc per unit, E_loc in chunks of 128 walkers): two graph replays equal two eager units bit
  for bit; the final E_loc of one walker of four geometries (the largest, n >= 40, among them) matches the
  float64 oracle.
* Serving form (24 requests of the seeded Poisson stream: 2-60 electrons, 256-4096 walkers, 50 steps):
  every request's energy, variance and acceptance are identical alone and in the mix.
"""
import numpy as np
import pytest
import torch

from oracle import hamiltonian as oh
from oracle import wavefunction as owf

pytestmark = pytest.mark.gpu

U32 = 2.0 ** -24


def _dev(x, dtype=torch.float32):
    return torch.as_tensor(np.asarray(x)).to("cuda", dtype).contiguous()


def test_precision_form_c4h2(cuda):
    from deeperwin_amd._native import FLAG_FP64_DET, NativeModel
    from workloads import system, walkers

    shape, params, R, Z, mapping = system("c4h2", seed=0)
    B = 4096
    r = _dev(walkers(R, mapping, B, seed=9))
    Rd, Zd = _dev(R), _dev(Z)
    mixed = NativeModel(shape, "bf16x3", fp64_cond_threshold=1e6, chunk=1024)
    mixed.set_params(params)
    ref = NativeModel(shape, "fp64", chunk=1024)
    ref.set_params(params)
    e_mix = torch.empty(B, device="cuda")
    e64 = torch.empty(B, device="cuda")
    fl = torch.zeros(B, dtype=torch.int32, device="cuda")
    print("c4h2: bf16x3 + fp64>1e6 on 4096 walkers", flush=True)
    mixed.local_energy(r, Rd, Zd, e_mix, None, fl)
    ref.local_energy(r, Rd, Zd, e64)
    torch.cuda.synchronize()
    a, b = e_mix.double().cpu().numpy(), e64.double().cpu().numpy()
    flagged = (fl.cpu().numpy() & FLAG_FP64_DET) != 0
    d_mean = np.mean(a) - np.mean(b)
    print(f"c4h2: {flagged.sum()} of {B} recomputed in FP64, dE_mean {d_mean * 1e3:+.4f} mHa", flush=True)
    assert flagged.any()
    np.testing.assert_allclose(a[flagged], b[flagged], rtol=1e-9, atol=1e-9)
    assert abs(d_mean) <= 1e-4
    idx = np.array([0, 1000, 3000])
    rs = r.cpu().numpy()[idx].astype(np.float64)
    e_ref = oh.local_energy(params, shape.n_up, shape.n_dn, rs, R, Z, vmapped=True).numpy()
    kappa = owf.conditioning(params, shape.n_up, shape.n_dn, rs, R)
    u64 = 2.0 ** -53
    assert np.all(np.abs(b[idx] - e_ref) <= (1 + np.abs(e_ref)) * (2e-7 + 64 * u64 * kappa)), (b[idx], e_ref)


def test_batching_form_64_geometries(cuda):
    from deeperwin_amd.batching import GeometryBatch
    from deeperwin_amd.params import ModelShape, init_params, truncate_params
    from deeperwin_amd.physics import batch_geometries

    geoms = [(np.asarray(R, np.float32).astype(np.float64), Z)
             for R, Z in batch_geometries(seed=4, n_geom=64, n_min=4, n_max=42)]
    kw = dict(n_det=8, walkers=512, n_steps=20, seed=3, n_el_max=42, n_ion_max=max(len(Z) for _, Z in geoms))
    a = GeometryBatch(geoms, n_streams=8, **kw)
    b = GeometryBatch(geoms, n_streams=8, **kw)
    for bt in (a, b):
        for gm in bt.geoms:
            gm.model.set_chunk(128)
    ga, gb = torch.Generator(device="cuda"), torch.Generator(device="cuda")
    ga.manual_seed(5)
    gb.manual_seed(5)
    print("batch64: eager units and graph capture", flush=True)
    a.refresh_draws(ga)
    a.run_eager()
    b.refresh_draws(gb)
    b.run_eager()
    b.capture()
    for rep in range(2):
        a.refresh_draws(ga)
        b.refresh_draws(gb)
        a.run_eager()
        b.replay()
        torch.cuda.synchronize()
        for k, (sa, sb) in enumerate(zip(a.state(), b.state())):
            for x, y in zip(sa, sb):
                np.testing.assert_array_equal(x, y, err_msg=f"replay {rep}, geometry {k}")
    n_els = [gm.shape.n_el for gm in a.geoms]
    assert max(n_els) >= 40 and min(n_els) <= 8
    big = init_params(ModelShape(42, 21, kw["n_ion_max"], 8), 3)
    pick = [int(np.argmax(n_els)), int(np.argmin(n_els)), 10, 40]
    for k in pick:
        gm = a.geoms[k]
        params = truncate_params(big, kw["n_ion_max"], 42, gm.shape)
        R, Z = geoms[k]
        r = gm.r[:1].cpu().numpy()
        print(f"batch64: oracle E_loc of geometry {k} (n = {gm.shape.n_el})", flush=True)
        e_ref = oh.local_energy(params, gm.shape.n_up, gm.shape.n_dn, r.astype(np.float64), R, Z,
                                vmapped=True).numpy()
        kappa = owf.conditioning(params, gm.shape.n_up, gm.shape.n_dn, r, R)
        err = np.abs(gm.e[:1].double().cpu().numpy() - e_ref)
        assert np.all(err <= (1 + np.abs(e_ref)) * (1e-4 + 4 * U32 * kappa)), (k, err, e_ref, kappa)


def test_serving_form_full_stream(cuda):
    from deeperwin_amd.serving import Server, SharedWeights, poisson_stream

    w = SharedWeights(n_el_max=60, n_ion_max=60, n_det=8, seed=4, chunk=128)
    reqs = poisson_stream(24, rate=50.0, seed=5)
    print(f"serving: {len(reqs)} requests, n_el {min(q.n_el for q in reqs)}-{max(q.n_el for q in reqs)}, walkers "
          f"{min(q.n_walkers for q in reqs)}-{max(q.n_walkers for q in reqs)}", flush=True)
    mixed = Server(w, max_active=8, slice_steps=10).serve(reqs, realtime=False)
    alone = [Server(w, max_active=1, slice_steps=50).serve([q], realtime=False)[0] for q in reqs]
    for x, y in zip(alone, sorted(mixed, key=lambda q: q.rid)):
        assert x.rid == y.rid
        assert x.energy_mean == y.energy_mean and x.energy_var == y.energy_var, (x, y)
        assert x.acceptance == y.acceptance
    assert all(np.isfinite(q.energy_mean) for q in mixed)
