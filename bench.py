#!/usr/bin/env python
"""Throughput of the walker-parallel VMC inner loop on one MI355X.

One step = the reference's evaluation unit for the N2 workload: 20 Metropolis-Hastings steps (Gaussian
proposals, stepsize 0.02 bohr, draws generated on the device from a seeded generator) followed by the
local energy of every walker with the full Laplacian over 3 n_el = 42 directions, for 2048 walkers of
N2 (Z = [7, 7], 2.07 bohr, 14 electrons, 7 up) with a FermiNet-style net of 4 iterations, 256 / 32 widths,
16 full determinants, random init from seed 1.  Walkers start Gaussian (sigma 1 bohr) around the nuclei
and the chain continues from step to step.

value = walkers * steps / seconds  ("walker local-energies/s"; one walker local energy = the chain steps
plus one Laplacian evaluation for that walker).

Extra fields: `roofline` for the dominant kernel (the 1-el embedding GEMM in Laplacian mode, timed with
HIP events on its launch stream inside the timed region), `cpu_baseline` (the CPU oracle in FP32, same
per-walker work on a bounded walker sample) and `accuracy` (FP64 oracle on a walker sample of the final
state).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "walker local-energies/s at |dE|<=0.1 mHa vs FP64; p50/p99 request latency"
PEAK_TFLOPS = {"fp32": 157.3, "bf16x3": 2500.0, "bf16": 2500.0}
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", default="fp32", choices=["fp32", "bf16x3", "bf16"])
    ap.add_argument("--walkers", type=int, default=2048)
    ap.add_argument("--mcmc-steps", type=int, default=20)
    ap.add_argument("--cpu-walkers", type=int, default=512)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--workload", default="n2", choices=["n2", "c4h2-precision", "serving", "batch64", "copy", "grad",
                                                          "grad-dpe4", "dpe4", "vendor-gemm"],
                    help="n2: the headline line; c4h2-precision: the precision-form study (E_loc only); "
                         "serving: Poisson request stream with continuous batching; batch64: 64 geometries "
                         "under shared weights in one graph launch per unit; copy: a plain device copy "
                         "of known size (calibrates the HBM byte counters); vendor-gemm: the vendor FP32 GEMM "
                         "(torch.matmul) on the N2 path's GEMM shapes, a reference point for the MFMA fraction")
    ap.add_argument("--requests", type=int, default=24, help="serving: requests in the stream")
    ap.add_argument("--rate", type=float, default=50.0, help="serving: offered load, requests/s")
    ap.add_argument("--fp64-threshold", type=float, default=1e6)
    ap.add_argument("--eager", action="store_true",
                    help="n2: launch the MCMC steps eagerly instead of replaying them as one HIP graph")
    ap.add_argument("--tune", action="store_true",
                    help="n2: pick launch shapes with the autotuner first (default: the built-in heuristics, so "
                         "every run times the same kernels)")
    ap.add_argument("--burn-in", type=int, default=300, help="c4h2-precision: MCMC steps before the second study")
    return ap.parse_args()


def traffic_from_profiles():
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3 --pmc summary, if any."""
    f = ROOT / "profiles" / "gemm_1el_lap_n2_traffic.txt"
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        if line.startswith("bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def precision_study(args):
    """C4H2-like molecule (4 C on a 2.6 bohr rhombus + 2 H), 26 electrons, 16 dets, 4096 walkers: E_loc of
    every walker in each precision form, timed, and compared with the FP64 form on the SAME walkers (the
    FP64 form matches the FP64 CPU oracle to ~1e-5 Ha; see tests/test_gpu_parity.py).  Two walker sets:
    the Gaussian initial walkers, and the same walkers after `--burn-in` Metropolis-Hastings steps of the
    FP32 form (stepsize 0.02 with the reference's adaptation), i.e. closer to |psi|^2."""
    import torch

    from deeperwin_amd._native import FLAG_FP64_DET, NativeModel
    from deeperwin_amd.mcmc import make_chain
    from deeperwin_amd.params import ModelShape, init_params
    from deeperwin_amd.physics import generate_el_ion_mapping, geometry_c4h2, init_walkers_gaussian

    R, Z = geometry_c4h2(3)
    n, u, D, B = 26, 13, 16, args.walkers if args.walkers != 2048 else 4096
    shape = ModelShape(n_el=n, n_up=u, n_ion=6, n_det=D)
    params = init_params(shape, seed=2)
    r0 = init_walkers_gaussian(np.random.default_rng(2), B, R, generate_el_ion_mapping(R, Z, n, u))
    dev = "cuda"
    r = torch.as_tensor(r0, dtype=torch.float32, device=dev).contiguous()
    Rd = torch.as_tensor(R, dtype=torch.float32, device=dev)
    Zd = torch.as_tensor(Z, dtype=torch.float32, device=dev)

    def run(prec, thr, reps):
        m = NativeModel(shape, prec, fp64_cond_threshold=thr, chunk=1024)  # bounds the workspace
        m.set_params(params)
        e = torch.empty(B, device=dev)
        fl = torch.zeros(B, dtype=torch.int32, device=dev)
        m.local_energy(r, Rd, Zd, e, None, fl)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            m.local_energy(r, Rd, Zd, e, None, fl)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / reps
        return e.double().cpu().numpy(), fl.cpu().numpy(), dt

    def study():
        e64, _, t64 = run("fp64", 0.0, max(1, args.steps // 2))
        forms = {"fp64": {"walker_eloc_per_s": round(B / t64, 1), "ms": round(t64 * 1e3, 3), "dE_mean_mHa": 0.0,
                          "median_abs_dE": 0.0, "max_abs_dE": 0.0, "fp64_recomputed": B}}
        for prec, thr in [("fp32", 0.0), ("fp32", args.fp64_threshold), ("bf16x3", 0.0),
                          ("bf16x3", args.fp64_threshold), ("bf16", 0.0), ("bf16", args.fp64_threshold)]:
            e, fl, dt = run(prec, thr, args.steps)
            de = e - e64
            key = prec + ("+fp64>%.0e" % thr if thr > 0 else "")
            forms[key] = {"walker_eloc_per_s": round(B / dt, 1), "ms": round(dt * 1e3, 3),
                          "dE_mean_mHa": round(float(np.mean(de)) * 1e3, 4),
                          "median_abs_dE": float(np.median(np.abs(de))), "max_abs_dE": float(np.max(np.abs(de))),
                          "fp64_recomputed": int(((fl & FLAG_FP64_DET) != 0).sum()) if thr > 0 else 0}
        best = max((k for k in forms if abs(forms[k]["dE_mean_mHa"]) <= 0.1),
                   key=lambda k: forms[k]["walker_eloc_per_s"], default=None)
        return forms, best, float(np.mean(e64))

    forms_g, best_g, emean_g = study()
    # burn-in with the FP32 form
    m = NativeModel(shape, "fp32")
    m.set_params(params)
    lp = torch.empty(B, device=dev)
    age = torch.zeros(B, dtype=torch.int32, device=dev)
    chain = make_chain(0.02)
    gen = torch.Generator(device=dev)
    gen.manual_seed(2)
    left = args.burn_in
    while left > 0:
        k = min(50, left)
        normals = torch.randn((k, B, n, 3), generator=gen, device=dev)
        uniforms = torch.rand((k, B), generator=gen, device=dev)
        m.mcmc_steps(r, Rd, Zd, lp, age, normals, uniforms, chain, 100, 100, 0.5, 1e-2, 1.0)
        left -= k
    torch.cuda.synchronize()
    forms_b, best_b, emean_b = study()
    base = forms_b["fp32"]["walker_eloc_per_s"]
    out = {"metric": METRIC, "value": forms_b[best_b]["walker_eloc_per_s"] if best_b else None,
           "unit": "walker local-energies/s", "n_gpus": 1, "steps": args.steps, "warmup": 1,
           "ms_per_step": forms_b[best_b]["ms"] if best_b else None, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": best_b, "data": "synthetic",
           "config": {"workload": f"C4H2-like (4 C on a seed-3 rhombus, 2.6 bohr sides, + 2 H), 26 electrons, 16 dets, "
                                  f"{B} walkers after {args.burn_in} MCMC steps, E_loc only", "walkers": B,
                      "burn_in_steps": args.burn_in},
           "reference": {"form": "fp64 (GPU, pinned to the FP64 CPU oracle)", "energy_mean_gaussian": emean_g,
                         "energy_mean_after_burn_in": emean_b},
           "forms_after_burn_in": forms_b, "best_form_within_0.1mHa_after_burn_in": best_b,
           "speedup_vs_plain_fp32_after_burn_in": round(forms_b[best_b]["walker_eloc_per_s"] / base, 3) if best_b else None,
           "forms_gaussian_walkers": forms_g, "best_form_within_0.1mHa_gaussian": best_g}
    print(json.dumps(out))


def copy_calibration(args):
    """A plain device-to-device copy of 256 MiB, launched 5 times: FETCH_SIZE / WRITE_SIZE of this kernel
    against its known 256 MiB read + 256 MiB write give the scale of the HBM byte counters."""
    import torch

    x = torch.ones(64 << 20, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"workload": "copy", "bytes_read_per_launch": x.numel() * 4,
                      "bytes_written_per_launch": y.numel() * 4, "launches": 5}))


def vendor_gemm_calibration(args):
    """The vendor library's FP32 GEMM (torch.matmul, no TF32) on the GEMM shapes of the N2 workload (2048
    walkers, 14 electrons, 7 up, 48 channel rows per electron in the Laplacian pass, widths 256/32, 16 full
    dets): what a tuned plain GEMM reaches on these shapes, without the fused epilogues (V sums, tanh rule)
    of this library's kernels.  Rows are the padded ones the kernels run."""
    import torch

    torch.backends.cuda.matmul.allow_tf32 = False
    B, n, nu, Cp, W1, W2, D = 2048, 14, 7, 48, 256, 32, 16
    shapes = [("laplacian 1-el GEMM", B * n * Cp, W1 + 2 * W2, W1),
              ("laplacian orbital GEMM (one spin)", B * nu * Cp, W1, D * n),
              ("value 1-el GEMM", B * n, W1 + 2 * W2, W1),
              ("value orbital GEMM (one spin)", B * nu, W1, D * n)]
    rows = []
    g = torch.Generator(device="cuda").manual_seed(0)
    for name, M, K, N in shapes:
        A = torch.randn(M, K, device="cuda", generator=g)
        Bt = torch.randn(N, K, device="cuda", generator=g)
        C = torch.empty(M, N, device="cuda")
        for _ in range(3):
            torch.matmul(A, Bt.t(), out=C)
        torch.cuda.synchronize()
        reps = 20
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            torch.matmul(A, Bt.t(), out=C)
        e1.record()
        torch.cuda.synchronize()
        us = e0.elapsed_time(e1) / reps * 1e3
        tf = 2.0 * M * K * N / us / 1e6
        rows.append({"gemm": name, "M": M, "K": K, "N": N, "us": round(us, 1), "tflops": round(tf, 2),
                     "frac_fp32_peak": round(tf / PEAK_TFLOPS["fp32"], 3)})
        del A, Bt, C
    print(json.dumps({"workload": "vendor-gemm", "library": "torch.matmul (FP32, TF32 off)", "gemms": rows}))


def grad_bench(args):
    """One optimisation step's gradient on the N2 workload (2048 walkers, 16 dets): E_loc, the clipped
    statistics and the parameter gradient G = sum_b c_b d log_psi_sqr(r_b) / d theta (the reference's
    value_and_grad of the clipped total energy, optimization/loss_function.py:75-158), timed per step;
    the gradient pass alone is timed too.  CPU baseline: the same step (E_loc by 3n HVPs in vmapped chunks,
    gradient by autograd) in FP32 PyTorch on a walker sample."""
    import torch

    from deeperwin_amd._native import NativeModel, clip_energies
    from deeperwin_amd.batching import executed_forward_flops
    from deeperwin_amd.params import ModelShape, init_params
    from deeperwin_amd.physics import generate_el_ion_mapping, geometry_n2, init_walkers_gaussian

    R, Z = geometry_n2()
    n, u, D, B = 14, 7, 16, args.walkers
    dpe4 = args.workload == "grad-dpe4"
    shape = (ModelShape(n_el=n, n_up=u, n_ion=2, n_det=D, embedding="dpe4", z_min=7, n_ion_types=1) if dpe4
             else ModelShape(n_el=n, n_up=u, n_ion=2, n_det=D))
    params = init_params(shape, seed=1)
    r0 = init_walkers_gaussian(np.random.default_rng(1), B, R, generate_el_ion_mapping(R, Z, n, u))
    m = NativeModel(shape, args.precision)
    m.set_params(params)
    r = torch.as_tensor(r0, dtype=torch.float32, device="cuda").contiguous()
    Rd = torch.as_tensor(R, dtype=torch.float32, device="cuda")
    Zd = torch.as_tensor(Z, dtype=torch.float32, device="cuda")
    e = torch.empty(B, device="cuda")
    state = torch.tensor([0.0, 1e12], dtype=torch.float64, device="cuda")

    def step():
        m.local_energy(r, Rd, Zd, e)
        w, _, stats = clip_energies(e, state)
        return m.param_grad(r, Rd, Zd, w), w

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g, w = step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(args.steps):
        m.param_grad(r, Rd, Zd, w)
    b.record()
    torch.cuda.synchronize()
    g_ms = a.elapsed_time(b) / args.steps
    # FLOPs of the gradient pass: one value pass, plus about twice its GEMM work backwards (input and
    # weight gradients of every GEMM), counted as three value passes
    gflop = 3 * B * executed_forward_flops(shape) / 1e9
    peak = PEAK_TFLOPS.get(args.precision, 157.3)
    tflops = None if dpe4 else round(gflop / (g_ms / 1e3) / 1e3, 2)  # no FLOP model of the dpe4 gradient
    out = {"metric": "walker gradient contributions/s (E_loc + clipped loss + parameter gradient)",
           "value": round(B / dt, 1), "unit": "walkers/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(dt * 1e3, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": args.precision, "data": "synthetic",
           "config": {"workload": ("N2, 14 electrons, 16 dets, " + ("dpe4 model" if dpe4 else "256/32 widths")
                                   + ", seed 1; one optimisation step's gradient (E_loc with the full Laplacian, "
                                     "tanh clipping, parameter gradient)"),
                      "walkers": B, "n_params": int(m.lib.dw_param_count(m.handle))},
           "gradient_pass_ms": round(g_ms, 3),
           "gradient_pass_tflops": tflops,
           "gradient_pass_frac_of_peak": None if dpe4 else round(tflops / peak, 4)}
    if not args.no_cpu_baseline:
        import torch as _t

        from oracle import hamiltonian as oh
        from oracle import loss as ol

        cores = len(os.sched_getaffinity(0))
        _t.set_num_threads(cores)
        k = 64
        rs = r0[:k].astype(np.float32)

        def cpu_step():
            ek = oh.local_energy(params, u, n - u, rs, R, Z, dtype=_t.float32, vmapped=True).numpy()
            ol.loss_gradient(params, u, n - u, rs, R, ek, (0.0, 1e12), Z=Z)

        cpu_step()
        ts = []
        for _ in range(3):
            c0 = time.perf_counter()
            cpu_step()
            ts.append(time.perf_counter() - c0)
        out["cpu_baseline"] = {"value": round(k / float(np.median(ts)), 3), "unit": "walkers/s", "cores": cores,
                               "kind": "port", "reps": 3,
                               "sample": f"{k} walkers: E_loc (FP32, vmapped chunk) + clipping + autograd gradient "
                                         f"(float64 restatement); median of 3 after a warm-up",
                               "baseline_only": True}
    print(json.dumps(out))


def serving_bench(args):
    """Poisson stream of energy-evaluation requests (2-60 electrons, 256-4096 walkers log-uniform, 50 MCMC
    steps + mean E_loc each) under one shared max-size weight set, served with continuous batching; the
    same requests are then run one at a time for the isolated latencies, and every request's result is
    compared between the two runs."""
    import dataclasses

    import torch

    from deeperwin_amd.serving import Server, SharedWeights, poisson_stream, summarize

    w = SharedWeights(n_el_max=60, n_ion_max=60, n_det=8, seed=4, precision=args.precision, chunk=128)
    reqs = poisson_stream(args.requests, rate=args.rate, seed=5)
    # warm-up: every request once (loads each kernel variant and touches the allocator for every size)
    warm = [dataclasses.replace(q, arrival=0.0, n_steps=2) for q in reqs]
    Server(w, max_active=1, slice_steps=2).serve(warm, realtime=False)
    iso_res = {}
    for q in reqs:
        q0 = dataclasses.replace(q, arrival=0.0)
        iso_res[q.rid] = Server(w, max_active=1, slice_steps=q.n_steps).serve([q0], realtime=False)[0]
        print(f"isolated request {q.rid}: n_el={q.n_el} walkers={q.n_walkers} "
              f"{iso_res[q.rid].latency * 1e3:.1f} ms", file=sys.stderr, flush=True)
    isolated = {k: v.latency for k, v in iso_res.items()}
    busy = sum(isolated.values())
    capacity = len(reqs) / busy
    # concurrency: requests in flight at once, swept at the offered rate; the best is used for both rates
    sweep = {}
    for ma in (2, 4, 8):
        res = Server(w, max_active=ma, slice_steps=10).serve(reqs)
        sweep[ma] = summarize(res, isolated)["makespan_s"]
        print(f"max_active {ma}: makespan {sweep[ma]:.3f} s", file=sys.stderr, flush=True)
    max_active = min(sweep, key=sweep.get)
    runs = {}
    for label, rate in (("offered_%g_per_s" % args.rate, args.rate), ("at_0.7_capacity", 0.7 * capacity)):
        stream = [dataclasses.replace(q, arrival=q.arrival * args.rate / rate) for q in reqs]
        if label == "at_0.7_capacity":  # below capacity the latency tail, not the makespan, picks the concurrency
            tails = {}
            for ma in (2, 4, 8):
                tails[ma] = summarize(Server(w, max_active=ma, slice_steps=10).serve(stream), isolated)["p99_latency_s"]
            ma_run = min(tails, key=tails.get)
        else:
            ma_run = max_active
        res = Server(w, max_active=ma_run, slice_steps=10).serve(stream)
        summ = summarize(res, isolated)
        summ["max_active"] = ma_run
        walkers = sum(q.n_walkers for q in reqs)
        summ["rate_per_s"] = round(rate, 3)
        summ["walker_eloc_per_s"] = round(walkers / summ["makespan_s"], 1)
        summ["identical_to_isolated"] = all(
            r.energy_mean == iso_res[r.rid].energy_mean and r.acceptance == iso_res[r.rid].acceptance for r in res)
        runs[label] = summ
        print(f"{label}: {json.dumps(summ)}", file=sys.stderr, flush=True)
    head = runs["offered_%g_per_s" % args.rate]
    out = {"metric": METRIC, "value": head["walker_eloc_per_s"], "unit": "walker local-energies/s",
           "n_gpus": 1, "steps": len(reqs), "warmup": len(warm), "ms_per_step": round(1e3 * head["makespan_s"] / len(reqs), 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": args.precision,
           "data": "synthetic",
           "config": {"workload": f"serving: {len(reqs)} Poisson requests at {args.rate:g}/s, 2-60 electrons, "
                                  "256-4096 walkers, 50 MCMC steps + mean E_loc, 8 dets, shared weights",
                      "max_active": max_active, "slice_steps": 10, "max_active_sweep_makespan_s": sweep},
           "isolated": {"p50_s": head["isolated_p50_s"], "p99_s": head["isolated_p99_s"],
                        "sum_s": round(busy, 4), "capacity_req_per_s": round(capacity, 3)},
           "runs": runs}
    print(json.dumps(out))


def batch_bench(args):
    """64 seeded geometries (electron counts uniform in [4, 42], Z in {1, 6, 7, 8}, nearest-neighbour
    distances 1.8-2.8 bohr), shared random-init weights (256/32 widths, 8 dets, truncated per geometry),
    512 walkers each; one unit = 20 MCMC steps + E_loc for every geometry, recorded as one HIP graph.
    Launch shapes come from the autotuner (one representative geometry per electron-count bucket).
    Compared with the same batch run eagerly geometry after geometry on one stream with the built-in
    launch-shape heuristics."""
    import torch

    from deeperwin_amd._native import NativeModel
    from deeperwin_amd.batching import GeometryBatch
    from deeperwin_amd.params import ModelShape, init_params, truncate_params
    from deeperwin_amd.physics import batch_geometries, generate_el_ion_mapping, init_walkers_gaussian
    from deeperwin_amd.tuning import BUCKETS, bucket_of, tune_buckets

    geoms = batch_geometries(seed=4, n_geom=64, n_min=4, n_max=42)
    W, S, chunk = 512, args.mcmc_steps, 128
    n_el_max, n_ion_max, D = 42, max(len(Z) for _, Z in geoms), 8
    big = init_params(ModelShape(n_el_max, (n_el_max + 1) // 2, n_ion_max, D), 3)
    reps = {}
    for R, Z in geoms:  # the largest geometry of each bucket represents it
        n = int(np.sum(Z))
        b = bucket_of(n)
        if b not in reps or n > int(np.sum(reps[b][1])):
            reps[b] = (R, Z)
    by_n = {int(np.sum(Z)): (R, Z) for R, Z in reps.values()}

    def make_model(n):
        R, Z = by_n[n]
        shape = ModelShape(n, (n + 1) // 2, len(Z), D)
        m = NativeModel(shape, args.precision, chunk=chunk)
        m.set_params(truncate_params(big, n_ion_max, n_el_max, shape))
        r0 = init_walkers_gaussian(np.random.default_rng(0), W, R, generate_el_ion_mapping(R, Z, n, shape.n_up))
        return (m, torch.as_tensor(r0, dtype=torch.float32, device="cuda").contiguous(),
                torch.as_tensor(R, dtype=torch.float32, device="cuda"),
                torch.as_tensor(Z, dtype=torch.float32, device="cuda"))

    t0 = time.perf_counter()
    table, tune_log = tune_buckets(make_model, sorted(by_n), reps=2)
    t_tune = time.perf_counter() - t0
    print(f"tuned {len(by_n)} buckets in {t_tune:.1f} s", file=sys.stderr, flush=True)

    def unit_time(bt, fn, k):
        gen = torch.Generator(device="cuda")
        gen.manual_seed(7)
        for _ in range(max(1, args.warmup)):
            bt.refresh_draws(gen)
            fn()
        torch.cuda.synchronize()
        a = time.perf_counter()
        for _ in range(k):
            bt.refresh_draws(gen)
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - a) / k

    # plain form: eager, one stream, built-in heuristics
    plain = GeometryBatch(geoms, n_det=D, walkers=W, n_steps=S, precision=args.precision, seed=3, n_streams=1,
                          n_el_max=n_el_max, n_ion_max=n_ion_max)
    for gm in plain.geoms:
        gm.model.set_chunk(chunk)
    t_plain = unit_time(plain, plain.run_eager, max(1, args.steps // 2))
    print(f"plain eager unit {t_plain * 1e3:.1f} ms", file=sys.stderr, flush=True)
    del plain
    torch.cuda.empty_cache()

    bt = GeometryBatch(geoms, n_det=D, walkers=W, n_steps=S, precision=args.precision, seed=3, n_streams=8,
                       table=table, n_el_max=n_el_max, n_ion_max=n_ion_max)
    for gm in bt.geoms:
        gm.model.set_chunk(chunk)
    t_eager = unit_time(bt, bt.run_eager, max(1, args.steps // 2))
    print(f"tuned multi-stream eager unit {t_eager * 1e3:.1f} ms", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    bt.capture()
    t_capture = time.perf_counter() - t0
    t_graph = unit_time(bt, bt.replay, args.steps)
    print(f"graph unit {t_graph * 1e3:.1f} ms (capture {t_capture:.1f} s)", file=sys.stderr, flush=True)

    # accuracy: FP64 form on the final walkers of 8 geometries (every 8th)
    acc = []
    for gm in bt.geoms[::8]:
        m64 = NativeModel(gm.shape, "fp64", chunk=chunk)
        m64.set_params(truncate_params(big, n_ion_max, n_el_max, gm.shape))
        e64 = torch.empty(W, device="cuda")
        m64.local_energy(gm.r, gm.R, gm.Z, e64)
        torch.cuda.synchronize()
        a, b = gm.e.double().cpu().numpy(), e64.double().cpu().numpy()
        ok = np.isfinite(a) & np.isfinite(b)
        acc.append({"n_el": gm.shape.n_el, "dE_mean_mHa": round(float(np.mean(a[ok]) - np.mean(b[ok])) * 1e3, 4),
                    "median_abs_dE": float(np.median(np.abs(a[ok] - b[ok])))})
    walkers = W * len(geoms)
    flops = bt.flops_per_unit(executed=True)
    flops_scope = bt.flops_per_unit()
    peak = PEAK_TFLOPS.get(args.precision, 157.3)
    achieved = flops / t_graph / 1e12
    out = {"metric": METRIC, "value": round(walkers / t_graph, 1), "unit": "walker local-energies/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(t_graph * 1e3, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": args.precision,
           "data": "synthetic",
           "config": {"workload": "64 seeded geometries, 4-42 electrons, shared weights 256/32, 8 dets, 512 walkers "
                                  f"each; one unit = {S} MCMC steps + E_loc per geometry in one graph launch",
                      "geometries": len(geoms), "walkers_per_geometry": W, "mcmc_steps": S,
                      "n_el": sorted(int(np.sum(Z)) for _, Z in geoms), "eloc_chunk": chunk},
           "roofline": {"bound": "mfma", "achieved": round(achieved, 2), "peak": peak, "unit": "TFLOP/s",
                        "frac": round(achieved / peak, 4), "traffic": None,
                        "scope": "whole unit: FLOPs of the algorithm as run (n_steps forwards + one Laplacian pass "
                                 "per walker, spin means once per walker) / wall time of a graph replay",
                        "scope_table_frac": round(flops_scope / t_graph / 1e12 / peak, 4)},
           "plain_eager_ms": round(t_plain * 1e3, 3), "tuned_eager_ms": round(t_eager * 1e3, 3),
           "graph_capture_s": round(t_capture, 2), "speedup_vs_plain": round(t_plain / t_graph, 3),
           "tuning": {"seconds": round(t_tune, 1), "table": json.loads(table.to_json()), "log": tune_log},
           "accuracy_vs_fp64_form": acc}
    print(json.dumps(out))


def whole_path_fractions(shape, n_steps, walkers, unit_s, peak_tflops):
    """Whole-unit FLOP rates: `executed` counts the algorithm as this library runs it (spin means multiplied
    once per walker), `scope_table` the reference's layer shapes (every electron multiplies the full
    symmetric feature vector).  Only the executed fraction is a utilisation; the other is the reference
    work done per second, as a fraction of the same peak."""
    from deeperwin_amd.batching import executed_unit_flops, unit_flops

    out = {}
    for key, f in (("executed", executed_unit_flops), ("scope_table", unit_flops)):
        fl = walkers * f(shape, n_steps)
        rate = fl / unit_s / 1e12
        out[key] = {"gflop_per_walker_unit": round(f(shape, n_steps) / 1e9, 4), "tflops": round(rate, 2),
                    "frac_of_peak": round(rate / peak_tflops, 4)}
    return out


def cpu_baseline(params, n_up, n_dn, r0, R, Z, n_steps, k, reps=3, chunk=64):
    """The reference algorithm on the host cores, FP32 PyTorch CPU (the oracle's restatement): n_steps
    Metropolis steps over k walkers (batched forward), then E_loc as 3 n Hessian-vector products per walker
    in chunks of `chunk` walkers batched by vmap (hamiltonian.py:234-266, 280).  One warm-up on a
    `chunk`-walker slice, then the median of `reps` timed repetitions; MCMC and E_loc are timed
    separately and their sum gives the end-to-end rate."""
    import torch as _t

    from oracle import hamiltonian as oh
    from oracle import mcmc as om
    from oracle import wavefunction as owf

    cores = len(os.sched_getaffinity(0))
    _t.set_num_threads(cores)
    rng = np.random.default_rng(7)
    n = n_up + n_dn

    def f(x):
        return owf.log_psi_sqr(params, n_up, n_dn, _t.tensor(x), R, Z, dtype=_t.float32)[1].numpy()

    def run(m):
        rc = r0[:m].astype(np.float32)
        normals = rng.standard_normal((n_steps, m, n, 3)).astype(np.float32)
        uniforms = rng.uniform(size=(n_steps, m)).astype(np.float32)
        st = om.ChainState(r=rc, log_psi_sqr=np.full(m, -1000.0, np.float32), walker_age=np.zeros(m, np.int64),
                           stepsize=0.02)
        a = time.perf_counter()
        st, _, _ = om.mcmc_steps(f, st, normals, uniforms, max_age=100, update_interval=100)
        b = time.perf_counter()
        oh.local_energy(params, n_up, n_dn, st.r, R, Z, max_batch_size=chunk, dtype=_t.float32, vmapped=True)
        c = time.perf_counter()
        return b - a, c - b

    run(min(chunk, k))  # warm-up
    t_mcmc, t_eloc = zip(*[run(k) for _ in range(reps)])
    tm, te = float(np.median(t_mcmc)), float(np.median(t_eloc))
    return {"value": round(k / (tm + te), 3), "unit": "walker local-energies/s", "cores": cores, "kind": "port",
            "sample": f"{k} walkers of the same workload per repetition: {n_steps} MCMC steps + E_loc (3n HVPs, "
                      f"vmapped chunks of {chunk}) in FP32 PyTorch CPU; warm-up, then median of {reps} "
                      f"(end-to-end {tm + te:.1f} s per repetition)",
            "reps": reps, "chunk": chunk, "median_s": round(tm + te, 3),
            "mcmc_walker_steps_per_s": round(k * n_steps / tm, 1), "eloc_walkers_per_s": round(k / te, 3),
            "baseline_only": True}


def dpe4_bench(args):
    """The N2 unit (walkers x [20 MCMC steps + E_loc]) with the reference's default dpe4 model: same/diff
    pair streams, el-ion stream and SchNet-style convolutions (configuration.py:875-891), 16 dets, seed 1."""
    import torch

    from deeperwin_amd._native import TIMER_LOCAL_ENERGY, NativeModel
    from deeperwin_amd.mcmc import make_chain
    from deeperwin_amd.params import ModelShape, init_params
    from deeperwin_amd.physics import generate_el_ion_mapping, geometry_n2, init_walkers_gaussian

    R, Z = geometry_n2()
    n, u, D, B, S = 14, 7, 16, args.walkers, args.mcmc_steps
    shape = ModelShape(n_el=n, n_up=u, n_ion=2, n_det=D, embedding="dpe4", z_min=7, n_ion_types=1)
    params = init_params(shape, seed=1)
    r0 = init_walkers_gaussian(np.random.default_rng(1), B, R, generate_el_ion_mapping(R, Z, n, u))
    dev = "cuda"
    model = NativeModel(shape, args.precision)
    model.set_params(params)
    r = torch.as_tensor(r0, dtype=torch.float32, device=dev).contiguous()
    Rd = torch.as_tensor(R, dtype=torch.float32, device=dev)
    Zd = torch.as_tensor(Z, dtype=torch.float32, device=dev)
    lp = torch.full((B,), -1000.0, device=dev)
    age = torch.zeros(B, dtype=torch.int32, device=dev)
    chain = make_chain(0.02)
    e = torch.empty(B, device=dev)
    gen = torch.Generator(device=dev)
    gen.manual_seed(1)
    events = []

    def step(record=False):
        normals = torch.randn((S, B, n, 3), generator=gen, device=dev)
        uniforms = torch.rand((S, B), generator=gen, device=dev)
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        model.mcmc_steps(r, Rd, Zd, lp, age, normals, uniforms, chain, 100, 100, 0.5, 1e-2, 1.0)
        b.record()
        if record:
            events.append((a, b))
        model.local_energy(r, Rd, Zd, e)

    for _ in range(max(args.warmup, 1)):
        step()
    torch.cuda.synchronize()
    model.set_timing(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(record=True)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    lap_ms, lap_n = model.timer(TIMER_LOCAL_ENERGY)
    model.set_timing(False)
    mcmc_ms = float(np.mean([a.elapsed_time(b) for a, b in events]))

    def eloc_with(m):
        out_e = torch.empty(B, device=dev)
        m.local_energy(r, Rd, Zd, out_e)
        torch.cuda.synchronize()
        return out_e.double().cpu().numpy()

    m64 = NativeModel(shape, "fp64", chunk=1024)
    m64.set_params(params)
    e64, ex = eloc_with(m64), eloc_with(model)
    ok = np.isfinite(e64) & np.isfinite(ex)
    out = {
        "metric": METRIC, "value": round(B / dt, 1), "unit": "walker local-energies/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": args.precision, "data": "synthetic",
        "config": {"workload": f"N2 with the dpe4 model (same/diff pair streams, el-ion stream, SchNet-style "
                               f"convolutions, 4 iterations 256/32), 16 full dets, seed 1; {B} walkers; {S} MCMC "
                               f"steps (stepsize 0.02) + E_loc over {3 * n} directions",
                   "walkers": B, "mcmc_steps": S, "precision": args.precision},
        "breakdown": {"mcmc_step_ms": round(mcmc_ms / S, 4),
                      "local_energy_pass_ms": round(lap_ms / max(lap_n, 1), 4)},
        "vs_fp64_form": {"dE_mean_mHa": round(float(np.mean(ex[ok]) - np.mean(e64[ok])) * 1e3, 4),
                         "median_abs_dE": float(np.median(np.abs(ex[ok] - e64[ok]))),
                         "nonfinite": int((~ok).sum())},
    }
    if not args.no_accuracy:
        from oracle import hamiltonian as oh

        idx = np.linspace(0, B - 1, 8).astype(int)
        r_fin = r.cpu().numpy().astype(np.float64)
        e_ref = oh.local_energy(params, u, n - u, r_fin[idx], R, Z).numpy()
        out["accuracy"] = {"sample_walkers": 8, "vs": "FP64 CPU oracle on the same walkers",
                           "max_abs_dE_fp32": float(np.max(np.abs(ex[idx] - e_ref))),
                           "max_abs_dE_fp64_form": float(np.max(np.abs(e64[idx] - e_ref)))}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(params, u, n - u, r0, R, Z, S, max(8, args.cpu_walkers // 4))
    print(json.dumps(out))


def main():
    args = parse()
    if args.workload != "n2" and args.gpus != 1:
        print("bench.py measures one GPU", file=sys.stderr)
        sys.exit(2)
    if args.workload == "c4h2-precision":
        precision_study(args)
        return
    if args.workload == "serving":
        serving_bench(args)
        return
    if args.workload == "batch64":
        batch_bench(args)
        return
    if args.workload == "copy":
        copy_calibration(args)
        return
    if args.workload == "vendor-gemm":
        vendor_gemm_calibration(args)
        return
    if args.workload in ("grad", "grad-dpe4"):
        grad_bench(args)
        return
    if args.workload == "dpe4":
        dpe4_bench(args)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (walkers shard across GPUs "
              "without any per-step exchange beyond one scalar)", file=sys.stderr)
        sys.exit(2)
    import torch

    from deeperwin_amd._native import TIMER_GEMM_1EL_LAP, TIMER_LOCAL_ENERGY, NativeModel
    from deeperwin_amd.mcmc import make_chain
    from deeperwin_amd.params import ModelShape, init_params
    from deeperwin_amd.physics import generate_el_ion_mapping, geometry_n2, init_walkers_gaussian

    if not torch.cuda.is_available():
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)
    R, Z = geometry_n2()
    n, u, D, B, S = 14, 7, 16, args.walkers, args.mcmc_steps
    shape = ModelShape(n_el=n, n_up=u, n_ion=2, n_det=D)
    params = init_params(shape, seed=1)
    mapping = generate_el_ion_mapping(R, Z, n, u)
    r0 = init_walkers_gaussian(np.random.default_rng(1), B, R, mapping)

    dev = "cuda"
    model = NativeModel(shape, args.precision)
    model.set_params(params)
    r = torch.as_tensor(r0, dtype=torch.float32, device=dev).contiguous()
    Rd = torch.as_tensor(R, dtype=torch.float32, device=dev)
    Zd = torch.as_tensor(Z, dtype=torch.float32, device=dev)
    lp = torch.full((B,), -1000.0, device=dev)
    age = torch.zeros(B, dtype=torch.int32, device=dev)
    chain = make_chain(0.02)
    e = torch.empty(B, device=dev)
    flags = torch.empty(B, dtype=torch.int32, device=dev)
    acc = torch.empty(S, device=dev)
    gen = torch.Generator(device=dev)
    gen.manual_seed(1)

    tuning = None
    if args.tune:  # launch shapes from the autotuner, on the walkers of this run
        from deeperwin_amd.tuning import autotune_model

        tuning, _ = autotune_model(model, r, Rd, Zd, reps=3)

    # proposal normals and accept uniforms live in fixed buffers, refilled in place before every unit, so
    # the S MCMC steps (one dw_mcmc_steps call, ~18 launches per step) can be replayed as one HIP graph
    normals = torch.empty((S, B, n, 3), device=dev)
    uniforms = torch.empty((S, B), device=dev)

    def mcmc():
        model.mcmc_steps(r, Rd, Zd, lp, age, normals, uniforms, chain, 100, 100, 0.5, 1e-2, 1.0, out_acc=acc)

    graph = None
    mcmc_events = []

    def step(record=False):
        normals.normal_(generator=gen)
        uniforms.uniform_(generator=gen)
        if record:
            a = torch.cuda.Event(enable_timing=True)
            a.record()
        if graph is not None:
            graph.replay()
        else:
            mcmc()
        if record:
            b = torch.cuda.Event(enable_timing=True)
            b.record()
            mcmc_events.append((a, b))
        model.local_energy(r, Rd, Zd, e, None, flags)

    # graph mode needs one eager unit first (it sizes the workspace that the graph then reuses)
    for _ in range(args.warmup if args.eager else max(args.warmup, 1)):
        step()
    torch.cuda.synchronize()
    if not args.eager:
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            mcmc()
        torch.cuda.synchronize()
    model.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(record=True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    dt = (t1 - t0) / args.steps
    value = B / dt

    gemm_ms, gemm_n = model.timer(TIMER_GEMM_1EL_LAP)
    lap_ms, lap_n = model.timer(TIMER_LOCAL_ENERGY)
    model.set_timing(False)
    mcmc_ms = float(np.mean([a.elapsed_time(b) for a, b in mcmc_events]))
    C = 3 * n + 2
    M = B * n * C
    K = shape.width1 + 2 * shape.width2
    flops = 2.0 * M * K * shape.width1
    avg_s = (gemm_ms / max(gemm_n, 1)) / 1e3
    achieved = flops / avg_s / 1e12
    peak = PEAK_TFLOPS[args.precision]
    roofline = {"bound": "mfma", "achieved": round(achieved, 2), "peak": peak, "unit": "TFLOP/s",
                "frac": round(achieved / peak, 4), "traffic": traffic_from_profiles(),
                "kernel": "1-el embedding GEMM, Laplacian mode (k_gemm_lap)",
                "flops_per_launch": flops, "avg_launch_ms": round(avg_s * 1e3, 4), "launches": gemm_n}

    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "walker local-energies/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": args.precision, "data": "synthetic",
        "config": {"workload": "N2 (Z=[7,7], 2.07 bohr), 14 electrons (7 up), FermiNet 4 iterations 256/32 widths, "
                               f"16 full dets, random init seed 1; {B} walkers; {S} MCMC steps (stepsize 0.02) + "
                               f"E_loc with the full Laplacian over {3 * n} directions per step",
                   "walkers": B, "mcmc_steps": S, "n_el": n, "n_det": D, "precision": args.precision,
                   "parallelism": "single GPU (walker-parallel)", "launch_shapes": tuning or "heuristic",
                   "mcmc_launch": "eager" if args.eager else f"one HIP graph replay for the {S} MCMC steps"},
        "roofline": roofline,
        "whole_path": whole_path_fractions(shape, S, B, dt, peak),
        "breakdown": {"mcmc_step_ms": round(mcmc_ms / S, 4),
                      "local_energy_pass_ms": round(lap_ms / max(lap_n, 1), 4),
                      "mcmc_walker_steps_per_s": round(B * S / (mcmc_ms / 1e3), 1),
                      "eloc_only_walkers_per_s": round(B / (lap_ms / max(lap_n, 1) / 1e3), 1)},
    }

    e_dev = e.double().cpu().numpy()
    r_fin = r.cpu().numpy().astype(np.float64)
    out["energy_mean"] = float(np.nanmean(e_dev))

    # accuracy over the whole batch against the FP64 form (pinned to the FP64 CPU oracle by the parity
    # tests), and the precision form the metric is quoted at: the fastest whose |dE_mean| <= 0.1 mHa
    def eloc_with(m):
        out_e = torch.empty(B, device=dev)
        m.local_energy(r, Rd, Zd, out_e)
        torch.cuda.synchronize()
        return out_e.double().cpu().numpy()

    m64 = NativeModel(shape, "fp64", chunk=1024)
    m64.set_params(params)
    e64_all = eloc_with(m64)

    def dE(ex):
        ok = np.isfinite(ex) & np.isfinite(e64_all)
        d = ex[ok] - e64_all[ok]
        return {"dE_mean_mHa": round(float(np.mean(ex[ok]) - np.mean(e64_all[ok])) * 1e3, 4),
                "median_abs_dE": float(np.median(np.abs(d))), "max_abs_dE": float(np.max(np.abs(d))),
                "nonfinite": int((~ok).sum())}

    forms = {args.precision: {"walker_eloc_per_s": round(value, 1), "ms_per_step": round(dt * 1e3, 3),
                              **dE(eloc_with(model))}}
    if abs(forms[args.precision]["dE_mean_mHa"]) > 0.1 and args.precision == "fp32":
        def timed_with(m_e, k):
            def st():
                normals = torch.randn((S, B, n, 3), generator=gen, device=dev)
                uniforms = torch.rand((S, B), generator=gen, device=dev)
                model.mcmc_steps(r, Rd, Zd, lp, age, normals, uniforms, chain, 100, 100, 0.5, 1e-2, 1.0, out_acc=acc)
                m_e.local_energy(r, Rd, Zd, e)
            st()
            torch.cuda.synchronize()
            a = time.perf_counter()
            for _ in range(k):
                st()
            torch.cuda.synchronize()
            return (time.perf_counter() - a) / k

        for name, m_e in (("fp32+fp64>1e6", NativeModel(shape, "fp32", fp64_cond_threshold=1e6)),
                          ("fp64 E_loc", m64)):
            if m_e is not m64:
                m_e.set_params(params)
            t = timed_with(m_e, max(1, args.steps // 2))
            e64_all = eloc_with(m64)  # reference on the walkers the chain has reached now
            forms[name] = {"walker_eloc_per_s": round(B / t, 1), "ms_per_step": round(t * 1e3, 3),
                           **dE(eloc_with(m_e))}
    passing = [k for k, v in forms.items() if abs(v["dE_mean_mHa"]) <= 0.1]
    best = max(passing, key=lambda k: forms[k]["walker_eloc_per_s"]) if passing else None
    out["forms"] = forms
    out["form_at_metric"] = best
    if best is not None and best != args.precision:
        out["value"] = forms[best]["walker_eloc_per_s"]
        out["ms_per_step"] = forms[best]["ms_per_step"]
        out["dtype"] = best

    if not args.no_accuracy:
        import torch as _t

        from oracle import hamiltonian as oh

        k = 128  # the mean of a 16-walker sample has a standard error of ~0.2 mHa here
        idx = np.linspace(0, B - 1, k).astype(int)
        e64 = oh.local_energy(params, u, n - u, r_fin[idx], R, Z, vmapped=True).numpy()
        de = e_dev[idx] - e64
        out["accuracy"] = {"sample_walkers": k, "vs": "FP64 CPU oracle on the same walkers",
                           "median_abs_dE": float(np.median(np.abs(de))), "max_abs_dE": float(np.max(np.abs(de))),
                           "dE_mean_mHa": float(np.mean(de) * 1e3)}

    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(params, u, n - u, r0, R, Z, S, args.cpu_walkers)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
