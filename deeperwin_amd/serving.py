"""Request server for energy evaluations with continuous batching on one GPU.

A request is a molecule (R, Z), a walker count and a number of MCMC steps; its answer is the mean local
energy after the chain (the evaluation unit of the reference: MCMC inter-steps then E_loc,
evaluation.py:143-151, 187-198).  Weights are shared: one max-size parameter set stays resident in HBM
and every request gets the per-geometry slice of it (``params.truncate_params``), so requests of any
electron count (2..n_el_max) and ion count run under the same weights.

Scheduling (continuous batching at MCMC-slice granularity):
* requests are admitted in arrival order while the sum of their device footprints (walker state plus
  the model workspace, ``dw_workspace_size``) stays under the HBM budget and a slot is free;
* every admitted request owns a slot: a HIP stream, the slot's grow-only workspace arena
  (``dw_model_set_workspace``: no device allocation per request once the slot has seen a request that
  large), a native model instance from the shape pool and its walker state on the device; the host loop enqueues the next slice of
  MCMC steps of every active request round-robin, so slices of different requests overlap on the device;
* after the last slice the request's local energies and their FP64 mean are enqueued on its stream;
  completion is detected with a stream event, the slot and its memory budget are released and the next
  waiting request is admitted.
Each request's random draws come from its own seeded generator and every kernel is deterministic and
walker-independent, so a request's result does not depend on what else is in flight.
"""
from __future__ import annotations

import math
import time
from collections import OrderedDict
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np

from ._native import energy_moments
from .params import ModelShape, init_params, truncate_params
from .physics import generate_el_ion_mapping, init_walkers_gaussian, random_molecule


@dataclass
class Request:
    rid: int
    R: np.ndarray
    Z: np.ndarray
    n_walkers: int
    n_steps: int = 50
    seed: int = 0
    arrival: float = 0.0       # seconds after the stream starts

    @property
    def n_el(self) -> int:
        return int(np.sum(self.Z))

    @property
    def n_up(self) -> int:
        return (self.n_el + 1) // 2


@dataclass
class Result:
    rid: int
    energy_mean: float
    energy_var: float
    acceptance: float
    arrival: float
    admitted: float
    finished: float

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def poisson_stream(n_requests: int, rate: float = 50.0, seed: int = 5, n_el_range=(2, 60),
                   walker_range=(256, 4096), n_steps: int = 50) -> List[Request]:
    """Seeded request stream: Poisson arrivals at `rate`, electron counts uniform in n_el_range, walker
    counts log-uniform in walker_range (rounded to a multiple of 64)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for k in range(n_requests):
        t += rng.exponential(1.0 / rate)
        n = int(rng.integers(n_el_range[0], n_el_range[1] + 1))
        R, Z = random_molecule(rng, n)
        B = int(np.exp(rng.uniform(np.log(walker_range[0]), np.log(walker_range[1]))))
        B = max(64, (B + 32) // 64 * 64)
        out.append(Request(rid=k, R=R, Z=Z, n_walkers=B, n_steps=n_steps, seed=1000 + k, arrival=t))
    return out


class SharedWeights:
    """One max-size parameter set; every request runs a native model built from its truncated slice.

    The per-geometry slice depends only on the shape (n_el, n_up, n_ion), so models are kept in a pool per
    shape and handed out again: their weights stay resident in HBM, and a request whose shape has been
    seen before costs no model creation, weight upload or device allocation.  Idle models beyond
    `max_models` are destroyed, oldest first."""

    def __init__(self, n_el_max: int = 60, n_ion_max: int = 60, n_det: int = 8, seed: int = 0,
                 precision: str = "fp32", chunk: int = 128, max_models: int = 64):
        self.n_el_max, self.n_ion_max, self.n_det = n_el_max, n_ion_max, n_det
        self.big = init_params(ModelShape(n_el_max, (n_el_max + 1) // 2, n_ion_max, n_det), seed)
        self.precision = precision
        self.chunk = chunk
        self.max_models = max_models
        self._idle: "OrderedDict[tuple, List[object]]" = OrderedDict()
        self._n_idle = 0

    def shape_for(self, req: Request) -> ModelShape:
        return ModelShape(req.n_el, req.n_up, len(req.Z), self.n_det)

    def params_for(self, req: Request):
        return truncate_params(self.big, self.n_ion_max, self.n_el_max, self.shape_for(req))

    def acquire(self, req: Request):
        from ._native import NativeModel

        key = (req.n_el, req.n_up, len(req.Z))
        if self._idle.get(key):
            self._n_idle -= 1
            m = self._idle[key].pop()
            if not self._idle[key]:
                del self._idle[key]
            return m
        m = NativeModel(self.shape_for(req), self.precision, chunk=self.chunk)
        m.set_params(self.params_for(req))
        m.pool_key = key
        return m

    def release(self, model) -> None:
        """Return a model to the pool (its caller must have finished using it)."""
        self._idle.setdefault(model.pool_key, []).append(model)
        self._idle.move_to_end(model.pool_key)
        self._n_idle += 1
        while self._n_idle > self.max_models:
            key, lst = next(iter(self._idle.items()))
            lst.pop(0)  # destroyed with its last reference
            self._n_idle -= 1
            if not lst:
                del self._idle[key]

    def arena(self, slot: int, nbytes: int):
        """Workspace buffer of a request slot: grow-only and kept across requests and servers, so device
        memory is allocated only when a slot first meets a larger request."""
        import torch

        if not hasattr(self, "_arenas"):
            self._arenas = {}
        buf = self._arenas.get(slot)
        if buf is None or buf.numel() < nbytes:
            self._arenas[slot] = None
            del buf
            buf = torch.empty(max(nbytes, 1 << 20), dtype=torch.uint8, device="cuda")
            self._arenas[slot] = buf
        return buf[:nbytes]

    def workspace_bytes(self, req: Request, model) -> int:
        """The larger of the value-pass and Laplacian-pass workspaces (the latter per chunk of walkers)."""
        return int(max(model.workspace_bytes(req.n_walkers, 0), model.workspace_bytes(req.n_walkers, 1)))

    def footprint(self, req: Request, model) -> int:
        """Device bytes a request holds while active: its workspace, walker state and the draws of a slice."""
        state = req.n_walkers * (req.n_el * 3 * 4 * 3 + 16)
        return int(self.workspace_bytes(req, model) + state)


class _Active:
    def __init__(self, req: Request, model, stream, admitted: float, slice_steps: int):
        import torch

        self.req, self.model, self.stream, self.admitted = req, model, stream, admitted
        self.steps_done = 0
        self.slice_steps = slice_steps
        dev = "cuda"
        mapping = generate_el_ion_mapping(req.R, req.Z, req.n_el, req.n_up)
        r0 = init_walkers_gaussian(np.random.default_rng(req.seed), req.n_walkers, req.R, mapping)
        from .mcmc import make_chain

        with torch.cuda.stream(stream):
            self.r = torch.as_tensor(r0, dtype=torch.float32).to(dev, non_blocking=False).contiguous()
            self.R = torch.as_tensor(req.R, dtype=torch.float32).to(dev)
            self.Z = torch.as_tensor(req.Z, dtype=torch.float32).to(dev)
            self.lp = torch.empty(req.n_walkers, device=dev)
            model.logpsi(self.r, self.R, self.Z, self.lp)
            self.age = torch.zeros(req.n_walkers, dtype=torch.int32, device=dev)
            self.chain = make_chain(0.02)
            self.acc = torch.zeros(max(1, req.n_steps), device=dev)
            self.e = torch.empty(req.n_walkers, device=dev)
            self.gen = torch.Generator(device=dev)
            self.gen.manual_seed(req.seed)
        self.done_event = None
        self.moments = None

    def enqueue_slice(self) -> None:
        import torch

        k = min(self.slice_steps, self.req.n_steps - self.steps_done)
        B, n = self.req.n_walkers, self.req.n_el
        with torch.cuda.stream(self.stream):
            if k > 0:
                # one generator call per step and kind, in step order: the draws do not depend on how
                # the chain is cut into slices
                normals = torch.empty((k, B, n, 3), device="cuda")
                uniforms = torch.empty((k, B), device="cuda")
                for t in range(k):
                    normals[t].normal_(generator=self.gen)
                    uniforms[t].uniform_(generator=self.gen)
                self.model.mcmc_steps(self.r, self.R, self.Z, self.lp, self.age, normals, uniforms, self.chain,
                                      100, 100, 0.5, 1e-2, 1.0,
                                      out_acc=self.acc[self.steps_done:self.steps_done + k])
                self.steps_done += k
            if self.steps_done >= self.req.n_steps:
                self.model.local_energy(self.r, self.R, self.Z, self.e)
                # nanmean semantics of get_observables (evaluation.py:143-151), FP64 on the device
                mv = energy_moments(self.e)
                self.moments = torch.cat([mv, self.acc[:max(1, self.req.n_steps)].double().mean().reshape(1)])
                # result to pinned host memory on the request's own stream: reading it after the event
                # needs no further synchronisation
                self.host = torch.empty(3, dtype=torch.float64, pin_memory=True)
                self.host.copy_(self.moments, non_blocking=True)
                self.done_event = torch.cuda.Event()
                self.done_event.record(self.stream)

    @property
    def finished_enqueueing(self) -> bool:
        return self.done_event is not None


class Server:
    """Continuous-batching evaluation server on the current GPU."""

    def __init__(self, weights: SharedWeights, hbm_budget_bytes: int = 160 << 30, max_active: int = 8,
                 slice_steps: int = 10):
        self.w = weights
        self.budget = hbm_budget_bytes
        self.max_active = max_active
        self.slice_steps = slice_steps

    def serve(self, requests: List[Request], realtime: bool = True) -> List[Result]:
        import torch

        pending = sorted(requests, key=lambda q: q.arrival)
        streams = [torch.cuda.Stream() for _ in range(self.max_active)]
        free_streams = list(range(self.max_active))
        active: List[Tuple[_Active, int, int]] = []   # (state, stream index, bytes)
        used = 0
        results: List[Result] = []
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0
        while pending or active:
            # admit in arrival order
            while pending and free_streams and (not realtime or pending[0].arrival <= now()):
                req = pending[0]
                model = self.w.acquire(req)
                need = self.w.footprint(req, model)
                if need > self.budget:
                    raise ValueError(f"request {req.rid} needs {need} bytes, more than the whole budget")
                if used + need > self.budget:
                    self.w.release(model)
                    break
                pending.pop(0)
                si = free_streams.pop(0)
                # the slot's workspace arena: no device allocation once the slot has seen a request this big
                model.set_workspace(self.w.arena(si, self.w.workspace_bytes(req, model)))
                active.append((_Active(req, model, streams[si], now(), self.slice_steps), si, need))
                used += need
            # one slice for every active request that still has steps to enqueue
            for st, _, _ in active:
                if not st.finished_enqueueing:
                    st.enqueue_slice()
            # retire finished requests
            still = []
            for st, si, need in active:
                if st.done_event is not None and st.done_event.query():
                    m = st.host.numpy()
                    results.append(Result(st.req.rid, float(m[0]), float(m[1]), float(m[2]), st.req.arrival,
                                          st.admitted, now()))
                    st.model.set_workspace(None)  # the request's work is done: its workspace returns to the cache
                    self.w.release(st.model)
                    st.model = None
                    free_streams.append(si)
                    used -= need
                else:
                    still.append((st, si, need))
            active = still
            if not active and pending and realtime and pending[0].arrival > now():
                time.sleep(min(0.002, pending[0].arrival - now()))
            elif active and all(st.finished_enqueueing for st, _, _ in active):
                time.sleep(0.0002)
        torch.cuda.synchronize()
        return sorted(results, key=lambda x: x.rid)


def summarize(results: List[Result], isolated: Optional[Dict[int, float]] = None) -> dict:
    lat = np.array([r.latency for r in results])
    out = {"requests": len(results), "p50_latency_s": float(np.percentile(lat, 50)),
           "p99_latency_s": float(np.percentile(lat, 99)), "makespan_s": float(max(r.finished for r in results))}
    if isolated:
        iso = np.array([isolated[r.rid] for r in results])
        out["isolated_p50_s"] = float(np.percentile(iso, 50))
        out["isolated_p99_s"] = float(np.percentile(iso, 99))
        out["p99_over_isolated_p99"] = out["p99_latency_s"] / out["isolated_p99_s"]
    return out
