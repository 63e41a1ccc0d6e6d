"""Many geometries of uneven electron counts under shared weights, one graph launch per evaluation unit.

Every geometry gets an ordinary per-geometry parameter set cut out of one max-size set
(``params.truncate_params``: layer 0 and the orbital/envelope head are truncated, the 1-el layers 1-3 and
all 2-el layers are shared as they are), its own native model (with the launch shapes of its
electron-count bucket from ``tuning.TuningTable``) and its walker state resident on the device.

One evaluation unit = `n_steps` Metropolis-Hastings steps (mcmc.py:345-406) followed by the local energy
of every walker (hamiltonian.py:272-291), for every geometry.  The unit is recorded once into a HIP graph:
geometries are spread over `n_streams` capture streams (largest first, each to the least-loaded stream),
forked from and joined back into the capture stream, so one ``graph.replay()`` launches the whole batch
and independent geometries run concurrently on the device.  The proposal normals and accept uniforms of
all geometries live in two flat device buffers that are refilled (two launches) before each replay;
the graph reads them through fixed per-geometry views.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np

from .params import ModelShape, init_params, truncate_params
from .physics import generate_el_ion_mapping, init_walkers_gaussian


@dataclass
class _Geom:
    idx: int
    shape: ModelShape
    model: object
    r: object
    R: object
    Z: object
    lp: object
    age: object
    chain: object
    acc: object
    e: object
    normals: object = None
    uniforms: object = None
    stream: int = 0


class GeometryBatch:
    def __init__(self, geoms: Sequence[Tuple[np.ndarray, np.ndarray]], n_det: int = 8, walkers: int = 512,
                 n_steps: int = 20, precision: str = "fp32", seed: int = 0, n_streams: int = 8,
                 table=None, stepsize: float = 0.02, n_el_max: Optional[int] = None,
                 n_ion_max: Optional[int] = None):
        import torch

        from ._native import NativeModel
        from .mcmc import make_chain

        self.n_steps = n_steps
        self.walkers = walkers
        n_els = [int(np.sum(Z)) for _, Z in geoms]
        self.n_el_max = n_el_max or max(n_els)
        self.n_ion_max = n_ion_max or max(len(Z) for _, Z in geoms)
        self.n_det = n_det
        big = init_params(ModelShape(self.n_el_max, (self.n_el_max + 1) // 2, self.n_ion_max, n_det), seed)
        self.geoms: List[_Geom] = []
        dev = "cuda"
        for g, (R, Z) in enumerate(geoms):
            n = int(np.sum(Z))
            shape = ModelShape(n, (n + 1) // 2, len(Z), n_det)
            m = NativeModel(shape, precision)
            m.set_params(truncate_params(big, self.n_ion_max, self.n_el_max, shape))
            if table is not None:
                table.apply(m)
            mapping = generate_el_ion_mapping(R, Z, n, shape.n_up)
            r0 = init_walkers_gaussian(np.random.default_rng(seed * 1000 + g), walkers, R, mapping)
            self.geoms.append(_Geom(
                g, shape, m,
                torch.as_tensor(r0, dtype=torch.float32, device=dev).contiguous(),
                torch.as_tensor(R, dtype=torch.float32, device=dev),
                torch.as_tensor(np.asarray(Z), dtype=torch.float32, device=dev),
                torch.empty(walkers, device=dev), torch.zeros(walkers, dtype=torch.int32, device=dev),
                make_chain(stepsize), torch.zeros(max(1, n_steps), device=dev), torch.empty(walkers, device=dev)))
        # flat draw buffers with per-geometry views [n_steps, B, n, 3] and [n_steps, B]
        sizes = [n_steps * walkers * gm.shape.n_el * 3 for gm in self.geoms]
        self.normals = torch.empty(max(1, sum(sizes)), device=dev)
        self.uniforms = torch.empty(max(1, n_steps * walkers * len(self.geoms)), device=dev)
        off = 0
        for k, gm in enumerate(self.geoms):
            gm.normals = self.normals[off:off + sizes[k]].view(n_steps, walkers, gm.shape.n_el, 3)
            gm.uniforms = self.uniforms[k * n_steps * walkers:(k + 1) * n_steps * walkers].view(n_steps, walkers)
            off += sizes[k]
        # longest-processing-time assignment of geometries to streams (cost ~ n_el^2 per walker)
        self.n_streams = max(1, n_streams)
        load = [0.0] * self.n_streams
        for gm in sorted(self.geoms, key=lambda x: -x.shape.n_el):
            k = int(np.argmin(load))
            gm.stream = k
            load[k] += float(gm.shape.n_el) ** 2
        self.streams = [torch.cuda.Stream() for _ in range(self.n_streams)]
        self.graph = None
        for gm in self.geoms:  # log|psi|^2 of the initial walkers (the chain re-evaluates it anyway)
            gm.model.logpsi(gm.r, gm.R, gm.Z, gm.lp)
        torch.cuda.synchronize()

    # ---- draws ----
    def refresh_draws(self, generator=None) -> None:
        self.normals.normal_(generator=generator)
        self.uniforms.uniform_(generator=generator)

    # ---- one evaluation unit ----
    def _enqueue(self) -> None:
        import torch

        main = torch.cuda.current_stream()
        for s in self.streams:
            s.wait_stream(main)
        for gm in self.geoms:
            with torch.cuda.stream(self.streams[gm.stream]):
                gm.model.mcmc_steps(gm.r, gm.R, gm.Z, gm.lp, gm.age, gm.normals, gm.uniforms, gm.chain,
                                    100, 1000, 0.5, 1e-2, 1.0, out_acc=gm.acc)
                gm.model.local_energy(gm.r, gm.R, gm.Z, gm.e)
        for s in self.streams:
            main.wait_stream(s)

    def run_eager(self) -> None:
        self._enqueue()

    def capture(self) -> None:
        """Record one unit into a graph.  Every workspace must already be sized: run one eager unit first."""
        import torch

        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self._enqueue()
        torch.cuda.synchronize()

    def replay(self) -> None:
        if self.graph is None:
            raise RuntimeError("capture() first")
        self.graph.replay()

    # ---- results ----
    def energies(self) -> List[np.ndarray]:
        return [gm.e.double().cpu().numpy() for gm in self.geoms]

    def state(self):
        """Host copies of (r, log_psi_sqr, age, local energies) per geometry."""
        return [(gm.r.cpu().numpy(), gm.lp.cpu().numpy(), gm.age.cpu().numpy(), gm.e.cpu().numpy())
                for gm in self.geoms]

    def flops_per_unit(self, executed: bool = False) -> float:
        """FLOPs of one unit (n_steps forwards + one Laplacian pass per walker), summed over geometries: the
        scope table's count of the reference's layer shapes, or with `executed` the count of the algorithm
        this library runs (spin means multiplied once per walker)."""
        f = executed_unit_flops if executed else unit_flops
        return float(sum(self.walkers * f(gm.shape, self.n_steps) for gm in self.geoms))


def forward_flops(shape: ModelShape) -> float:
    """2 n sum_l d_in d_out (1-el) + 2 n^2 sum_l e_in e_out (2-el) + 2 n W1 D n (orbitals) + D (2/3) n^3."""
    n, I, D, W1, W2, L = shape.n_el, shape.n_ion, shape.n_det, shape.width1, shape.width2, shape.n_iter
    one = (12 * I + 8) * W1 + (L - 1) * (3 * W1 + 2 * W2) * W1
    two = 4 * W2 + (L - 2) * W2 * W2
    return 2 * n * one + 2 * n * n * two + 2 * n * W1 * D * n + D * (2.0 / 3.0) * n ** 3


def laplacian_flops(shape: ModelShape) -> float:
    """(2 + 3 n) (1-el + orbital GEMM FLOPs) + 8 (2-el FLOPs) + D 2 n^3 + D 3 n (2 n^3 + 4 n^2)."""
    n, I, D, W1, W2, L = shape.n_el, shape.n_ion, shape.n_det, shape.width1, shape.width2, shape.n_iter
    one = 2 * n * ((12 * I + 8) * W1 + (L - 1) * (3 * W1 + 2 * W2) * W1) + 2 * n * W1 * D * n
    two = 2 * n * n * (4 * W2 + (L - 2) * W2 * W2)
    return (2 + 3 * n) * one + 8 * two + D * 2 * n ** 3 + D * 3 * n * (2 * n ** 3 + 4 * n ** 2)


def unit_flops(shape: ModelShape, n_steps: int) -> float:
    return n_steps * forward_flops(shape) + laplacian_flops(shape)


# ---- FLOPs of the algorithm as this library runs it ----
# The counts above follow the reference's layer shapes: every electron multiplies the full symmetric
# feature vector f = [h, mean_up h, mean_dn h, pair_up, pair_dn] (width 3 d + 2 e) by W.  The library
# splits W by rows: the per-electron part [h | pair_up | pair_dn] (width d + 2 e) is one GEMM per electron
# and the spin means [mean_up h | mean_dn h] (width 2 d) are multiplied once per walker, so fewer FLOPs run
# for the same linear map.  The Laplacian pass carries C = 3 n + 2 channels (value, 3 n tangents,
# Laplacian) through the 1-el, spin-mean and orbital GEMMs and 5 channels (value, 3 tangents along
# r_j - r_i, Laplacian) through the 2-el stream.

def _layer_widths(shape: ModelShape):
    I, W1, W2, L = shape.n_ion, shape.width1, shape.width2, shape.n_iter
    for l in range(L):
        kh = 4 * I if l == 0 else W1
        ke = 4 if l == 0 else W2
        yield kh, ke


def executed_forward_flops(shape: ModelShape) -> float:
    """One value pass per walker: 2 n (d + 2 e) W1 per electron and 2 (2 d) W1 per walker per layer, the 2-el
    stream, the orbital GEMM and the LU."""
    n, D, W1, W2, L = shape.n_el, shape.n_det, shape.width1, shape.width2, shape.n_iter
    one = sum(2 * n * (kh + 2 * ke) * W1 + 2 * (2 * kh) * W1 for kh, ke in _layer_widths(shape))
    two = 2 * n * n * (4 * W2 + (L - 2) * W2 * W2)
    return one + two + 2 * n * W1 * D * n + D * (2.0 / 3.0) * n ** 3


def executed_laplacian_flops(shape: ModelShape) -> float:
    """One Laplacian pass per walker: C = 3 n + 2 channels through the 1-el, spin-mean and orbital GEMMs,
    5 channels through the 2-el stream, and the determinant stage of the scope table."""
    n, D, W1, W2, L = shape.n_el, shape.n_det, shape.width1, shape.width2, shape.n_iter
    C = 3 * n + 2
    one = sum(2 * n * C * (kh + 2 * ke) * W1 + 2 * C * (2 * kh) * W1 for kh, ke in _layer_widths(shape))
    two = 5 * 2 * n * n * (4 * W2 + (L - 2) * W2 * W2)
    return one + two + C * 2 * n * W1 * D * n + D * 2 * n ** 3 + D * 3 * n * (2 * n ** 3 + 4 * n ** 2)


def executed_unit_flops(shape: ModelShape, n_steps: int) -> float:
    return n_steps * executed_forward_flops(shape) + executed_laplacian_flops(shape)
