"""Drop-in for the reference's wavefunction construction and call (model/wavefunction.py:262-299).

``build_log_psi_squared`` returns ``(log_psi_sqr, get_slater_mat, get_cache, params, fixed_params)`` like
the reference.  ``log_psi_sqr(params, n_up, n_dn, r, R, Z, fixed_params) -> (phase, log_psi_sqr)`` keeps
the reference signature; it evaluates on the GPU through ``libdwmi.so`` (no CPU fallback).  ``r`` may
have any leading batch dimensions; numpy inputs are moved to the current GPU, torch CUDA tensors are
used in place.
"""
from __future__ import annotations

from typing import Dict, Optional, Tuple

import numpy as np

from ._native import NativeModel
from .config import ModelConfigFermiNet, PhysicalConfig
from .params import ModelShape, infer_shape, init_params


def _as_device(x, dtype=None):
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("the dwmi wavefunction runs on a HIP device; none is visible")
    if isinstance(x, torch.Tensor):
        t = x
    else:
        t = torch.as_tensor(np.asarray(x))
    return t.to(device="cuda", dtype=dtype or torch.float32).contiguous()


class NativeLogPsiSquared:
    """Callable with the reference signature; holds one native model per (shape, params) it has seen."""

    def __init__(self, precision: str = "fp32", fp64_cond_threshold: float = 0.0, chunk: int = 0, z_min: int = 1):
        self.precision = precision
        self.fp64_cond_threshold = fp64_cond_threshold
        self.chunk = chunk
        self.z_min = z_min  # dpe4: Z of the ion embedding table's first row
        self._models: Dict[ModelShape, NativeModel] = {}

    def model_for(self, params: Dict, n_up: int, n_dn: int) -> NativeModel:
        shape = infer_shape(params, n_up + n_dn, n_up, self.z_min)
        m = self._models.get(shape)
        if m is None:
            m = NativeModel(shape, self.precision, self.fp64_cond_threshold, self.chunk)
            self._models[shape] = m
        m.set_params(params)
        return m

    def __call__(self, params: Dict, n_up: int, n_dn: int, r, R, Z, fixed_params: Optional[Dict] = None
                 ) -> Tuple["torch.Tensor", "torch.Tensor"]:
        import torch

        model = self.model_for(params, n_up, n_dn)
        rt = _as_device(r)
        lead = rt.shape[:-2]
        rt = rt.reshape(-1, n_up + n_dn, 3)
        Rt, Zt = _as_device(R), _as_device(Z)
        if Rt.shape[0] != model.shape.n_ion:
            raise ValueError(f"R has {Rt.shape[0]} ions, parameters were built for {model.shape.n_ion}")
        lp = torch.empty(rt.shape[0], device="cuda", dtype=torch.float32)
        ph = torch.empty_like(lp)
        model.logpsi(rt, Rt, Zt, lp, ph)
        return ph.reshape(lead), lp.reshape(lead)


def build_log_psi_squared(config: ModelConfigFermiNet, physical_config: PhysicalConfig, baseline_config=None,
                          fixed_params: Optional[Dict] = None, rng_seed: int = 0, precision: str = "fp32",
                          fp64_cond_threshold: float = 0.0):
    """Build the native wavefunction and random-initialise its parameters from ``rng_seed``.

    Mirrors model/wavefunction.py:262-299 for the FermiNet-style model (ModelConfigFermiNet) and the
    reference's default dpe4 model (ModelConfigDeepErwin4); baseline (CASSCF) orbitals, TAOs and analytic
    envelope initialisation are not part of either and are rejected by ``config.validate``.
    """
    config.validate()
    if isinstance(physical_config, (list, tuple)):
        physical_config = physical_config[0]
    if baseline_config is not None:
        raise NotImplementedError("baseline orbitals are not used by the FermiNet-style model")
    phys = physical_config
    emb = config.embedding
    z_min, types = 1, 0
    if emb.name == "dpe4":  # WavefunctionDefinition (model/wavefunction.py:306-326)
        z_min = int(min(phys.Z)) if getattr(config, "Z_min", None) is None else int(config.Z_min)
        z_max = int(max(phys.Z)) if getattr(config, "Z_max", None) is None else int(config.Z_max)
        types = z_max - z_min + 1
    shape = ModelShape(n_el=phys.n_electrons, n_up=phys.n_up, n_ion=phys.n_ions,
                       n_det=config.orbitals.n_determinants, width1=emb.n_hidden_one_el,
                       width2=emb.n_hidden_two_el, n_iter=emb.n_iterations, embedding=emb.name, z_min=z_min,
                       n_ion_types=types)
    fixed_params = fixed_params or dict(input={}, baseline_energies=dict(E_ref=phys.E_ref))
    params = init_params(shape, rng_seed)
    log_psi_sqr = NativeLogPsiSquared(precision=precision, fp64_cond_threshold=fp64_cond_threshold, z_min=z_min)

    def get_slater_mat(*args, **kwargs):
        raise NotImplementedError("Slater matrices are internal to the fused native path")

    def get_cache(*args, **kwargs):
        return {}  # no geometry-only sub-networks in the FermiNet-style model (wavefunction.py:165-211)

    return log_psi_sqr, get_slater_mat, get_cache, params, fixed_params
