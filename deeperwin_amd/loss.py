"""Drop-in for the reference's loss and its parameter gradient (optimization/loss_function.py:12-158).

``build_value_and_grad_func(log_psi_sqr_func, get_local_energy, clipping_config)`` returns
``value_and_grad(params, state, spin_state, batch) -> ((loss, (state, aux)), grads)``, the shape of
``jax.value_and_grad(total_energy, has_aux=True)`` in the reference: E_loc on the device (dw_local_energy),
the clipped statistics and per-walker weights c_b = (E_clip_b - mean E_clip) / B (dw_clip_energies), and
G = sum_b c_b d log_psi_sqr(r_b) / d theta by a reverse-mode pass through the network (dw_param_grad).
``state`` is the clipping state (center, width); ``init_clipping_state()`` gives the reference's initial
(0, 1e12).  Gradients come back as a dict shaped like ``params`` (CUDA float32 tensors).
"""
from __future__ import annotations

from typing import Dict

import torch

from ._native import clip_energies
from .config import ClippingConfig
from .wavefunction import NativeLogPsiSquared, _as_device


def init_clipping_state(is_complex: bool = False):
    """(center, width) = (0, 1e12) (loss_function.py:12-16)."""
    if is_complex:
        raise NotImplementedError("complex wavefunctions are not on the native path")
    return torch.tensor([0.0, 1e12], dtype=torch.float64, device="cuda")


def build_value_and_grad_func(log_psi_sqr_func, get_local_energy, clipping_config: ClippingConfig,
                              is_complex: bool = False, kfac_register_complex: bool = False):
    if not isinstance(log_psi_sqr_func, NativeLogPsiSquared):
        raise TypeError("build_value_and_grad_func expects the native log_psi_squared from build_log_psi_squared")
    if is_complex:
        raise NotImplementedError("complex wavefunctions are not on the native path")
    clipping_config.validate()
    cc = clipping_config

    def value_and_grad(params: Dict, state, spin_state, batch):
        r, R, Z, fixed_params = batch
        n_up, n_dn = spin_state
        model = log_psi_sqr_func.model_for(params, n_up, n_dn)
        rt = _as_device(r).reshape(-1, n_up + n_dn, 3)
        Rt, Zt = _as_device(R), _as_device(Z)
        e_loc = get_local_energy(params, spin_state, rt, Rt, Zt, fixed_params)
        st = None if state is None else torch.as_tensor(state, dtype=torch.float64, device="cuda").reshape(2)
        weights, e_clip, stats = clip_energies(e_loc, st, cc.name, cc.center, cc.width_metric, cc.clip_by,
                                               cc.from_previous_step)
        g = model.param_grad(rt, Rt, Zt, weights)
        grads = model.unpack_grad(g, params)
        aux = dict(E_mean=stats[1], E_var=stats[2], E_mean_clipped=stats[0], E_var_clipped=stats[3],
                   E_loc_clipped=e_clip, E_loc=e_loc)
        return (stats[0], (stats[4:6].clone(), aux)), grads

    return value_and_grad
