"""CPU restatement of the reference's clipped-energy loss and its parameter gradient — TEST INFRASTRUCTURE
ONLY (see ``oracle/__init__.py``).

* ``clipping_center_and_width``, ``clip_energies``  <- optimization/loss_function.py:12-72
* ``total_energy``                                   <- build_value_and_grad_func's total_energy (75-110)
* ``param_grad``                                     <- the custom JVP (112-156) contracted into a gradient:
  G = sum_b c_b d log_psi_sqr(r_b) / d theta with c_b = (E_clip_b - E_mean_clip) / B, obtained by autograd
  through ``oracle.wavefunction.log_psi_sqr`` in float64.
"""
from __future__ import annotations

from typing import Dict, Tuple

import numpy as np
import torch

from . import wavefunction as wf


def clipping_center_and_width(E: np.ndarray, center: str = "mean", width_metric: str = "std",
                              clip_by: float = 5.0) -> Tuple[float, float]:
    """_get_clipping_center_and_width (loss_function.py:19-31), one device: center = nanmean or nanmedian;
    width = nanmean |E - center| ("mae") or sqrt(nanmean (E - center)^2) ("std"), times clip_by."""
    E = np.asarray(E, dtype=np.float64)
    c = {"mean": np.nanmean, "median": np.nanmedian}[center](E)
    if width_metric == "mae":
        w = np.nanmean(np.abs(E - c))
    elif width_metric == "std":
        w = np.sqrt(np.nanmean((E - c) ** 2))
    else:
        raise NotImplementedError(width_metric)
    return float(c), float(w * clip_by)


def clip(E: np.ndarray, center: float, width: float, name: str = "tanh") -> np.ndarray:
    """_get_clip_func (loss_function.py:43-60): hard = clip to center +- width; tanh = center + tanh((E -
    center) / width) width."""
    E = np.asarray(E, dtype=np.float64)
    if name == "hard":
        return np.clip(E, center - width, center + width)
    if name == "tanh":
        return center + np.tanh((E - center) / width) * width
    raise ValueError(name)


def clip_energies(E, clipping_state, name="tanh", center="mean", width_metric="std", clip_by=5.0,
                  from_previous_step=True):
    """_clip_energies (loss_function.py:63-72): the clipping state (center, width) of the previous step is
    used when from_previous_step, else it is recomputed from E; the new state comes from the clipped E."""
    c, w = clipping_state if clipping_state is not None else (None, None)
    if (not from_previous_step) or c is None:
        c, w = clipping_center_and_width(E, center, width_metric, clip_by)
    Ec = clip(E, c, w, name)
    return Ec, clipping_center_and_width(Ec, center, width_metric, clip_by)


def total_energy(E_loc, clipping_state, **clip_cfg):
    """total_energy's forward (loss_function.py:86-110): loss = nanmean(E_clip) and the aux statistics."""
    E = np.asarray(E_loc, dtype=np.float64)
    E_mean = np.nanmean(E)
    E_var = np.nanmean((E - E_mean) ** 2)
    Ec, state = clip_energies(E, clipping_state, **clip_cfg)
    Ecm = np.nanmean(Ec)
    aux = dict(E_mean=E_mean, E_var=E_var, E_mean_clipped=Ecm, E_var_clipped=np.nanmean((Ec - Ecm) ** 2),
               E_loc_clipped=Ec, E_loc=E)
    return float(Ecm), (state, aux)


def param_grad(params: Dict, n_up: int, n_dn: int, r, R, c, Z=None) -> Dict:
    """G = sum_b c_b d log_psi_sqr(r_b) / d theta in float64, as a dict shaped like `params` (Z: the nuclear
    charges, read by the dpe4 model's ion embedding)."""
    leaves = {(p, k): torch.tensor(np.asarray(v, dtype=np.float64), requires_grad=True)
              for p, lv in params.items() for k, v in lv.items()}
    tparams = {}
    for (p, k), t in leaves.items():
        tparams.setdefault(p, {})[k] = t
    r = torch.as_tensor(np.asarray(r, dtype=np.float64))
    lp = wf.log_psi_sqr(tparams, n_up, n_dn, r, R, Z)[1]
    (lp * torch.as_tensor(np.asarray(c, dtype=np.float64))).sum().backward()
    return {p: {k: leaves[(p, k)].grad.numpy() for k in lv} for p, lv in params.items()}


def loss_gradient(params: Dict, n_up: int, n_dn: int, r, R, E_loc, clipping_state, Z=None, **clip_cfg):
    """value_and_grad of total_energy (loss_function.py:75-158): (loss, (state, aux)), grads, with the
    tangent jnp.dot(d log_psi_sqr, E_clip - E_mean_clip) / B."""
    loss, (state, aux) = total_energy(E_loc, clipping_state, **clip_cfg)
    c = (aux["E_loc_clipped"] - aux["E_mean_clipped"]) / len(aux["E_loc"])
    return (loss, (state, aux)), param_grad(params, n_up, n_dn, r, R, c, Z)
