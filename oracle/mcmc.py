"""CPU restatement of the reference's Metropolis-Hastings chain with supplied draws — TEST
INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates MetropolisHastingsMonteCarlo.make_mcmc_step / _adjust_stepsize / _run_mcmc_steps
(src/deeperwin/mcmc.py:345-406) with the Gaussian proposal (mcmc.py:175-180) and the local-stepsize and
approximate-Langevin proposals (mcmc.py:204-236, 255-284).  The reference draws the proposal normals and
the accept uniforms from per-walker threefry keys; here they are supplied (normals [S, B, n, 3], uniforms
[S, B]) so the same chain can be replayed on the device.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Callable

import numpy as np


@dataclass
class ChainState:
    r: np.ndarray            # [B, n, 3]
    log_psi_sqr: np.ndarray  # [B]
    walker_age: np.ndarray   # [B] int
    stepsize: float
    step_nr: int = 0
    acc_rate: float = 0.0


def adjust_stepsize(stepsize: float, acc_rate: float, target: float, lo: float, hi: float, dtype=np.float64) -> float:
    """mcmc.py:381-387: /1.05 below the target acceptance, *1.05 otherwise, then clip.  The arithmetic is in
    `dtype`: in the reference's float32 mode the stepsize is a float32 array and the Python constants are
    weakly typed, so the factor is float32(1.05) and the quotient is rounded to float32."""
    s, f = dtype(stepsize), dtype(1.05)
    s = s / f if acc_rate < target else s * f
    return float(np.clip(s, dtype(lo), dtype(hi)))


def scale_and_drift(r, R, Z, stepsize, r_min, r_max, langevin_scale=1.0, dtype=np.float32):
    """_calculate_stepsize_and_langevin_bias (mcmc.py:208-215) per electron: s = stepsize * clip(distance to
    the closest nucleus, r_min, r_max) [..., n] and g = -scale sum_I (r - R_I) Z_I / |r - R_I| [..., n, 3]."""
    r = np.asarray(r, dtype)
    R = np.asarray(R, dtype)
    Z = np.asarray(Z, dtype)
    diff = r[..., :, None, :] - R                        # get_el_ion_distance_matrix (utils.py:262-273)
    dist = np.sqrt(np.sum(diff * diff, axis=-1)).astype(dtype)
    s = (dtype(stepsize) * np.clip(dist.min(axis=-1), dtype(r_min), dtype(r_max))).astype(dtype)
    g = (-dtype(langevin_scale) * np.sum(diff * Z[:, None] / dist[..., None], axis=-2)).astype(dtype)
    return s, g


def log_q_ratio(kind: str, r, r_new, R, Z, stepsize, r_min=0.1, r_max=1.0, langevin_scale=1.0, dtype=np.float32):
    """log q [B] of the move r -> r_new, as the reference forms it after drawing r_new (mcmc.py:228-235,
    273-283): 'local' sum_i 3 (log s - log s') + |r' - r|^2 (1/s^2 - 1/s'^2) / 2; 'langevin' sum_i
    3 (log s - log s') + (|r' - r - g s^2|^2 / s^2 - |r - r' - g' s'^2|^2 / s'^2) / 2; 'normal' 0."""
    r = np.asarray(r, dtype)
    r_new = np.asarray(r_new, dtype)
    if kind == "normal":
        return np.zeros(r.shape[0], dtype)
    scale = langevin_scale if kind == "langevin" else 0.0
    s, g = scale_and_drift(r, R, Z, stepsize, r_min, r_max, scale, dtype)
    s_new, g_new = scale_and_drift(r_new, R, Z, stepsize, r_min, r_max, scale, dtype)
    lq = dtype(3) * (np.log(s) - np.log(s_new))
    if kind == "local":
        dist_sqr = np.sum((r_new - r) ** 2, axis=-1)
        lq = lq + dtype(0.5) * dist_sqr * (dtype(1) / s ** 2 - dtype(1) / s_new ** 2)
    elif kind == "langevin":
        d_fwd = np.sum((r_new - r - g * s[..., None] ** 2) ** 2, axis=-1)
        d_rev = np.sum((r - r_new - g_new * s_new[..., None] ** 2) ** 2, axis=-1)
        lq = lq + dtype(0.5) * (d_fwd / s ** 2 - d_rev / s_new ** 2)
    else:
        raise ValueError(f"unknown proposal {kind!r}")
    return np.sum(lq, axis=-1).astype(dtype)


def propose(kind: str, r, R, Z, normals, stepsize, r_min=0.1, r_max=1.0, langevin_scale=1.0, dtype=np.float32,
            step_nr: int = 0):
    """One proposal (r' [B, n, 3], log q [B]) of the reference's rules, in the reference's operation order:

    * 'normal'   r' = r + stepsize xi                      (mcmc.py:175-180; 'cauchy' is the same with
                                                            Cauchy draws, 195-200)
    * 'local'    r' = r + s xi                             (mcmc.py:218-236)
    * 'langevin' r' = r + s xi + g s^2                     (mcmc.py:255-284)
    * 'normal_one_el' / 'local_one_el': only electron k = step_nr mod n moves, by xi [B, 3] (times stepsize,
      or its local s); the local form's log q is that electron's term alone (mcmc.py:183-192, 239-252)
    with s, g from scale_and_drift at r, and log q from log_q_ratio."""
    r = np.asarray(r, dtype)
    xi = np.asarray(normals, dtype)
    if kind in ("normal_one_el", "local_one_el"):
        k = step_nr % r.shape[-2]
        r_new = r.copy()
        if kind == "normal_one_el":
            r_new[:, k] = (r[:, k] + xi * dtype(stepsize)).astype(dtype)
            return r_new, np.zeros(r.shape[0], dtype)
        s, _ = scale_and_drift(r[:, k:k + 1], R, Z, stepsize, r_min, r_max, 0.0, dtype)
        r_new[:, k] = (r[:, k] + xi * s).astype(dtype)
        s_new, _ = scale_and_drift(r_new[:, k:k + 1], R, Z, stepsize, r_min, r_max, 0.0, dtype)
        s, s_new = s[:, 0], s_new[:, 0]
        dist_sqr = np.sum((r_new[:, k] - r[:, k]) ** 2, axis=-1)
        lq = dtype(3) * (np.log(s) - np.log(s_new)) + dtype(0.5) * dist_sqr * (dtype(1) / s ** 2 - dtype(1) / s_new ** 2)
        return r_new, lq.astype(dtype)
    if kind == "normal":
        r_new = r + xi * dtype(stepsize)
    elif kind in ("local", "langevin"):
        scale = langevin_scale if kind == "langevin" else 0.0
        s, g = scale_and_drift(r, R, Z, stepsize, r_min, r_max, scale, dtype)
        s = s[..., None]
        r_new = r + xi * s + g * s ** 2 if kind == "langevin" else r + xi * s
    else:
        raise ValueError(f"unknown proposal {kind!r}")
    r_new = r_new.astype(dtype)
    return r_new, log_q_ratio(kind, r, r_new, R, Z, stepsize, r_min, r_max, langevin_scale, dtype)


def mcmc_steps(logpsi_fn: Callable[[np.ndarray], np.ndarray], state: ChainState, normals: np.ndarray,
               uniforms: np.ndarray, max_age: int, update_interval: int, target: float = 0.5,
               lo: float = 1e-2, hi: float = 1.0, dtype=np.float32, proposal=None):
    """Run S = len(normals) steps.  ``logpsi_fn(r[B, n, 3]) -> log_psi_sqr[B]``.

    ``proposal(r, xi, stepsize, step_nr) -> (r', log q)`` (default: the Gaussian rule, log q = 0).
    Per step (mcmc.py:345-379): r', log q from the proposal; p = exp(lp' - lp + log q); accept = (p > u) or
    (old age >= max_age); age <- 0 if accepted else age + 1; r, lp selected; acc = mean(accept);
    step_nr += 1; acc_rate <- 0.9 acc_rate_old + 0.1 acc; when step_nr % interval == 0 the stepsize is
    adjusted from the OLD stepsize and OLD acc_rate.  log_psi_sqr is re-evaluated at entry (mcmc.py:396).
    Returns (state, per-step acceptance fractions, per-step accept masks).
    """
    r = state.r.astype(dtype)
    lp = logpsi_fn(r).astype(dtype)
    age = state.walker_age.astype(np.int64).copy()
    stepsize, step_nr, acc_rate = dtype(state.stepsize), state.step_nr, dtype(state.acc_rate)
    accs, masks = [], []
    for s in range(len(normals)):
        if proposal is None:
            r_new, log_q = (r + normals[s].astype(dtype) * stepsize).astype(dtype), dtype(0)
        else:
            r_new, log_q = proposal(r, normals[s], stepsize, step_nr)
        lp_new = logpsi_fn(r_new).astype(dtype)
        with np.errstate(over="ignore"):
            p_accept = np.exp((lp_new - lp + log_q).astype(dtype))
        accept = (p_accept > uniforms[s].astype(dtype)) | (age >= max_age)
        age = np.where(accept, 0, age + 1)
        lp = np.where(accept, lp_new, lp)
        r = np.where(accept[:, None, None], r_new, r)
        acc = float(np.mean(accept))
        step_nr += 1
        old_rate = acc_rate
        acc_rate = dtype(0.9 * old_rate + 0.1 * acc)
        if step_nr % update_interval == 0:
            stepsize = dtype(adjust_stepsize(float(stepsize), float(old_rate), target, lo, hi, dtype))
        accs.append(acc)
        masks.append(accept)
    out = ChainState(r=r, log_psi_sqr=lp, walker_age=age, stepsize=float(stepsize), step_nr=step_nr,
                     acc_rate=float(acc_rate))
    return out, np.asarray(accs), np.asarray(masks)
