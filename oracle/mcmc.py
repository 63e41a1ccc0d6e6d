"""CPU restatement of the reference's Metropolis-Hastings chain with supplied draws — TEST
INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates MetropolisHastingsMonteCarlo.make_mcmc_step / _adjust_stepsize / _run_mcmc_steps
(src/deeperwin/mcmc.py:345-406) with the Gaussian proposal (mcmc.py:175-180).  The reference draws the
proposal normals and the accept uniforms from per-walker threefry keys; here they are supplied
(normals [S, B, n, 3], uniforms [S, B]) so the same chain can be replayed on the device.
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


def mcmc_steps(logpsi_fn: Callable[[np.ndarray], np.ndarray], state: ChainState, normals: np.ndarray,
               uniforms: np.ndarray, max_age: int, update_interval: int, target: float = 0.5,
               lo: float = 1e-2, hi: float = 1.0, dtype=np.float32):
    """Run S = len(normals) steps.  ``logpsi_fn(r[B, n, 3]) -> log_psi_sqr[B]``.

    Per step (mcmc.py:345-379): r' = r + stepsize * xi; p = exp(lp' - lp); accept = (p > u) or
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
        r_new = (r + normals[s].astype(dtype) * stepsize).astype(dtype)
        lp_new = logpsi_fn(r_new).astype(dtype)
        with np.errstate(over="ignore"):
            p_accept = np.exp((lp_new - lp).astype(dtype))
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
