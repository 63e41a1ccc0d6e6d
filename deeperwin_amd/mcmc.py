"""Drop-in for the reference's MCMC sampler (mcmc.py:20-151, 314-412).

``MetropolisHastingsMonteCarlo(mcmc_config).run_mcmc_steps(func, state, params, n_up, n_dn,
fixed_params, n_steps)`` keeps the reference signature and returns a new ``MCMCState``.  The whole chain
of ``n_steps`` steps runs on the device without host synchronisation: proposals (the config's rule:
'normal', 'cauchy', 'normal_one_el', 'local', 'local_one_el' or 'langevin', mcmc.py:175-284), log|psi|^2 of
the proposals, accept/reject, walker ages, the running acceptance rate and the stepsize schedule
(stepsize, acc_rate and step_nr live in a 16-byte device record).

Random draws: the reference derives proposals and accept thresholds from per-walker threefry keys
(mcmc.py:178, 360-361).  When ``state.rng_state`` holds such keys (a CUDA int32 tensor [B, 2] of uint32
words, as ``initialize_around_nuclei`` makes from a PRNG key), the draws are the reference's own streams,
generated on the device by ``deeperwin_amd.threefry`` (jax 0.4.23's threefry2x32; the proposal normal and
the accept uniform of a step share one subkey, as in the reference).  A ``torch.Generator`` in
``rng_state`` draws instead from that generator; ``draws=(normals[S, B, n, 3], uniforms[S, B])`` supplies
them explicitly so that a chain can be replayed exactly by the CPU oracle.
"""
from __future__ import annotations

import copy
import struct
from dataclasses import dataclass, field
from typing import Optional, Tuple

import numpy as np
import torch

from .config import MCMCConfig, PhysicalConfig
from .physics import init_walkers_gaussian
from .wavefunction import NativeLogPsiSquared, _as_device


def make_chain(stepsize: float, acc_rate: float = 0.0, step_nr: int = 0) -> torch.Tensor:
    raw = struct.pack("<ffii", float(stepsize), float(acc_rate), int(step_nr), 0)
    return torch.frombuffer(bytearray(raw), dtype=torch.int32).clone().to("cuda")


def read_chain(chain: torch.Tensor) -> Tuple[float, float, int]:
    raw = chain.detach().cpu().numpy().astype(np.int32).tobytes()
    s, a, n, _ = struct.unpack("<ffii", raw)
    return s, a, n


@dataclass
class MCMCState:
    """Walker state (mcmc.py:20-37): r [B, n, 3], R [n_ion, 3], Z [n_ion], log_psi_sqr [B], walker_age [B]."""
    r: torch.Tensor
    R: torch.Tensor
    Z: torch.Tensor
    log_psi_sqr: Optional[torch.Tensor] = None
    walker_age: Optional[torch.Tensor] = None
    rng_state: Optional[torch.Generator] = None
    chain: Optional[torch.Tensor] = None
    last_acceptance: Optional[torch.Tensor] = field(default=None, repr=False)

    @property
    def stepsize(self) -> float:
        return read_chain(self.chain)[0]

    @property
    def acc_rate(self) -> float:
        return read_chain(self.chain)[1]

    @property
    def step_nr(self) -> int:
        return read_chain(self.chain)[2]

    def build_batch(self, fixed_params):
        return self.r, self.R, self.Z, fixed_params

    @classmethod
    def initialize_around_nuclei(cls, n_walkers: int, physical_config: PhysicalConfig, init_method: str = "gaussian",
                                 spin_initialization: str = "el_ion_mapping", rng=0, stepsize: float = 1e-2):
        """Gaussian initialisation around the mapped nuclei (mcmc.py:40-91, branch 64-67); log_psi_sqr
        starts at -1000 so the first move is always accepted; ages start at 0.

        With a PRNG key for `rng` (2 uint32 words: ``threefry.PRNGKey(seed)``, a numpy uint32 pair or a
        CUDA int32 tensor) this follows the reference step by step: rng_r, rng_spin, rng = split(rng, 3);
        r0 = R[mapping] + normal(rng_r, [B, n, 3]); the walker keys are split(rng, B).  With an integer
        seed, r0 comes from numpy's generator and later draws from a seeded torch.Generator."""
        if init_method != "gaussian":
            raise NotImplementedError("only the Gaussian walker initialisation is on the native path")
        if spin_initialization != "el_ion_mapping":
            raise NotImplementedError("only spin_initialization='el_ion_mapping' is on the native path")
        if not isinstance(rng, (int, np.integer)):
            from . import threefry

            rng_r, _rng_spin, rng_k = threefry.split(rng, 3)
            n = physical_config.n_electrons
            offset = torch.as_tensor(np.asarray(physical_config.R, np.float32)[np.asarray(
                physical_config.el_ion_mapping)], device="cuda")
            r0 = (threefry.normal(rng_r, (n_walkers, n, 3)) + offset).contiguous()
            return cls(r=r0, R=_as_device(physical_config.R),
                       Z=_as_device(np.asarray(physical_config.Z, dtype=np.float32)),
                       log_psi_sqr=torch.full((n_walkers,), -1000.0, device="cuda"),
                       walker_age=torch.zeros(n_walkers, dtype=torch.int32, device="cuda"),
                       rng_state=threefry.split(rng_k, n_walkers), chain=make_chain(stepsize))
        seed = int(rng)
        gen = np.random.default_rng(seed)
        r0 = init_walkers_gaussian(gen, n_walkers, physical_config.R, physical_config.el_ion_mapping)
        tg = torch.Generator(device="cuda")
        tg.manual_seed(seed + 1)
        return cls(r=_as_device(r0), R=_as_device(physical_config.R),
                   Z=_as_device(np.asarray(physical_config.Z, dtype=np.float32)),
                   log_psi_sqr=torch.full((n_walkers,), -1000.0, device="cuda"),
                   walker_age=torch.zeros(n_walkers, dtype=torch.int32, device="cuda"),
                   rng_state=tg, chain=make_chain(stepsize))


class MetropolisHastingsMonteCarlo:
    def __init__(self, mcmc_config: MCMCConfig):
        mcmc_config.validate()
        self.config = mcmc_config

    def draw(self, state: MCMCState, n_steps: int):
        """(step draws, uniforms, next rng_state) for n_steps steps of the walkers in `state`: normals
        [S, B, n, 3] ([S, B, 3] for the one-electron proposals, whose key draws normal(subkey, (3,))),
        standard Cauchy variates for the 'cauchy' proposal."""
        from ._native import ONE_EL_PROPOSALS

        B, n, _ = state.r.shape
        name = self.config.proposal.name
        m = 1 if name in ONE_EL_PROPOSALS else n
        g = state.rng_state
        if isinstance(g, torch.Tensor):  # per-walker threefry keys: the reference's own streams
            from . import threefry

            if name == "cauchy":
                raise NotImplementedError("JAX's Cauchy stream is not replayed; use a torch.Generator or draws=")
            normals, uniforms, keys = threefry.mcmc_draws(g, m, n_steps)
            return (normals.reshape(n_steps, B, 3) if m == 1 else normals), uniforms, keys
        shape = (n_steps, B, 3) if m == 1 else (n_steps, B, n, 3)
        if name == "cauchy":
            normals = torch.empty(shape, device="cuda", dtype=torch.float32).cauchy_(generator=g)
        else:
            normals = torch.randn(shape, generator=g, device="cuda", dtype=torch.float32)
        uniforms = torch.rand((n_steps, B), generator=g, device="cuda", dtype=torch.float32)
        return normals, uniforms, g

    def run_mcmc_steps(self, func, state: MCMCState, params, n_up: int, n_dn: int, fixed_params, n_steps: int,
                       draws=None) -> MCMCState:
        if not isinstance(func, NativeLogPsiSquared):
            raise TypeError("run_mcmc_steps expects the native log_psi_squared from build_log_psi_squared")
        model = func.model_for(params, n_up, n_dn)
        new = copy.copy(state)
        new.r = state.r.to(device="cuda", dtype=torch.float32).clone().contiguous()
        B = new.r.shape[0]
        new.log_psi_sqr = (state.log_psi_sqr.clone() if state.log_psi_sqr is not None
                           else torch.full((B,), -1000.0, device="cuda"))
        new.walker_age = (state.walker_age.to(torch.int32).clone() if state.walker_age is not None
                          else torch.zeros(B, dtype=torch.int32, device="cuda"))
        new.chain = state.chain.clone() if state.chain is not None else make_chain(1e-2)
        if n_steps > 0:
            if draws is not None:
                normals, uniforms = draws
            else:
                normals, uniforms, new.rng_state = self.draw(state, n_steps)
            normals = _as_device(normals)
            uniforms = _as_device(uniforms)
            from ._native import ONE_EL_PROPOSALS

            step_shape = (n_steps, B, 3) if self.config.proposal.name in ONE_EL_PROPOSALS else (
                n_steps, B, n_up + n_dn, 3)
            if normals.shape != step_shape or uniforms.shape != (n_steps, B):
                raise ValueError(f"draws must have shapes {list(step_shape)} and [n_steps, B]")
        else:
            normals = uniforms = None
        acc = torch.empty(max(n_steps, 1), device="cuda", dtype=torch.float32)
        c = self.config
        p = c.proposal
        model.set_proposal(p.name, getattr(p, "r_min", 0.1), getattr(p, "r_max", 1.0),
                           getattr(p, "langevin_scale", 1.0))
        model.mcmc_steps(new.r, _as_device(state.R), _as_device(state.Z), new.log_psi_sqr, new.walker_age, normals,
                         uniforms, new.chain, c.max_age, c.stepsize_update_interval, c.target_acceptance_rate,
                         c.min_stepsize_scale, c.max_stepsize_scale, out_acc=acc)
        new.last_acceptance = acc[:n_steps]
        return new

    def run_inter_steps(self, func, state, params, n_up, n_dn, fixed_params):
        return self.run_mcmc_steps(func, state, params, n_up, n_dn, fixed_params, self.config.n_inter_steps)

    def run_burn_in(self, func, state, params, n_up, n_dn, fixed_params):
        return self.run_mcmc_steps(func, state, params, n_up, n_dn, fixed_params, self.config.n_burn_in)
