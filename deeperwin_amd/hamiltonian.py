"""Drop-in for the reference's local-energy builder (hamiltonian.py:272-291).

``build_local_energy(log_psi_squared, ...)`` returns ``get_local_energy(params, spin_state, r, R, Z,
fixed_params) -> E_loc[B]``.  The kinetic term is a forward-mode Laplacian over all 3 n_el directions
computed by the native library (the value of ``forward_lap`` does not change the result, only how the
reference computes it); the potential is the non-periodic Coulomb sum (hamiltonian.py:17-39).
"""
from __future__ import annotations

from typing import Dict, Optional

import torch

from ._native import FLAG_NONFINITE
from .wavefunction import NativeLogPsiSquared, _as_device


def build_local_energy(log_psi_squared, is_complex: bool = False, is_periodic: bool = False,
                       include_heg_background: bool = False, forward_lap: bool = False, max_batch_size: int = 64,
                       native_chunk: int = 0):
    if not isinstance(log_psi_squared, NativeLogPsiSquared):
        raise TypeError("build_local_energy expects the native log_psi_squared from build_log_psi_squared")
    if is_complex:
        raise NotImplementedError("complex wavefunctions are not on the native path")
    if is_periodic or include_heg_background:
        raise NotImplementedError("periodic systems (Ewald sums) are not on the native path")

    def get_local_energy(params: Dict, spin_state, r, R, Z, fixed_params: Optional[Dict] = None,
                         return_flags: bool = False):
        n_up, n_dn = spin_state
        model = log_psi_squared.model_for(params, n_up, n_dn)
        if native_chunk:
            model.set_chunk(native_chunk)
        rt = _as_device(r).reshape(-1, n_up + n_dn, 3)
        Rt, Zt = _as_device(R), _as_device(Z)
        e = torch.empty(rt.shape[0], device="cuda", dtype=torch.float32)
        flags = torch.empty(rt.shape[0], device="cuda", dtype=torch.int32)
        model.local_energy(rt, Rt, Zt, e, None, flags)
        if return_flags:
            return e, flags
        return e

    return get_local_energy


def nonfinite_mask(flags: torch.Tensor) -> torch.Tensor:
    return (flags & FLAG_NONFINITE) != 0
