"""JAX's random streams on the device (jax 0.4.23 default: threefry2x32), so that the sampler draws what the
reference draws: ``PRNGKey``, ``split``, ``normal`` and the per-walker MCMC stream (mcmc.py:43, 90,
175-180, 345-378).  Keys are CUDA int32 tensors holding the uint32 words ([2] or [B, 2]).
"""
from __future__ import annotations

import numpy as np

from ._native import _check, _ptr, _stream, load_library


def _torch():
    import torch

    return torch


def PRNGKey(seed: int):
    """jax.random.PRNGKey(seed) = [seed >> 32, seed & 0xffffffff] as a CUDA int32 tensor of uint32 bits."""
    k = np.array([(seed >> 32) & 0xFFFFFFFF, seed & 0xFFFFFFFF], np.uint32).view(np.int32)
    return _torch().tensor(k, device="cuda")


def as_key(key):
    torch = _torch()
    if isinstance(key, torch.Tensor):
        return key.to(device="cuda", dtype=torch.int32).contiguous()
    return torch.tensor(np.asarray(key, np.uint32).view(np.int32), device="cuda")


def split(key, num: int = 2):
    """jax.random.split(key, num) -> [num, 2]."""
    torch = _torch()
    key = as_key(key)
    out = torch.empty((num, 2), dtype=torch.int32, device="cuda")
    _check(load_library().dw_threefry_split(_ptr(key), int(num), _ptr(out), _stream()))
    return out


def normal(key, shape):
    """jax.random.normal(key, shape), float32."""
    torch = _torch()
    key = as_key(key)
    out = torch.empty(tuple(shape), dtype=torch.float32, device="cuda")
    _check(load_library().dw_threefry_normal(_ptr(key), int(out.numel()), _ptr(out), _stream()))
    return out


def mcmc_draws(keys, n_el: int, n_steps: int):
    """Proposal normals [S, B, n, 3], accept uniforms [S, B] and the walker keys after S steps."""
    torch = _torch()
    keys = as_key(keys)
    B = keys.shape[0]
    normals = torch.empty((n_steps, B, n_el, 3), dtype=torch.float32, device="cuda")
    uniforms = torch.empty((n_steps, B), dtype=torch.float32, device="cuda")
    new_keys = torch.empty_like(keys)
    _check(load_library().dw_threefry_mcmc_draws(_ptr(keys), _ptr(new_keys), int(B), int(n_el), int(n_steps),
                                                 _ptr(normals), _ptr(uniforms), _stream()))
    return normals, uniforms, new_keys


def uint32(t) -> np.ndarray:
    """Host copy of a key tensor as uint32."""
    return t.detach().cpu().numpy().astype(np.int32).view(np.uint32)
