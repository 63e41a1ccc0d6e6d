"""Molecule set-up: electron-to-ion mapping, Gaussian walker initialisation and the seeded synthetic
geometries of the five workloads (LiH, N2, C4H2-like, a 64-geometry batch and a request stream).

Reference behaviour restated here:
* ``generate_el_ion_mapping``   <- PhysicalConfig._generate_el_ion_mapping (configuration.py:1572-1615)
* ``init_walkers_gaussian``     <- MCMCState.initialize_around_nuclei, Gaussian branch (mcmc.py:40-91, 64-67)
"""
from __future__ import annotations

import math
from typing import List, Sequence, Tuple

import numpy as np


def generate_el_ion_mapping(R, Z, n_el: int, n_up: int) -> List[int]:
    """Greedy placement of electrons on ions that keeps the exp(-distance)-weighted local spin small.

    For every electron, all 2*n_ion (spin, ion) choices are scored by the largest absolute weighted local
    spin they would produce; the best-scoring choice whose ion is not yet full and whose spin still has
    electrons left wins.  The result lists ion indices, spin-up electrons first (configuration.py:1572-1615).
    """
    R = np.asarray(R, dtype=np.float64)
    n_ions = len(Z)
    w = np.exp(-np.linalg.norm(R[:, None, :] - R[None, :, :], axis=-1))
    counts = np.zeros((n_ions, 2), dtype=np.int64)
    left = [n_up, n_el - n_up]
    for _ in range(n_el):
        spin_now = w @ (counts[:, 0] - counts[:, 1])
        scores = []
        for s in (0, 1):
            for ion in range(n_ions):
                # adding one electron of spin s on `ion` changes the local spin by +-w[:, ion]
                scores.append(np.max(np.abs(spin_now + (1 if s == 0 else -1) * w[:, ion])))
        order = np.argsort(np.asarray(scores))
        for idx in order:
            s, ion = divmod(int(idx), n_ions)
            if counts[ion].sum() == Z[ion] or left[s] == 0:
                continue
            counts[ion, s] += 1
            left[s] -= 1
            break
    mapping: List[int] = []
    for s in (0, 1):
        for ion in range(n_ions):
            mapping += [ion] * int(counts[ion, s])
    return mapping


def init_walkers_gaussian(rng: np.random.Generator, n_walkers: int, R, el_ion_mapping, sigma: float = 1.0):
    """r0 = R[el_ion_mapping] + sigma * N(0, 1)  (mcmc.py:64-67); returns float64 [B, n_el, 3]."""
    R = np.asarray(R, dtype=np.float64)
    n_el = len(el_ion_mapping)
    return R[np.asarray(el_ion_mapping)][None] + sigma * rng.standard_normal((n_walkers, n_el, 3))


# --------------------------------------------------------------------------------------------------
# Seeded synthetic geometries
# --------------------------------------------------------------------------------------------------

def geometry_lih() -> Tuple[np.ndarray, np.ndarray]:
    """2 nuclei Z=[3,1] at [[0,0,0],[3,0,0]] bohr."""
    return np.array([[0.0, 0.0, 0.0], [3.0, 0.0, 0.0]]), np.array([3, 1])


def geometry_n2() -> Tuple[np.ndarray, np.ndarray]:
    """N2 with 2.07 bohr bond."""
    return np.array([[0.0, 0.0, 0.0], [2.07, 0.0, 0.0]]), np.array([7, 7])


def geometry_c4h2(seed: int) -> Tuple[np.ndarray, np.ndarray]:
    """4 carbons on a seed-generated planar rhombus with 2.6 bohr sides (interior angle drawn in
    [60, 120] degrees), plus two hydrogens 2.0 bohr outward from two opposite carbons."""
    rng = np.random.default_rng(seed)
    theta = math.radians(rng.uniform(60.0, 120.0))
    a = 2.6
    v = a * np.array([math.cos(theta), math.sin(theta), 0.0])
    C = np.array([[0.0, 0.0, 0.0], [a, 0.0, 0.0], [a, 0.0, 0.0] + v, v])
    C -= C.mean(axis=0)
    H = []
    for idx in (0, 2):
        d = C[idx] / np.linalg.norm(C[idx])
        H.append(C[idx] + 2.0 * d)
    R = np.concatenate([C, np.array(H)], axis=0)
    Z = np.array([6, 6, 6, 6, 1, 1])
    return R, Z


def random_molecule(rng: np.random.Generator, n_el_target: int, charges: Sequence[int] = (1, 6, 7, 8),
                    d_min: float = 1.8, d_max: float = 2.8) -> Tuple[np.ndarray, np.ndarray]:
    """Neutral random molecule with exactly ``n_el_target`` electrons.

    Atoms with Z drawn from ``charges`` (never overshooting the electron budget) are attached one by one
    at a distance in [d_min, d_max] from a random existing atom, rejecting placements closer than d_min
    to any atom; the budget left at the end is filled with hydrogens placed the same way.
    """
    R = [np.zeros(3)]
    first = [z for z in charges if z <= n_el_target]
    Z = [int(rng.choice(first))]
    while sum(Z) < n_el_target:
        left = n_el_target - sum(Z)
        allowed = [z for z in charges if z <= left]
        z = int(rng.choice(allowed))
        for _ in range(1000):
            anchor = R[int(rng.integers(len(R)))]
            direction = rng.standard_normal(3)
            direction /= np.linalg.norm(direction)
            pos = anchor + rng.uniform(d_min, d_max) * direction
            if min(np.linalg.norm(pos - p) for p in R) >= d_min - 1e-9:
                break
        R.append(pos)
        Z.append(z)
    R = np.array(R)
    R -= R.mean(axis=0)
    return R, np.array(Z)


def batch_geometries(seed: int = 4, n_geom: int = 64, n_min: int = 4, n_max: int = 42):
    """64 seeded molecules with electron counts uniform in [4, 42]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_geom):
        n = int(rng.integers(n_min, n_max + 1))
        out.append(random_molecule(rng, n))
    return out
