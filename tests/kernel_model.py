"""NumPy float64 mirror of the native kernel pipeline (test helper, not the oracle).

It restates, stage by stage, the algorithm the HIP kernels run — channel rows (value, 3 n tangents,
Laplacian), the spin-mean split of the symmetric features, the pair-mean tangent rows of the 2-el
stream, the tanh rule, the envelope product rule, the determinant traces and the epsilon-literal
sum-of-determinants combination — so that the mathematics of that design can be checked against the
oracle on the CPU, independently of the GPU.
"""
from __future__ import annotations

import numpy as np

from deeperwin_amd.params import find_leaf


def _pair_means(val, T, Lp, u, lap):
    """val [B,n,n,F], T [B,n,n,3,F], Lp [B,n,n,F] -> rows [B, n, C, 2F]."""
    B, n, _, F = val.shape
    C = 3 * n + 2 if lap else 1
    out = np.zeros((B, n, C, 2 * F), val.dtype)
    blocks = [(0, u), (u, n)]
    for s, (j0, j1) in enumerate(blocks):
        ns = j1 - j0
        out[:, :, 0, s * F:(s + 1) * F] = val[:, :, j0:j1].sum(2) / ns
        if lap:
            out[:, :, -1, s * F:(s + 1) * F] = Lp[:, :, j0:j1].sum(2) / ns
            Tsum = T[:, :, j0:j1].sum(2) / ns  # [B, n, 3, F]
            for e in range(j0, j1):
                for c in range(3):
                    out[:, :, 1 + 3 * e + c, s * F:(s + 1) * F] += T[:, :, e, c] / ns
            for i in range(n):
                for c in range(3):
                    out[:, i, 1 + 3 * i + c, s * F:(s + 1) * F] -= Tsum[:, i, c]
    return out


def _tanh_rule(Z, lap):
    t = np.tanh(Z[:, :, 0])
    if not lap:
        return t[:, :, None]
    d1 = 1 - t ** 2
    d2 = -2 * t * d1
    out = d1[:, :, None] * Z
    out[:, :, 0] = t
    S = (Z[:, :, 1:-1] ** 2).sum(2)
    out[:, :, -1] = d1 * Z[:, :, -1] + d2 * S
    return out


def run(params, n_up, r, R, Z, n_det, lap=True, dtype=np.float64, det_dtype=np.float64):
    """dtype: arithmetic of the network stages; det_dtype: arithmetic of the determinant stage."""
    r = np.asarray(r, dtype)
    R = np.asarray(R, dtype)
    B, n, _ = r.shape
    u, I, nd = n_up, R.shape[0], 3 * n
    C = nd + 2 if lap else 1
    # ---- layer-0 one-electron rows ----
    d = r[:, :, None, :] - R[None, None]
    rho = np.linalg.norm(d, axis=-1)
    h = np.zeros((B, n, C, I, 4), dtype)
    h[:, :, 0, :, 0] = rho
    h[:, :, 0, :, 1:] = d
    if lap:
        for i in range(n):
            for c in range(3):
                h[:, i, 1 + 3 * i + c, :, 0] = d[:, i, :, c] / rho[:, i]
                h[:, i, 1 + 3 * i + c, :, 1 + c] = 1.0
        h[:, :, -1, :, 0] = 2.0 / rho
    h = h.reshape(B, n, C, 4 * I)
    # ---- two-electron stream ----
    dd = r[:, None, :, :] - r[:, :, None, :]  # r_j - r_i
    eye = np.eye(n, dtype=bool)
    rho_ee = np.linalg.norm(dd, axis=-1)
    rho_safe = np.where(eye, 1.0, rho_ee)
    val = np.concatenate([np.where(eye, 0.0, rho_ee)[..., None], dd], -1)
    T = np.zeros((B, n, n, 3, 4), dtype)
    for c in range(3):
        T[:, :, :, c, 0] = np.where(eye, 0.0, dd[..., c] / rho_safe)
        T[:, :, :, c, 1 + c] = np.where(eye, 0.0, 1.0)
    Lp = np.zeros((B, n, n, 4), dtype)
    Lp[..., 0] = np.where(eye, 0.0, 4.0 / rho_safe)
    L = 0
    while True:
        try:
            find_leaf(params, f"h_el_{L}/linear_0", "w")
        except KeyError:
            break
        L += 1
    pms = [_pair_means(val, T, Lp, u, lap)]
    for l2 in range(L - 1):
        W = np.asarray(find_leaf(params, f"h_el_el_{l2}/linear_0", "w"), dtype)
        b = np.asarray(find_leaf(params, f"h_el_el_{l2}/linear_0", "b"), dtype)
        z = val @ W + b
        zT = T @ W
        zL = Lp @ W
        t = np.tanh(z)
        d1 = 1 - t ** 2
        d2 = -2 * t * d1
        nval, nT = t, d1[:, :, :, None] * zT
        nL = d1 * zL + d2 * 2.0 * (zT ** 2).sum(3)
        if W.shape[0] == W.shape[1]:
            s2 = dtype(np.sqrt(2.0))  # a float64 scalar would promote the stream under NEP 50
            nval, nT, nL = (val + nval) / s2, (T + nT) / s2, (Lp + nL) / s2
        val, T, Lp = nval, nT, nL
        pms.append(_pair_means(val, T, Lp, u, lap))
    # ---- one-electron layers with the spin-mean split ----
    for l in range(L):
        W = np.asarray(find_leaf(params, f"h_el_{l}/linear_0", "w"), dtype)
        b = np.asarray(find_leaf(params, f"h_el_{l}/linear_0", "b"), dtype)
        Kh = h.shape[-1]
        mu_up = h[:, :u].mean(1)   # [B, C, Kh]
        mu_dn = h[:, u:].mean(1)
        V = mu_up @ W[Kh:2 * Kh] + mu_dn @ W[2 * Kh:3 * Kh]
        V[:, 0] += b
        Zp = h @ W[:Kh] + pms[l] @ W[3 * Kh:] + V[:, None]
        h = _tanh_rule(Zp, lap)
    # ---- orbitals ----
    NO = n_det * n
    bf = np.zeros((B, n, C, NO), dtype)
    Wu = np.asarray(find_leaf(params, "bf_up/linear_0", "w"), dtype)
    Wd = np.asarray(find_leaf(params, "bf_dn/linear_0", "w"), dtype)
    bf[:, :u] = h[:, :u] @ Wu
    bf[:, u:] = h[:, u:] @ Wd
    A = np.zeros((B, n_det, C, n, n), dtype)
    for s, (i0, i1) in enumerate([(0, u), (u, n)]):
        sfx = "up" if s == 0 else "dn"
        kap = np.log1p(np.exp(np.asarray(find_leaf(params, "envelope_orbitals", f"alpha_{sfx}"), dtype)))
        we = np.asarray(find_leaf(params, "envelope_orbitals", f"weights_{sfx}"), dtype)
        for i in range(i0, i1):
            e = we[None] * np.exp(-kap[None] * rho[:, i, :, None])        # [B, I, NO]
            env = e.sum(1)
            genv = np.stack([(-kap[None] * e / rho[:, i, :, None] * d[:, i, :, c, None]).sum(1)
                             for c in range(3)], 1)                          # [B, 3, NO]
            lenv = (e * (kap[None] ** 2 - 2 * kap[None] / rho[:, i, :, None])).sum(1)
            a = env[:, None] * bf[:, i]                                       # [B, C, NO]
            if lap:
                for c in range(3):
                    a[:, 1 + 3 * i + c] += genv[:, c] * bf[:, i, 0]
                a[:, -1] = lenv * bf[:, i, 0] + env * bf[:, i, -1] + 2 * (genv * bf[:, i, 1 + 3 * i:4 + 3 * i]).sum(1)
            A[:, :, :, i, :] = a.reshape(B, C, n_det, n).transpose(0, 2, 1, 3)
    # ---- determinants ----
    assert A.dtype == dtype and h.dtype == dtype, (A.dtype, h.dtype)
    A = A.astype(det_dtype)
    sgn, ld = np.linalg.slogdet(A[:, :, 0])
    out = dict(ld=ld, sd=sgn)
    m = ld.max(1, keepdims=True)
    ew = sgn * np.exp(ld - m)
    S = ew.sum(1)
    eps = 1e-8
    out["log_psi_sqr"] = 2 * (np.log(np.abs(S) + eps) + m[:, 0])
    if not lap:
        return out
    Minv = np.linalg.inv(A[:, :, 0])
    X = np.einsum("bdac,bdkcx->bdkax", Minv, A[:, :, 1:-1])
    g = np.einsum("bdkaa->bdk", X)
    q = np.einsum("bdkac,bdkca->bdk", X, X)
    lapd = np.einsum("bdac,bdca->bd", Minv, A[:, :, -1]) - q.sum(2)
    is_max = (ld == m)
    cnt = is_max.sum(1)
    dm = (g * is_max[..., None]).sum(1) / cnt[:, None]
    lapm = (lapd * is_max).sum(1) / cnt
    x = g - dm[:, None]
    dS = (ew[..., None] * x).sum(1)
    d2S = (ew * ((x ** 2).sum(2) + lapd - lapm[:, None])).sum(1)
    sg = np.sign(S)
    den = np.abs(S) + eps
    df = sg[:, None] * dS / den[:, None] + dm
    lapf = sg * d2S / den - (dS ** 2).sum(1) / den ** 2 + lapm
    out["ekin"] = -0.5 * (lapf + (df ** 2).sum(1))
    out["grad_logpsi"] = df
    return out
