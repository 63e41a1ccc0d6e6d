"""Scalar, per-index restatement of the FermiNet-style log|psi|^2 (test helper for pinning the oracle).

Every quantity is formed element by element with Python loops and the ``math`` module, read directly off
the reference lines, with no reshape, transpose, concatenation or broadcasting: a layout mistake in the
oracle's tensor code (block order of the symmetric features, ion-major el-ion concat, (det, orbital) column
order of the backflow and envelope) changes the oracle but not this code.  Only tiny sizes are meant
(n <= 4, widths of a few units).
"""
from __future__ import annotations

import itertools
import math
from typing import Dict, List

import numpy as np

from deeperwin_amd.params import find_leaf


def _mat(params: Dict, module: str, leaf: str) -> List[List[float]]:
    a = np.asarray(find_leaf(params, module, leaf), dtype=np.float64)
    if a.ndim == 1:
        return [float(x) for x in a]
    return [[float(a[i, j]) for j in range(a.shape[1])] for i in range(a.shape[0])]


def _n_iter(params: Dict) -> int:
    L = 0
    while True:
        try:
            find_leaf(params, f"h_el_{L}/linear_0", "b")
        except KeyError:
            return L
        L += 1


def _det_leibniz(M: List[List[float]]) -> float:
    n = len(M)
    tot = 0.0
    for perm in itertools.permutations(range(n)):
        inversions = sum(1 for a in range(n) for b in range(a + 1, n) if perm[a] > perm[b])
        prod = 1.0
        for i in range(n):
            prod *= M[i][perm[i]]
        tot += (-1.0) ** inversions * prod
    return tot


def softplus(x: float) -> float:
    return math.log1p(math.exp(x))


def el_ion_row(r_i, R) -> List[float]:
    """h_one^0 of one electron: for ion I = 0, 1, ... the four features [|r_i - R_I|, (r_i - R_I)_x, _y, _z]
    one ion after the other (input_features.py:237-239 reshapes [n, I, 4] -> [n, 4 I];
    diff = r_i - R_I, utils.py:271; features [dist, diff] at input_features.py:53-67)."""
    row = []
    for RI in R:
        d = [r_i[c] - RI[c] for c in range(3)]
        row.append(math.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]))
        row.extend(d)
    return row


def el_el_pair(r, i: int, j: int) -> List[float]:
    """h_ee^0[i, j] = [|r_j - r_i|, (r_j - r_i)_x, _y, _z], exactly zero on the diagonal (utils.py:287-305)."""
    if i == j:
        return [0.0, 0.0, 0.0, 0.0]
    d = [r[j][c] - r[i][c] for c in range(3)]
    return [math.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])] + d


def dense_tanh(x: List[float], W: List[List[float]], b: List[float]) -> List[float]:
    """tanh(x W + b), then (x + y)/sqrt(2) only when len(x) == len(y) (mlp.py:13-16, 45-68)."""
    out = []
    for o in range(len(b)):
        acc = b[o]
        for k in range(len(x)):
            acc += x[k] * W[k][o]
        out.append(math.tanh(acc))
    if len(out) == len(x):
        out = [(x[k] + out[k]) / math.sqrt(2.0) for k in range(len(x))]
    return out


def symmetric_row(h, hee, i: int, n_up: int) -> List[float]:
    """Features of electron i (ferminet_embedding.py:55-94, default flags): h_i, then the mean of h over the
    spin-up electrons, then over the spin-down electrons, then the mean over partners j < n_up of
    h_ee[i][j], then over partners j >= n_up; the partner means include j == i."""
    n = len(h)
    F = len(h[0])
    E = len(hee[0][0])
    row = list(h[i])
    row += [sum(h[j][k] for j in range(n_up)) / n_up for k in range(F)]
    row += [sum(h[j][k] for j in range(n_up, n)) / (n - n_up) for k in range(F)]
    row += [sum(hee[i][j][k] for j in range(n_up)) / n_up for k in range(E)]
    row += [sum(hee[i][j][k] for j in range(n_up, n)) / (n - n_up) for k in range(E)]
    return row


def embedding(params: Dict, r, R, n_up: int) -> List[List[float]]:
    n = len(r)
    L = _n_iter(params)
    h = [el_ion_row(r[i], R) for i in range(n)]
    hee = [[el_el_pair(r, i, j) for j in range(n)] for i in range(n)]
    for it in range(L):
        W = _mat(params, f"h_el_{it}/linear_0", "w")
        b = _mat(params, f"h_el_{it}/linear_0", "b")
        h = [dense_tanh(symmetric_row(h, hee, i, n_up), W, b) for i in range(n)]
        if it == L - 1:
            break
        W2 = _mat(params, f"h_el_el_{it}/linear_0", "w")
        b2 = _mat(params, f"h_el_el_{it}/linear_0", "b")
        hee = [[dense_tanh(hee[i][j], W2, b2) for j in range(n)] for i in range(n)]
    return h


def dense_tanh_nores(x: List[float], W: List[List[float]], b: List[float]) -> List[float]:
    """MLP([width]) without residual: tanh(x W + b)."""
    return [math.tanh(b[o] + sum(x[k] * W[k][o] for k in range(len(x)))) for o in range(len(b))]


def embedding_dpe4(params: Dict, r, R, Z, z_min: int, n_up: int) -> List[List[float]]:
    """The dpe4 embedding read off ferminet_embedding.py:52-94, 97-188, 196-267 step by step: the pair
    stream is split into the lists same = [uu pairs row-major, then dd pairs] and diff = [ud pairs, then du
    pairs] (_split_into_same_diff, 196-205); each list goes through its own MLPs; the convolution reshapes
    the mapped lists back into the four blocks w_uu [n_up, n_up], w_ud [n_up, n_dn], w_du [n_dn, n_up],
    w_dd [n_dn, n_dn] (150-153) and forms emb_up[i] = sum_j w_uu[i][j] hm[j] + sum_j w_ud[i][j] hm[n_up + j],
    emb_dn[i] = sum_j w_du[i][j] hm[j] + sum_j w_dd[i][j] hm[n_up + j] (154-164, sum aggregation); the
    el-ion convolution is sum_I h_ei[i][I] him[I] (166-171) with him from the ion embedding table row
    Z_I - Z_min (input_features.py:176-183)."""
    n = len(r)
    n_dn = n - n_up
    L = _n_iter(params)
    h = [el_ion_row(r[i], R) for i in range(n)]
    # el-el features: the distance only (use_el_el_differences=False), zero on the diagonal
    dist = [[el_el_pair(r, i, j)[0] for j in range(n)] for i in range(n)]
    same = [[dist[i][j]] for i in range(n_up) for j in range(n_up)] + \
           [[dist[n_up + i][n_up + j]] for i in range(n_dn) for j in range(n_dn)]
    diff = [[dist[i][n_up + j]] for i in range(n_up) for j in range(n_dn)] + \
           [[dist[n_up + i][j]] for i in range(n_dn) for j in range(n_up)]
    # el-ion stream: [dist, dx, dy, dz] per (electron, ion)
    hei = [[el_ion_row(r[i], [RI]) for RI in R] for i in range(n)]
    table = _mat(params, "input/h_ion", "embeddings")
    h_ion = [table[int(round(float(z))) - z_min] for z in Z]
    for it in range(L):
        conv = f"symm_features_{it}/convolutional_features"
        ws = [dense_tanh_nores(x, _mat(params, f"{conv}/w_same/linear_0", "w"),
                               _mat(params, f"{conv}/w_same/linear_0", "b")) for x in same]
        wd = [dense_tanh_nores(x, _mat(params, f"{conv}/w_diff/linear_0", "w"),
                               _mat(params, f"{conv}/w_diff/linear_0", "b")) for x in diff]
        w_uu = [[ws[i * n_up + j] for j in range(n_up)] for i in range(n_up)]
        w_dd = [[ws[n_up * n_up + i * n_dn + j] for j in range(n_dn)] for i in range(n_dn)]
        w_ud = [[wd[i * n_dn + j] for j in range(n_dn)] for i in range(n_up)]
        w_du = [[wd[n_up * n_dn + i * n_up + j] for j in range(n_up)] for i in range(n_dn)]
        hm = [dense_tanh_nores(h[j], _mat(params, f"{conv}/h_map/linear_0", "w"),
                               _mat(params, f"{conv}/h_map/linear_0", "b")) for j in range(n)]
        E = len(hm[0])
        emb = []
        for i in range(n_up):
            emb.append([sum(w_uu[i][j][f] * hm[j][f] for j in range(n_up)) +
                        sum(w_ud[i][j][f] * hm[n_up + j][f] for j in range(n_dn)) for f in range(E)])
        for i in range(n_dn):
            emb.append([sum(w_du[i][j][f] * hm[j][f] for j in range(n_up)) +
                        sum(w_dd[i][j][f] * hm[n_up + j][f] for j in range(n_dn)) for f in range(E)])
        him = [dense_tanh_nores(x, _mat(params, f"{conv}/h_ion_map/linear_0", "w"),
                                _mat(params, f"{conv}/h_ion_map/linear_0", "b")) for x in h_ion]
        G = len(him[0])
        eei = [[sum(hei[i][I][g] * him[I][g] for I in range(len(R))) for g in range(G)] for i in range(n)]
        F = len(h[0])
        rows = []
        for i in range(n):
            row = list(h[i])
            row += [sum(h[j][k] for j in range(n_up)) / n_up for k in range(F)]
            row += [sum(h[j][k] for j in range(n_up, n)) / n_dn for k in range(F)]
            rows.append(row + emb[i] + eei[i])
        W = _mat(params, f"h_el_{it}/linear_0", "w")
        b = _mat(params, f"h_el_{it}/linear_0", "b")
        h = [dense_tanh(rows[i], W, b) for i in range(n)]
        if it == L - 1:
            break
        same = [dense_tanh(x, _mat(params, f"h_same_{it}/linear_0", "w"), _mat(params, f"h_same_{it}/linear_0", "b"))
                for x in same]
        diff = [dense_tanh(x, _mat(params, f"h_diff_{it}/linear_0", "w"), _mat(params, f"h_diff_{it}/linear_0", "b"))
                for x in diff]
        hei = [[dense_tanh(x, _mat(params, f"h_el_ion_{it}/linear_0", "w"),
                           _mat(params, f"h_el_ion_{it}/linear_0", "b")) for x in hei[i]] for i in range(n)]
    return h


def slater_matrices(params: Dict, r, R, n_up: int, n_det: int, emb=None) -> List[List[List[float]]]:
    """M[d][i][k] for det d, electron row i (up rows first), orbital k.  The backflow of a row comes from the
    spin's bf_* matrix and the envelope from the spin's alpha/weights; column c = d * n + k of those
    matrices belongs to det d, orbital k (envelope_orbitals.py:46-59, 62-75, 96-127).  `emb` replaces the
    FermiNet embedding (a list of per-electron rows)."""
    n = len(r)
    h = embedding(params, r, R, n_up) if emb is None else emb
    mats = [[[0.0] * n for _ in range(n)] for _ in range(n_det)]
    for i in range(n):
        s = "up" if i < n_up else "dn"
        Wbf = _mat(params, f"bf_{s}/linear_0", "w")
        alpha = _mat(params, "envelope_orbitals", f"alpha_{s}")
        wenv = _mat(params, "envelope_orbitals", f"weights_{s}")
        for d in range(n_det):
            for k in range(n):
                c = d * n + k
                bf = sum(h[i][q] * Wbf[q][c] for q in range(len(h[i])))
                env = 0.0
                for I, RI in enumerate(R):
                    rho = math.sqrt(sum((r[i][x] - RI[x]) ** 2 for x in range(3)))
                    env += wenv[I][c] * math.exp(-softplus(alpha[I][c]) * rho)
                mats[d][i][k] = env * bf
    return mats


def log_psi_sqr(params: Dict, r, R, n_up: int, n_det: int, emb=None):
    """(phase, log|psi|^2) of one walker: psi~ = sum_d det_d e^{-m}, m = max_d log|det_d|, and
    log|psi|^2 = 2 (log(|psi~| + 1e-8) + m) (wavefunction.py:63-83)."""
    dets = [_det_leibniz(M) for M in slater_matrices(params, r, R, n_up, n_det, emb)]
    m = max(math.log(abs(x)) for x in dets)
    psi = sum(x * math.exp(-m) for x in dets)
    return (math.pi if psi < 0 else 0.0), 2.0 * (math.log(abs(psi) + 1e-8) + m)
