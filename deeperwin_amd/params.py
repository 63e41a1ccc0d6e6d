"""Trainable parameters of the FermiNet-style wavefunction, as a haiku-style nested dict.

Layout (module path -> leaf -> array), following the reference's module names:

* ``wf/ferminet_embedding/h_el_{i}/linear_0``      w [3*d_i + 2*e_i, W1], b [W1]     i = 0..L-1
  (ferminet_embedding.py:233-241, mlp.py:57-59; d_0 = 4*n_ion, e_0 = 4, then d = W1, e = W2)
* ``wf/ferminet_embedding/h_el_el_{i}/linear_0``   w [e_i, W2], b [W2]               i = 0..L-2
  (ferminet_embedding.py:254-257)
* ``wf/orbitals/envelope_orbitals/bf_up/linear_0`` w [W1, D*n_el]  (no bias; envelope_orbitals.py:46-52)
* ``wf/orbitals/envelope_orbitals/bf_dn/linear_0`` w [W1, D*n_el]  (envelope_orbitals.py:62-68)
* ``wf/orbitals/envelope_orbitals``  alpha_up/alpha_dn/weights_up/weights_dn [n_ion, D*n_el], init 1
  (envelope_orbitals.py:99-104)

The dpe4 model (the reference's default, configuration.py:875-891, 422-441, 666-674) replaces the 2-el
stream by same/diff-spin streams with SchNet-style convolutions and adds an el-ion stream
(ferminet_embedding.py:97-188, 207-267); its extra leaves, per iteration i (2-el leaves for i < L-1):

* ``.../symm_features_{i}/convolutional_features/w_same/linear_0``, ``w_diff``  w [f_i, E], b [E]
  (f_0 = 1: the el-el distance; then W2)
* ``.../symm_features_{i}/convolutional_features/h_map/linear_0``   w [d_i, E], b [E]
* ``.../symm_features_{i}/convolutional_features/h_ion_map/linear_0`` w [32, g_i], b [g_i]  (g_0 = 4, then W2)
* ``wf/ferminet_embedding/h_same_{i}/linear_0``, ``h_diff_{i}``, ``h_el_ion_{i}`` (w [f_i or g_i, W2])
* ``wf/input/h_ion`` embeddings [Z_max - Z_min + 1, 32]  (input_features.py:176-183, hk.Embed)
* ``h_el_{i}/linear_0`` w [3 d_i + E + g_i, W1]  (E = emb_dim = W2 = 32)

Initialisation is the reference's VarianceScaling(1, fan_avg, uniform) for weights and zero biases
(mlp.py:42-43, configuration.py:215-227), and haiku's truncated-normal (stddev 1) for the ion embedding
table, drawn from a NumPy generator: haiku's PRNG stream is not reproduced, so "seed s" means
``numpy.random.default_rng(s)``.

Module paths are matched by suffix (``find_leaf``), so dicts whose keys carry haiku's ``~`` segments
(``wf/~/orbitals/...``) are accepted as well.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Dict, Tuple

import numpy as np

EMB = "wf/ferminet_embedding"
ENV = "wf/orbitals/envelope_orbitals"
ION_EMB = "wf/input/h_ion"
N_ION_FEATURES = 32  # InputFeatureConfig.n_ion_features (configuration.py:590)


@dataclass(frozen=True)
class ModelShape:
    n_el: int
    n_up: int
    n_ion: int
    n_det: int
    width1: int = 256
    width2: int = 32
    n_iter: int = 4
    embedding: str = "ferminet"  # or "dpe4"
    z_min: int = 1               # dpe4: Z of the ion embedding table's first row
    n_ion_types: int = 0         # dpe4: rows of the ion embedding table (Z_max - Z_min + 1)

    @property
    def n_dn(self) -> int:
        return self.n_el - self.n_up

    def one_el_in(self, i: int) -> int:
        d = 4 * self.n_ion if i == 0 else self.width1
        if self.embedding == "dpe4":
            return 3 * d + self.width2 + self.el_ion_width(i)
        e = 4 if i == 0 else self.width2
        return 3 * d + 2 * e

    def two_el_in(self, i: int) -> int:
        if self.embedding == "dpe4":
            return 1 if i == 0 else self.width2
        return 4 if i == 0 else self.width2

    def el_ion_width(self, i: int) -> int:
        """dpe4: width of the el-ion stream entering iteration i ([dist, diff] first)."""
        return 4 if i == 0 else self.width2


def _uniform_fan_avg(rng: np.random.Generator, fan_in: int, fan_out: int) -> np.ndarray:
    limit = np.sqrt(3.0 / max(1.0, (fan_in + fan_out) / 2.0))
    return rng.uniform(-limit, limit, size=(fan_in, fan_out))


def _dense(rng, fin: int, fout: int, dtype):
    return dict(w=_uniform_fan_avg(rng, fin, fout).astype(dtype), b=np.zeros(fout, dtype))


def init_params(shape: ModelShape, seed: int, dtype=np.float32) -> Dict[str, Dict[str, np.ndarray]]:
    if shape.embedding == "dpe4":
        return init_params_dpe4(shape, seed, dtype)
    rng = np.random.default_rng(seed)
    p: Dict[str, Dict[str, np.ndarray]] = {}
    for i in range(shape.n_iter):
        fin = shape.one_el_in(i)
        p[f"{EMB}/h_el_{i}/linear_0"] = dict(w=_uniform_fan_avg(rng, fin, shape.width1).astype(dtype),
                                               b=np.zeros(shape.width1, dtype))
        if i < shape.n_iter - 1:
            fin2 = shape.two_el_in(i)
            p[f"{EMB}/h_el_el_{i}/linear_0"] = dict(w=_uniform_fan_avg(rng, fin2, shape.width2).astype(dtype),
                                                      b=np.zeros(shape.width2, dtype))
    _orbital_params(p, rng, shape, dtype)
    return p


def _orbital_params(p, rng, shape: ModelShape, dtype):
    n_orb = shape.n_det * shape.n_el
    p[f"{ENV}/bf_up/linear_0"] = dict(w=_uniform_fan_avg(rng, shape.width1, n_orb).astype(dtype))
    p[f"{ENV}/bf_dn/linear_0"] = dict(w=_uniform_fan_avg(rng, shape.width1, n_orb).astype(dtype))
    p[ENV] = dict(alpha_up=np.ones((shape.n_ion, n_orb), dtype), alpha_dn=np.ones((shape.n_ion, n_orb), dtype),
                  weights_up=np.ones((shape.n_ion, n_orb), dtype), weights_dn=np.ones((shape.n_ion, n_orb), dtype))


def init_params_dpe4(shape: ModelShape, seed: int, dtype=np.float32) -> Dict[str, Dict[str, np.ndarray]]:
    """Parameters of the dpe4 model (module paths of ferminet_embedding.py:97-188, 207-267 and
    input_features.py:176-183 under 'wf')."""
    if shape.n_ion_types < 1:
        raise ValueError("dpe4 needs n_ion_types >= 1 (rows of the ion embedding table)")
    rng = np.random.default_rng(seed)
    p: Dict[str, Dict[str, np.ndarray]] = {}
    E = shape.width2
    for i in range(shape.n_iter):
        conv = f"{EMB}/symm_features_{i}/convolutional_features"
        d = 4 * shape.n_ion if i == 0 else shape.width1
        p[f"{conv}/w_same/linear_0"] = _dense(rng, shape.two_el_in(i), E, dtype)
        p[f"{conv}/w_diff/linear_0"] = _dense(rng, shape.two_el_in(i), E, dtype)
        p[f"{conv}/h_map/linear_0"] = _dense(rng, d, E, dtype)
        p[f"{conv}/h_ion_map/linear_0"] = _dense(rng, N_ION_FEATURES, shape.el_ion_width(i), dtype)
        p[f"{EMB}/h_el_{i}/linear_0"] = _dense(rng, shape.one_el_in(i), shape.width1, dtype)
        if i < shape.n_iter - 1:
            p[f"{EMB}/h_same_{i}/linear_0"] = _dense(rng, shape.two_el_in(i), shape.width2, dtype)
            p[f"{EMB}/h_diff_{i}/linear_0"] = _dense(rng, shape.two_el_in(i), shape.width2, dtype)
            p[f"{EMB}/h_el_ion_{i}/linear_0"] = _dense(rng, shape.el_ion_width(i), shape.width2, dtype)
    table = np.clip(rng.standard_normal((shape.n_ion_types, N_ION_FEATURES)), -2.0, 2.0)
    p[ION_EMB] = dict(embeddings=table.astype(dtype))
    _orbital_params(p, rng, shape, dtype)
    return p


def is_dpe4(params: Dict) -> bool:
    try:
        find_leaf(params, "symm_features_0/convolutional_features/h_map/linear_0", "w")
        return True
    except KeyError:
        return False


def find_leaf(params: Dict, module_suffix: str, leaf: str):
    """Return params[path][leaf] for the unique path ending in ``module_suffix`` (ignoring ``~`` segments)."""
    hits = []
    for path, leaves in params.items():
        norm = "/".join(s for s in path.split("/") if s != "~")
        if norm == module_suffix or norm.endswith("/" + module_suffix):
            if leaf in leaves:
                hits.append(leaves[leaf])
    if len(hits) != 1:
        raise KeyError(f"expected exactly one parameter '{module_suffix}/{leaf}', found {len(hits)}")
    return hits[0]


def infer_shape(params: Dict, n_el: int, n_up: int, z_min: int = None) -> ModelShape:
    """The ModelShape of `params`.  For dpe4 parameters `z_min` (the Z of the ion table's first row, by
    default 1) is not stored in the parameters and has to be given."""
    w_up = np.asarray(find_leaf(params, "bf_up/linear_0", "w"))
    alpha = np.asarray(find_leaf(params, "envelope_orbitals", "alpha_up"))
    n_ion = alpha.shape[0]
    width1 = w_up.shape[0]
    n_det = w_up.shape[1] // n_el
    if n_det * n_el != w_up.shape[1]:
        raise ValueError(f"backflow width {w_up.shape[1]} is not a multiple of n_el={n_el}")
    n_iter = 0
    while True:
        try:
            find_leaf(params, f"h_el_{n_iter}/linear_0", "w")
        except KeyError:
            break
        n_iter += 1
    if is_dpe4(params):
        width2 = np.asarray(find_leaf(params, "symm_features_0/convolutional_features/w_same/linear_0",
                                      "w")).shape[1]
        types = np.asarray(find_leaf(params, "input/h_ion", "embeddings")).shape[0]
        shape = ModelShape(n_el, n_up, n_ion, n_det, width1, width2, n_iter, "dpe4",
                           1 if z_min is None else int(z_min), types)
    else:
        width2 = np.asarray(find_leaf(params, "h_el_el_0/linear_0", "w")).shape[1]
        shape = ModelShape(n_el, n_up, n_ion, n_det, width1, width2, n_iter)
    w0 = np.asarray(find_leaf(params, "h_el_0/linear_0", "w"))
    if w0.shape != (shape.one_el_in(0), width1):
        raise ValueError(f"h_el_0 weight has shape {w0.shape}, expected {(shape.one_el_in(0), width1)}")
    return shape


def truncate_params(params: Dict, n_ion_max: int, n_el_max: int, shape: ModelShape) -> Dict:
    """Per-geometry parameters cut out of a shared max-size set.

    Layer 0's weight rows are [h (4*I_max) | mean_up h (4*I_max) | mean_dn h (4*I_max) | pair up (4) |
    pair dn (4)]; each h block keeps its first 4*I rows.  Orbital columns are (det, orbital) pairs with
    n_el_max orbitals per det; each det keeps its first n_el orbitals; envelope tensors keep the first
    I ions.  Every other weight is shape-independent and shared as is.
    """
    I, n, D = shape.n_ion, shape.n_el, shape.n_det
    out: Dict[str, Dict[str, np.ndarray]] = {}
    for path, leaves in params.items():
        out[path] = dict(leaves)
    w0 = np.asarray(find_leaf(params, "h_el_0/linear_0", "w"))
    dmax = 4 * n_ion_max
    if w0.shape[0] != 3 * dmax + 8:
        raise ValueError("layer-0 weight does not match n_ion_max")
    rows = np.concatenate([np.arange(0, 4 * I), dmax + np.arange(0, 4 * I), 2 * dmax + np.arange(0, 4 * I),
                           3 * dmax + np.arange(8)])
    for path in out:
        if path.endswith("h_el_0/linear_0"):
            out[path]["w"] = w0[rows]
    cols = (np.arange(D)[:, None] * n_el_max + np.arange(n)[None, :]).reshape(-1)
    for path in out:
        if path.endswith("bf_up/linear_0") or path.endswith("bf_dn/linear_0"):
            out[path]["w"] = np.asarray(out[path]["w"])[:, cols]
        if path.endswith("envelope_orbitals") and "alpha_up" in out[path]:
            for k in ("alpha_up", "alpha_dn", "weights_up", "weights_dn"):
                out[path][k] = np.asarray(out[path][k])[:I][:, cols]
    return out


def param_count(params: Dict) -> int:
    return int(sum(np.asarray(v).size for leaves in params.values() for v in leaves.values()))


def cast(params: Dict, dtype) -> Dict:
    return {p: {k: np.asarray(v).astype(dtype) for k, v in leaves.items()} for p, leaves in params.items()}


def shape_tuple(shape: ModelShape) -> Tuple[int, ...]:
    return (shape.n_el, shape.n_up, shape.n_ion, shape.n_det, shape.width1, shape.width2, shape.n_iter,
            shape.embedding, shape.z_min, shape.n_ion_types)
