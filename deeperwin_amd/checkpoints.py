"""Checkpoint import and export in the reference's run-file format (checkpoints.py:24-93).

A run file is a zip archive holding ``config.yml``, ``history.csv``, ``summary.csv`` and one pickle per
non-empty RunData field (``params.pkl``, ``ema_params.pkl``, ``fixed_params.pkl``, ``mcmc_state.pkl``,
``clipping_state.pkl``, ...).  ``load_run`` reads it WITHOUT executing anything from the file: the pickles are
read by an unpickler that resolves only an allow-list of constructors (NumPy arrays, dtypes and scalars,
built-in containers, the dataclass reconstructor for the reference's MCMCState, and jax's array
reconstructor, which is answered with the NumPy array it wraps); any other global in the stream is refused
with ``pickle.UnpicklingError`` before it is loaded.  ``params_from_run`` returns the haiku parameter dict
(float32 NumPy leaves) that ``NativeModel.set_params`` / the drop-in ``log_psi_sqr`` take unchanged.
``save_run`` writes the same layout (the pickles hold NumPy arrays and plain containers).
"""
from __future__ import annotations

import copyreg
import io
import os
import pickle
import zipfile
from dataclasses import dataclass, fields
from typing import Any, Dict, List, Optional

import numpy as np


@dataclass
class RunData:
    config: Optional[Any] = None
    history: Optional[List[dict]] = None
    summary: Optional[dict] = None
    metadata: Optional[dict] = None
    params: Optional[dict] = None
    ema_params: Optional[dict] = None
    fixed_params: Optional[dict] = None
    opt_state: Optional[Any] = None
    mcmc_state: Optional[Any] = None
    clipping_state: Optional[Any] = None


class StateRecord:
    """Stand-in for the reference's pickled dataclasses (e.g. MCMCState): keeps their field dict."""

    def __setstate__(self, state):
        if isinstance(state, tuple) and len(state) == 2:  # (dict, slots)
            state = {**(state[0] or {}), **(state[1] or {})}
        self.__dict__.update(state)


def _jax_reconstruct_array(fun, args, arr_state, aval_state):
    """jax.Array pickles as (jax._src.array._reconstruct_array, (fun, args, arr_state, aval_state)) where
    (fun, args, arr_state) is the NumPy array's own reduction; `fun` has already passed the allow-list."""
    arr = fun(*args)
    arr.__setstate__(arr_state)
    return arr


def _np_multiarray():
    try:
        from numpy._core import multiarray  # NumPy 2
    except ImportError:  # pragma: no cover
        from numpy.core import multiarray
    return multiarray


_MA = _np_multiarray()
_ALLOWED = {
    ("numpy.core.multiarray", "_reconstruct"): _MA._reconstruct,
    ("numpy._core.multiarray", "_reconstruct"): _MA._reconstruct,
    ("numpy.core.multiarray", "scalar"): _MA.scalar,
    ("numpy._core.multiarray", "scalar"): _MA.scalar,
    ("numpy", "ndarray"): np.ndarray,
    ("numpy", "dtype"): np.dtype,
    ("builtins", "dict"): dict,
    ("builtins", "list"): list,
    ("builtins", "tuple"): tuple,
    ("builtins", "set"): set,
    ("builtins", "frozenset"): frozenset,
    ("builtins", "complex"): complex,
    ("collections", "OrderedDict"): dict,
    ("copyreg", "__newobj__"): copyreg.__newobj__,
    ("copyreg", "_reconstructor"): copyreg._reconstructor,
    ("builtins", "object"): object,
    ("jax._src.array", "_reconstruct_array"): _jax_reconstruct_array,
    ("deeperwin.mcmc", "MCMCState"): StateRecord,
}


class _AllowListUnpickler(pickle.Unpickler):
    def find_class(self, module, name):
        fn = _ALLOWED.get((module, name))
        if fn is None:
            raise pickle.UnpicklingError(f"refusing to load {module}.{name}: not an array or container type")
        return fn


def loads(data: bytes):
    """Unpickle `data` with the allow-list (no code from the stream runs)."""
    return _AllowListUnpickler(io.BytesIO(data)).load()


def load_run(fname: str, load_pkl: bool = True, parse_config: bool = True) -> RunData:
    """load_run (checkpoints.py:68-86): the pickles through the allow-list unpickler, config.yml with
    yaml.safe_load, history/summary CSVs as text rows."""
    data = RunData()
    with zipfile.ZipFile(fname, "r") as zf:
        for field in zf.namelist():
            key, ext = os.path.splitext(field)
            raw = zf.read(field)
            if ext == ".pkl" and load_pkl and key in {f.name for f in fields(RunData)}:
                setattr(data, key, loads(raw))
            elif field == "config.yml" and parse_config:
                import yaml

                data.config = yaml.safe_load(raw.decode("utf-8"))
            elif ext == ".csv":
                rows = [line.split(";") for line in raw.decode("utf-8").splitlines() if line]
                setattr(data, key, rows)
    return data


def _as_numpy_tree(x):
    if isinstance(x, dict):
        return {k: _as_numpy_tree(v) for k, v in x.items()}
    if hasattr(x, "detach"):
        return x.detach().cpu().numpy()
    return x


def save_run(fname: str, data: RunData) -> None:
    """save_run (checkpoints.py:49-65) with NumPy leaves: config.yml, history/summary CSV, one pickle per
    remaining non-empty field, bzip2-compressed like the reference's."""
    with zipfile.ZipFile(fname, "w", zipfile.ZIP_BZIP2) as zf:
        if data.config is not None:
            import yaml

            zf.writestr("config.yml", yaml.safe_dump(data.config))
        if data.history is not None:
            keys = sorted({k for h in data.history for k in h})
            lines = [";".join(keys)] + [";".join(str(h.get(k, "")) for k in keys) for h in data.history]
            zf.writestr("history.csv", "\n".join(lines) + "\n")
        if data.summary is not None:
            zf.writestr("summary.csv", "\n".join(f"{k};{v}" for k, v in data.summary.items()))
        for f in fields(RunData):
            v = getattr(data, f.name)
            if v is not None and f.name not in ("config", "history", "summary"):
                zf.writestr(f.name + ".pkl", pickle.dumps(_as_numpy_tree(v)))


def params_from_run(fname: str, which: str = "params") -> Dict[str, Dict[str, np.ndarray]]:
    """The trainable parameters (`params` or `ema_params`) of a run file as a haiku dict of float32 arrays."""
    run = load_run(fname, parse_config=False)
    p = getattr(run, which)
    if p is None:
        raise KeyError(f"{fname} holds no {which}")
    return {str(mod): {str(k): np.asarray(v, dtype=np.float32) for k, v in leaves.items()} for mod, leaves in p.items()}
