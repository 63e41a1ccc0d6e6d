"""Checkpoint import/export in the reference's run-file format (CPU; deeperwin_amd/checkpoints.py).

* a run file written by save_run loads back leaf for leaf, and its parameters drive the native model's
  parameter ingestion (host-side dw_model_set_param, no device needed);
* a jax.Array-shaped pickle (jax 0.4's reduction: _reconstruct_array(fun, args, arr_state, aval_state) with
  the NumPy array's own reduction inside) loads as the NumPy array;
* a pickle naming any other global (here os.system) is refused before anything from it runs.
"""
import io
import pickle
import sys
import types
import zipfile

import numpy as np
import pytest

from deeperwin_amd import checkpoints as ck
from deeperwin_amd.params import ModelShape, init_params


def test_round_trip_and_native_ingestion(tmp_path):
    from deeperwin_amd._native import NativeModel

    shape = ModelShape(14, 7, 2, 4)
    params = init_params(shape, 3)
    fname = str(tmp_path / "run.zip")
    ck.save_run(fname, ck.RunData(config={"model": {"name": "ferminet"}}, params=params,
                                  clipping_state=(0.0, 1e12), history=[{"E_mean": -109.1, "step": 0}],
                                  summary={"E_mean": -109.1}))
    with zipfile.ZipFile(fname) as zf:
        assert {"config.yml", "history.csv", "summary.csv", "params.pkl", "clipping_state.pkl"} <= set(zf.namelist())
    run = ck.load_run(fname)
    assert run.config == {"model": {"name": "ferminet"}}
    assert run.clipping_state == (0.0, 1e12)
    loaded = ck.params_from_run(fname)
    assert set(loaded) == set(params)
    for mod in params:
        for leaf in params[mod]:
            np.testing.assert_array_equal(loaded[mod][leaf], params[mod][leaf])
    m = NativeModel(shape)
    m.set_params(loaded)
    assert m.lib.dw_model_ready(m.handle) == 1


def test_jax_array_pickles_load_as_numpy():
    mod = types.ModuleType("jax._src.array")

    def _reconstruct_array(fun, args, arr_state, aval_state):  # pragma: no cover - never called here
        raise AssertionError("the allow-list answers this name with its own shim")

    _reconstruct_array.__module__ = "jax._src.array"
    _reconstruct_array.__qualname__ = "_reconstruct_array"
    mod._reconstruct_array = _reconstruct_array

    class FakeJaxArray:
        def __init__(self, value):
            self.value = value

        def __reduce__(self):
            fun, args, state = self.value.__reduce__()
            return (_reconstruct_array, (fun, args, state, {"weak_type": False}))

    saved = {k: sys.modules.get(k) for k in ("jax", "jax._src", "jax._src.array")}
    sys.modules.update({"jax": types.ModuleType("jax"), "jax._src": types.ModuleType("jax._src"),
                        "jax._src.array": mod})
    try:
        w = np.arange(12, dtype=np.float32).reshape(3, 4)
        blob = pickle.dumps({"wf/h_el_0/linear_0": {"w": FakeJaxArray(w)}})
    finally:
        for k, v in saved.items():
            if v is None:
                sys.modules.pop(k, None)
            else:
                sys.modules[k] = v
    out = ck.loads(blob)
    np.testing.assert_array_equal(out["wf/h_el_0/linear_0"]["w"], w)


def test_other_globals_are_refused():
    class Evil:
        def __reduce__(self):
            import os

            return (os.system, ("echo should-not-run",))

    blob = pickle.dumps({"params": Evil()})
    with pytest.raises(pickle.UnpicklingError, match="refusing"):
        ck.loads(blob)
