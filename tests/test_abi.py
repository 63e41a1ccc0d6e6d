"""The C-ABI library (CPU only: no compute call is made here).

Checks that libdwmi.so loads, exports every function declared in include/dwmi.h, validates model
descriptors and parameter leaves (names matched by suffix, shapes checked) and reports errors through
dw_last_error — all host-side paths that touch no device memory.
"""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from deeperwin_amd import _native
from deeperwin_amd.params import ModelShape, init_params

ROOT = Path(__file__).resolve().parents[1]


def _declared():
    text = (ROOT / "include" / "dwmi.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:[\w\s\*]+?)\b(dw_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = _native.load_library()
    names = _declared()
    assert len(names) >= 10
    for name in names:
        assert hasattr(lib, name), name
    assert set(_native.SIGNATURES) == set(names)
    assert lib.dw_version().decode().startswith("dwmi")


def _desc(**kw):
    d = dict(n_el=14, n_up=7, n_ion=2, n_det=16, width1=256, width2=32, n_iter=4, precision=0,
             fp64_cond_threshold=0.0)
    d.update(kw)
    return _native.ModelDesc(**d)


def test_model_create_validates_descriptor():
    lib = _native.load_library()
    h = ctypes.c_void_p()
    assert lib.dw_model_create(ctypes.byref(_desc()), ctypes.byref(h)) == 0
    assert lib.dw_model_ready(h) == 0
    assert lib.dw_model_destroy(h) == 0
    for bad in (dict(n_el=65), dict(n_up=0), dict(n_up=14), dict(width2=16), dict(width1=200), dict(n_iter=5),
                dict(precision=7), dict(embedding=2), dict(embedding=1, n_ion_types=0)):
        h = ctypes.c_void_p()
        assert lib.dw_model_create(ctypes.byref(_desc(**bad)), ctypes.byref(h)) == 1, bad
        assert lib.dw_last_error().decode()


def test_parameter_ingestion_and_errors():
    shape = ModelShape(14, 7, 2, 16)
    m = _native.NativeModel(shape)
    params = init_params(shape, 1)
    # tilde-segmented haiku paths are accepted
    tilde = {k.replace("wf/", "wf/~/"): v for k, v in params.items()}
    m.set_params(tilde)
    assert m.lib.dw_model_ready(m.handle) == 1
    # wrong shape -> error with a message naming the leaf
    bad = np.zeros((10, 256), np.float32)
    shp = (ctypes.c_int64 * 2)(*bad.shape)
    st = m.lib.dw_model_set_param(m.handle, b"wf/ferminet_embedding/h_el_0/linear_0", b"w",
                                  bad.ctypes.data_as(ctypes.c_void_p), shp, 2)
    assert st == 1 and "h_el_0" in m.lib.dw_last_error().decode()
    st = m.lib.dw_model_set_param(m.handle, b"wf/unknown/linear_0", b"w", bad.ctypes.data_as(ctypes.c_void_p), shp, 2)
    assert st == 1 and "unknown" in m.lib.dw_last_error().decode()
    with pytest.raises(_native.DwError):
        m.set_params({"wf/ferminet_embedding/h_el_0/linear_0": {"w": bad}})


def test_dpe4_parameter_ingestion():
    """The dpe4 model takes the reference's dpe4 leaves (conv maps, same/diff and el-ion streams, the ion
    table) and refuses the FermiNet 2-el leaves."""
    shape = ModelShape(14, 7, 2, 4, embedding="dpe4", z_min=7, n_ion_types=1)
    m = _native.NativeModel(shape)
    params = init_params(shape, 3)
    m.set_params({k.replace("wf/", "wf/~/"): v for k, v in params.items()})
    assert m.lib.dw_model_ready(m.handle) == 1
    w = np.zeros((4, 32), np.float32)
    shp = (ctypes.c_int64 * 2)(*w.shape)
    st = m.lib.dw_model_set_param(m.handle, b"wf/ferminet_embedding/h_el_el_0/linear_0", b"w",
                                  w.ctypes.data_as(ctypes.c_void_p), shp, 2)
    assert st == 1 and "unknown" in m.lib.dw_last_error().decode()
    m2 = _native.NativeModel(shape)
    incomplete = dict(params)
    incomplete.pop("wf/input/h_ion")
    with pytest.raises(_native.DwError, match="not every parameter"):
        m2.set_params(incomplete)


def test_missing_leaf_is_reported():
    shape = ModelShape(4, 2, 2, 4)
    m = _native.NativeModel(shape)
    params = init_params(shape, 0)
    params.pop("wf/orbitals/envelope_orbitals/bf_dn/linear_0")
    with pytest.raises(_native.DwError, match="not every parameter"):
        m.set_params(params)


def test_workspace_size_formula():
    shape = ModelShape(14, 7, 2, 16)
    m = _native.NativeModel(shape)
    v = m.workspace_bytes(2048, 0)
    lap = m.workspace_bytes(2048, 1)
    assert 0 < v < lap
    m.set_chunk(512)
    assert m.workspace_bytes(2048, 1) < lap


def test_noalloc_mcmc_rejects_before_touching_the_device():
    """dw_mcmc_steps_noalloc validates sizes and weight state on the host first: too many walkers for its
    fixed scratch array -> DW_ERR_INVALID naming the limit; weights not uploaded yet -> DW_ERR_STATE.
    The pointers below are never dereferenced."""
    shape = ModelShape(14, 7, 2, 16)
    m = _native.NativeModel(shape)
    m.set_params(init_params(shape, 1))
    fake = ctypes.c_void_p(256)
    lib = m.lib

    def call(B):
        return lib.dw_mcmc_steps_noalloc(m.handle, fake, B, fake, fake, fake, fake, None, None, 0, fake, 100, 100,
                                         0.5, 0.01, 1.0, None, None, None)

    assert call(1 << 20) == 1
    msg = lib.dw_last_error().decode()
    assert "walkers" in msg and "fit" in msg
    assert call(64) == 3
    assert "not on the device" in lib.dw_last_error().decode()
