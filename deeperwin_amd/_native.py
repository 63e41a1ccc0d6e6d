"""ctypes binding of ``libdwmi.so`` (the C ABI in ``include/dwmi.h``).

The library must exist (built by ``python -m deeperwin_amd.build``); there is no CPU fallback, so a
missing library raises immediately.  ``NativeModel`` owns one ``dw_model`` handle and moves torch
tensors across the ABI as raw device pointers on torch's current HIP stream.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path
from typing import Dict, Optional

import numpy as np

from .params import ModelShape

LIB_PATH = Path(os.environ.get("DWMI_LIB", Path(__file__).resolve().parent / "libdwmi.so"))

DW_FP32, DW_BF16X3, DW_BF16, DW_FP64 = 0, 1, 2, 3
PRECISIONS = {"fp32": DW_FP32, "bf16x3": DW_BF16X3, "bf16": DW_BF16, "fp64": DW_FP64}
FLAG_NONFINITE, FLAG_FP64_DET, FLAG_FORCED_ACCEPT = 1, 2, 4
TIMER_GEMM_1EL_LAP, TIMER_DET_LAP, TIMER_FORWARD, TIMER_LOCAL_ENERGY = 0, 1, 2, 3
PROPOSALS = {"normal": 0, "local": 1, "langevin": 2, "normal_one_el": 3, "local_one_el": 4}
ONE_EL_PROPOSALS = ("normal_one_el", "local_one_el")

# Exported symbols and their ctypes signatures (argtypes, restype).
_vp, _i32, _f32 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_float


class ModelDesc(ctypes.Structure):
    _fields_ = [("n_el", ctypes.c_int32), ("n_up", ctypes.c_int32), ("n_ion", ctypes.c_int32),
                ("n_det", ctypes.c_int32), ("width1", ctypes.c_int32), ("width2", ctypes.c_int32),
                ("n_iter", ctypes.c_int32), ("precision", ctypes.c_int32), ("fp64_cond_threshold", ctypes.c_float),
                ("embedding", ctypes.c_int32), ("z_min", ctypes.c_int32), ("n_ion_types", ctypes.c_int32)]


EMBEDDINGS = {"ferminet": 0, "dpe4": 1}


class Tuning(ctypes.Structure):
    _fields_ = [("fwd_main_tile", ctypes.c_int32), ("fwd_orb_tile", ctypes.c_int32),
                ("fwd_mu_split", ctypes.c_int32), ("lap_orb_tile", ctypes.c_int32),
                ("lap_mu_tile", ctypes.c_int32), ("lap_electrons", ctypes.c_int32), ("lap_cols", ctypes.c_int32)]


TUNING_FIELDS = [f for f, _ in Tuning._fields_]
TUNING_DEFAULT = {"fwd_main_tile": -1, "fwd_orb_tile": -1, "fwd_mu_split": 0, "lap_orb_tile": -1,
                  "lap_mu_tile": -1, "lap_electrons": 0, "lap_cols": 0}


SIGNATURES = {
    "dw_model_create": ([ctypes.POINTER(ModelDesc), ctypes.POINTER(_vp)], ctypes.c_int),
    "dw_model_destroy": ([_vp], ctypes.c_int),
    "dw_model_set_param": ([_vp, ctypes.c_char_p, ctypes.c_char_p, _vp, ctypes.POINTER(ctypes.c_int64), _i32],
                           ctypes.c_int),
    "dw_model_ready": ([_vp], ctypes.c_int32),
    "dw_logpsi": ([_vp, _vp, _i32, _vp, _vp, _vp, _vp, _vp], ctypes.c_int),
    "dw_local_energy": ([_vp, _vp, _i32, _vp, _vp, _vp, _vp, _vp, _vp], ctypes.c_int),
    "dw_mcmc_steps": ([_vp, _vp, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _i32, _vp, _i32, _i32, _f32, _f32, _f32, _vp,
                       _vp, _vp], ctypes.c_int),
    "dw_mcmc_steps_noalloc": ([_vp, _vp, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _i32, _vp, _i32, _i32, _f32, _f32, _f32,
                               _vp, _vp, _vp], ctypes.c_int),
    "dw_model_set_proposal": ([_vp, _i32, _f32, _f32, _f32], ctypes.c_int),
    "dw_mcmc_propose": ([_vp, _vp, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp], ctypes.c_int),
    "dw_energy_moments": ([_vp, _i32, _vp, _vp], ctypes.c_int),
    "dw_threefry_split": ([_vp, _i32, _vp, _vp], ctypes.c_int),
    "dw_threefry_normal": ([_vp, ctypes.c_int64, _vp, _vp], ctypes.c_int),
    "dw_threefry_mcmc_draws": ([_vp, _vp, _i32, _i32, _i32, _vp, _vp, _vp], ctypes.c_int),
    "dw_param_grad": ([_vp, _vp, _i32, _vp, _vp, _vp, _vp, _vp, _vp], ctypes.c_int),
    "dw_param_count": ([_vp], ctypes.c_int64),
    "dw_clip_energies": ([_vp, _i32, _i32, _i32, _i32, _f32, _i32, _vp, _vp, _vp, _vp, _vp], ctypes.c_int),
    "dw_param_layout": ([_vp, _i32, ctypes.c_char_p, _i32, ctypes.POINTER(ctypes.c_int64),
                         ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
    "dw_workspace_size": ([_vp, _i32, _i32], ctypes.c_size_t),
    "dw_model_set_chunk": ([_vp, _i32], ctypes.c_int),
    "dw_model_set_workspace": ([_vp, _vp, ctypes.c_size_t], ctypes.c_int),
    "dw_model_set_tuning": ([_vp, ctypes.POINTER(Tuning)], ctypes.c_int),
    "dw_model_get_tuning": ([_vp, ctypes.POINTER(Tuning)], ctypes.c_int),
    "dw_model_conditioning": ([_vp, _vp, _i32, _vp], ctypes.c_int),
    "dw_model_set_timing": ([_vp, _i32], ctypes.c_int),
    "dw_model_timer_query": ([_vp, _i32, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int32)],
                             ctypes.c_int),
    "dw_last_error": ([], ctypes.c_char_p),
    "dw_version": ([], ctypes.c_char_p),
}

_lib: Optional[ctypes.CDLL] = None


def load_library() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m deeperwin_amd.build` "
                           "(the native path has no CPU fallback)")
    lib = ctypes.CDLL(str(LIB_PATH))
    for name, (argtypes, restype) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = restype
    _lib = lib
    return lib


class DwError(RuntimeError):
    pass


def _check(status: int) -> None:
    if status != 0:
        msg = load_library().dw_last_error().decode()
        raise DwError(f"dwmi error {status}: {msg}")


def _ptr(t) -> int:
    return 0 if t is None else int(t.data_ptr())


def _stream() -> int:
    import torch

    return int(torch.cuda.current_stream().cuda_stream)


class NativeModel:
    """One compiled-shape wavefunction on the current GPU."""

    def __init__(self, shape: ModelShape, precision: str = "fp32", fp64_cond_threshold: float = 0.0,
                 chunk: int = 0):
        self.lib = load_library()
        self.shape = shape
        self.precision = precision
        desc = ModelDesc(shape.n_el, shape.n_up, shape.n_ion, shape.n_det, shape.width1, shape.width2,
                         shape.n_iter, PRECISIONS[precision], float(fp64_cond_threshold),
                         EMBEDDINGS[shape.embedding], int(shape.z_min), int(shape.n_ion_types))
        h = ctypes.c_void_p()
        _check(self.lib.dw_model_create(ctypes.byref(desc), ctypes.byref(h)))
        self.handle = h
        self._param_token = None
        if chunk:
            _check(self.lib.dw_model_set_chunk(self.handle, int(chunk)))

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                self.lib.dw_model_destroy(h)
            except Exception:
                pass
            self.handle = None

    @staticmethod
    def _leaf_token(v):
        # numpy arrays are compared by content (they may be mutated in place); torch tensors by identity and
        # their in-place version counter
        if hasattr(v, "_version"):
            return ("t", v._version)
        a = np.asarray(v)
        return ("a", a.shape, a.dtype.str, hash(np.ascontiguousarray(a).tobytes()))

    def set_params(self, params: Dict) -> None:
        """Upload the leaves of `params` unless they are the very objects of the last upload, unchanged.

        The model keeps references to the uploaded leaves, so an identity match cannot come from a freed
        and reused object; torch tensors also carry their version counter (in-place updates) and numpy
        arrays a content hash.  `invalidate()` forces the next call to upload."""
        leaves_now = [(p, k, v) for p, leaves in params.items() for k, v in leaves.items()]
        last = self._param_token
        if last is not None and len(last) == len(leaves_now) and all(
                p == lp and k == lk and v is lv and self._leaf_token(v) == lt
                for (p, k, v), (lp, lk, lv, lt) in zip(leaves_now, last)):
            return
        for path, leaves in params.items():
            for leaf, val in leaves.items():
                arr = val.detach().cpu().numpy() if hasattr(val, "detach") else np.asarray(val)
                arr = np.ascontiguousarray(arr, dtype=np.float32)
                shp = (ctypes.c_int64 * arr.ndim)(*arr.shape)
                _check(self.lib.dw_model_set_param(self.handle, path.encode(), leaf.encode(),
                                                   arr.ctypes.data_as(ctypes.c_void_p), shp, arr.ndim))
        if not self.lib.dw_model_ready(self.handle):
            raise DwError("not every parameter leaf of the wavefunction was provided")
        self._param_token = [(p, k, v, self._leaf_token(v)) for p, k, v in leaves_now]

    def invalidate(self) -> None:
        """Forget the last upload: the next set_params uploads every leaf."""
        self._param_token = None

    def set_tuning(self, **choices) -> None:
        """Launch-shape choices (fields of dw_tuning); unnamed fields keep their current value."""
        cur = self.get_tuning()
        unknown = set(choices) - set(TUNING_FIELDS)
        if unknown:
            raise KeyError(f"unknown tuning fields {sorted(unknown)}")
        cur.update(choices)
        t = Tuning(*[int(cur[f]) for f in TUNING_FIELDS])
        _check(self.lib.dw_model_set_tuning(self.handle, ctypes.byref(t)))

    def get_tuning(self) -> Dict[str, int]:
        t = Tuning()
        _check(self.lib.dw_model_get_tuning(self.handle, ctypes.byref(t)))
        return {f: int(getattr(t, f)) for f in TUNING_FIELDS}

    def set_workspace(self, buf) -> None:
        """Carve the activation buffers from `buf` (a CUDA uint8 tensor of >= workspace_bytes, kept alive by
        the caller), or return to library-owned buffers with None."""
        if buf is None:
            _check(self.lib.dw_model_set_workspace(self.handle, None, 0))
            self._ws = None
            return
        _check(self.lib.dw_model_set_workspace(self.handle, ctypes.c_void_p(buf.data_ptr()), buf.numel() * buf.element_size()))
        self._ws = buf

    def set_chunk(self, walkers: int) -> None:
        _check(self.lib.dw_model_set_chunk(self.handle, int(walkers)))

    def set_timing(self, enable: bool) -> None:
        _check(self.lib.dw_model_set_timing(self.handle, 1 if enable else 0))

    def timer(self, kind: int):
        """(total milliseconds, launches) of one timed region kind since timing was enabled."""
        ms = ctypes.c_double()
        cnt = ctypes.c_int32()
        _check(self.lib.dw_model_timer_query(self.handle, int(kind), ctypes.byref(ms), ctypes.byref(cnt)))
        return ms.value, cnt.value

    def conditioning(self, n_walkers: int):
        """Per-walker conditioning estimate of the last pass, as a torch CUDA tensor."""
        import torch

        out = torch.empty(n_walkers, device="cuda", dtype=torch.float32)
        _check(self.lib.dw_model_conditioning(self.handle, _ptr(out), int(n_walkers), _stream()))
        return out

    def param_layout(self):
        """[(module suffix, leaf, offset, shape)] of the packed parameter gradient, in order."""
        out = []
        key = ctypes.create_string_buffer(128)
        off = ctypes.c_int64()
        shp = (ctypes.c_int64 * 2)()
        idx = 0
        while self.lib.dw_param_layout(self.handle, idx, key, 128, ctypes.byref(off), shp) == 0:
            mod, leaf = key.value.decode().split(":")
            shape = (int(shp[0]),) if leaf == "b" else (int(shp[0]), int(shp[1]))
            out.append((mod, leaf, int(off.value), shape))
            idx += 1
        return out

    def param_grad(self, r, R, Z, walker_weights, out_lp=None):
        """Packed G = sum_b c_b d log|psi|^2(r_b) / d theta as a CUDA float32 tensor (dw_param_grad)."""
        import torch

        g = torch.empty(int(self.lib.dw_param_count(self.handle)), device="cuda", dtype=torch.float32)
        c = walker_weights.to(torch.float32).contiguous()
        _check(self.lib.dw_param_grad(self.handle, _ptr(r), r.shape[0], _ptr(R), _ptr(Z), _ptr(c), _ptr(g),
                                      _ptr(out_lp), _stream()))
        return g

    def unpack_grad(self, g, params: Dict) -> Dict:
        """The packed gradient as a dict shaped like `params` (module paths matched by suffix)."""
        out = {}
        layout = {(mod, leaf): (off, shape) for mod, leaf, off, shape in self.param_layout()}
        for path, leaves in params.items():
            norm = "/".join(seg for seg in path.split("/") if seg != "~")
            out[path] = {}
            for leaf in leaves:
                hit = [k for k in layout if k[1] == leaf and (norm == k[0] or norm.endswith("/" + k[0]))]
                if len(hit) != 1:
                    raise KeyError(f"{path}/{leaf}: {len(hit)} gradient leaves match")
                off, shape = layout[hit[0]]
                cnt = int(np.prod(shape))
                out[path][leaf] = g[off:off + cnt].reshape(shape)
        return out

    def workspace_bytes(self, n_walkers: int, mode: int) -> int:
        return int(self.lib.dw_workspace_size(self.handle, int(n_walkers), int(mode)))

    # ---- compute entry points (torch CUDA tensors, float32, contiguous) ----
    def logpsi(self, r, R, Z, out_lp, out_phase=None) -> None:
        _check(self.lib.dw_logpsi(self.handle, _ptr(r), r.shape[0], _ptr(R), _ptr(Z), _ptr(out_lp),
                                  _ptr(out_phase), _stream()))

    def local_energy(self, r, R, Z, out_eloc, out_lp=None, out_flags=None) -> None:
        _check(self.lib.dw_local_energy(self.handle, _ptr(r), r.shape[0], _ptr(R), _ptr(Z), _ptr(out_eloc),
                                        _ptr(out_lp), _ptr(out_flags), _stream()))

    def set_proposal(self, name: str = "normal", r_min: float = 0.1, r_max: float = 1.0,
                     langevin_scale: float = 1.0) -> None:
        """dw_model_set_proposal with a proposal name of MCMCConfig ('cauchy' is the normal kernel fed
        Cauchy draws)."""
        if name == "cauchy":
            name = "normal"
        if name not in PROPOSALS:
            raise ValueError(f"unknown proposal {name!r}; the device sampler has {sorted(PROPOSALS)}")
        _check(self.lib.dw_model_set_proposal(self.handle, PROPOSALS[name], float(r_min), float(r_max),
                                              float(langevin_scale)))

    def propose(self, r, R, Z, normals, chain, out_r, out_log_q) -> None:
        _check(self.lib.dw_mcmc_propose(self.handle, _ptr(r), r.shape[0], _ptr(R), _ptr(Z), _ptr(normals),
                                        _ptr(chain), _ptr(out_r), _ptr(out_log_q), _stream()))

    def mcmc_steps(self, r, R, Z, lp, age, normals, uniforms, chain, max_age, interval, target, lo, hi,
                   out_acc=None, out_flags=None, noalloc: bool = False) -> None:
        """dw_mcmc_steps, or with noalloc=True dw_mcmc_steps_noalloc (no device allocation: every buffer of
        the call comes from the library's fixed scratch array)."""
        n_steps = 0 if normals is None else normals.shape[0]
        fn = self.lib.dw_mcmc_steps_noalloc if noalloc else self.lib.dw_mcmc_steps
        _check(fn(self.handle, _ptr(r), r.shape[0], _ptr(R), _ptr(Z), _ptr(lp), _ptr(age),
                                      _ptr(normals), _ptr(uniforms), n_steps, _ptr(chain), int(max_age),
                                      int(interval), float(target), float(lo), float(hi), _ptr(out_acc),
                                      _ptr(out_flags), _stream()))


def energy_moments(e_loc):
    """(E_mean, E_var) of a CUDA float32 tensor of local energies as a CUDA float64 tensor of 2 (device
    side, ordered on the current stream): nanmean and nanmean of squared deviations
    (optimization/evaluation.py:143-151)."""
    import torch

    out = torch.empty(2, device=e_loc.device, dtype=torch.float64)
    e = e_loc.reshape(-1).to(torch.float32).contiguous()
    _check(load_library().dw_energy_moments(_ptr(e), int(e.numel()), _ptr(out), _stream()))
    return out


def clip_energies(e_loc, state=None, name: str = "tanh", center: str = "mean", width_metric: str = "std",
                  clip_by: float = 5.0, from_previous_step: bool = True):
    """dw_clip_energies on a CUDA tensor of local energies: (per-walker gradient weights, clipped energies,
    stats [loss, E_mean, E_var, E_var_clipped, next center, next width, center used, width used]) as CUDA
    tensors.  `state` is a CUDA float64 tensor (center, width) or None."""
    import torch

    e = e_loc.reshape(-1).to(torch.float32).contiguous()
    w = torch.empty_like(e)
    ec = torch.empty_like(e)
    stats = torch.empty(8, device=e.device, dtype=torch.float64)
    st = None if state is None else state.to(device=e.device, dtype=torch.float64).contiguous()
    _check(load_library().dw_clip_energies(_ptr(e), int(e.numel()), 1 if name == "tanh" else 0,
                                           1 if center == "median" else 0, 1 if width_metric == "mae" else 0,
                                           float(clip_by), 1 if from_previous_step else 0, _ptr(st), _ptr(w), _ptr(ec),
                                           _ptr(stats), _stream()))
    return w, ec, stats
