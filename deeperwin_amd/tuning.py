"""Autotuner of launch shapes per electron-count bucket.

The reference leaves tiling to XLA.  Here every model exposes its launch-shape choices through
``dw_model_set_tuning`` (FP32 GEMM tiles of the value and Laplacian passes, the K-split of the per-walker
spin-mean GEMM of the value pass, electrons and output columns per workgroup of the fused Laplacian
GEMM).  The tuner times
the pass each choice affects on real inputs of one representative geometry per bucket and keeps the
fastest value of each field in turn (one coordinate-descent sweep); the table then applies to every model
whose electron count falls in that bucket.  Choices change launch shapes only: every candidate computes
the same sums (the K-split reorders one FP32 sum per element), so tuning never changes what is computed.
"""
from __future__ import annotations

import json
from dataclasses import dataclass, field
from pathlib import Path
from typing import Callable, Dict, List, Optional, Sequence, Tuple

BUCKETS: Tuple[Tuple[int, int], ...] = ((2, 8), (9, 16), (17, 24), (25, 32), (33, 42), (43, 63))

TILE_CHOICES = (0, 1, 2)          # 64x64, 64x128, 128x128
SPLIT_CHOICES = (1, 2, 4)


def bucket_of(n_el: int) -> Tuple[int, int]:
    for lo, hi in BUCKETS:
        if lo <= n_el <= hi:
            return (lo, hi)
    raise ValueError(f"n_el={n_el} outside the supported range 2..63")


def channel_rows(n_el: int) -> int:
    """Rows per electron in the Laplacian pass: value, 3 n tangents and the Laplacian, padded to 16."""
    c = 3 * n_el + 2
    return (c + 15) // 16 * 16


def lap_electron_choices(n_el: int) -> List[int]:
    """Electrons per workgroup of the fused Laplacian GEMM that keep a tile within 192 rows."""
    cp = channel_rows(n_el)
    return [e for e in range(1, 7) if e * cp <= 192]


def search_space(n_el: int) -> List[Tuple[str, Sequence[int], str]]:
    """(field, candidates, pass it affects) in the order the sweep visits them."""
    return [
        ("fwd_main_tile", TILE_CHOICES + (3, 4), "value"),
        ("fwd_mu_split", SPLIT_CHOICES, "value"),
        ("fwd_orb_tile", TILE_CHOICES, "value"),
        ("lap_cols", (128, 256), "lap"),
        ("lap_electrons", lap_electron_choices(n_el), "lap"),
        ("lap_mu_tile", TILE_CHOICES, "lap"),
        ("lap_orb_tile", TILE_CHOICES, "lap"),
    ]


def time_ms(fn: Callable[[], None], reps: int = 3) -> float:
    """Median duration of fn() in milliseconds, measured with events on the current stream."""
    import torch

    fn()  # warm-up (first use may size the workspace)
    times = []
    for _ in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        fn()
        b.record()
        b.synchronize()
        times.append(a.elapsed_time(b))
    times.sort()
    return times[len(times) // 2]


def autotune_model(model, r, R, Z, reps: int = 3, include_lap: bool = True) -> Tuple[Dict[str, int], List[dict]]:
    """One coordinate-descent sweep over the search space on the given walkers.  Leaves the best choices set
    on `model` and returns them with the timing log."""
    import torch

    B = r.shape[0]
    lp = torch.empty(B, device=r.device)
    e = torch.empty(B, device=r.device)
    passes = {"value": lambda: model.logpsi(r, R, Z, lp), "lap": lambda: model.local_energy(r, R, Z, e)}
    best = model.get_tuning()
    log = []
    for name, cands, which in search_space(model.shape.n_el):
        if which == "lap" and not include_lap:
            continue
        if name == "fwd_mu_split" and model.precision != "fp32":
            continue
        timings = {}
        for c in cands:
            model.set_tuning(**{**best, name: c})
            timings[c] = time_ms(passes[which], reps)
        choice = min(timings, key=timings.get)
        best[name] = choice
        model.set_tuning(**best)
        log.append({"field": name, "pass": which, "ms": {str(k): round(v, 4) for k, v in timings.items()},
                    "choice": choice})
    return best, log


@dataclass
class TuningTable:
    """Choices per electron-count bucket; buckets without an entry keep the built-in heuristics."""
    entries: Dict[Tuple[int, int], Dict[str, int]] = field(default_factory=dict)

    def lookup(self, n_el: int) -> Optional[Dict[str, int]]:
        return self.entries.get(bucket_of(n_el))

    def apply(self, model) -> None:
        choice = self.lookup(model.shape.n_el)
        if choice is None:
            return
        choice = dict(choice)
        # a bucket's Laplacian tile height may not fit a larger member of the bucket
        if choice.get("lap_electrons", 0) and choice["lap_electrons"] not in lap_electron_choices(model.shape.n_el):
            choice["lap_electrons"] = 0
        model.set_tuning(**choice)

    def to_json(self) -> str:
        return json.dumps({f"{lo}-{hi}": v for (lo, hi), v in sorted(self.entries.items())}, indent=1)

    @classmethod
    def from_json(cls, text: str) -> "TuningTable":
        raw = json.loads(text)
        out = cls()
        for k, v in raw.items():
            lo, hi = (int(x) for x in k.split("-"))
            if (lo, hi) not in BUCKETS:
                raise ValueError(f"unknown bucket {k}")
            out.entries[(lo, hi)] = {str(f): int(c) for f, c in v.items()}
        return out

    def save(self, path) -> None:
        Path(path).write_text(self.to_json())

    @classmethod
    def load(cls, path) -> "TuningTable":
        return cls.from_json(Path(path).read_text())


def tune_buckets(make_model: Callable[[int], Tuple[object, object, object, object]], n_els: Sequence[int],
                 reps: int = 3) -> Tuple[TuningTable, Dict[str, List[dict]]]:
    """Tune one representative electron count per bucket.  make_model(n_el) returns (model, r, R, Z) with
    walkers of the production batch size on the device."""
    table = TuningTable()
    logs = {}
    for n in n_els:
        model, r, R, Z = make_model(n)
        best, log = autotune_model(model, r, R, Z, reps)
        table.entries[bucket_of(n)] = best
        logs[f"n_el={n}"] = log
    return table, logs
