"""Tuner bookkeeping and the per-unit FLOP counts used for the batched-form roofline (CPU only)."""
import pytest

from deeperwin_amd.batching import forward_flops, laplacian_flops, unit_flops
from deeperwin_amd.params import ModelShape
from deeperwin_amd.tuning import BUCKETS, TuningTable, bucket_of, channel_rows, lap_electron_choices


def test_buckets_cover_every_supported_size():
    for n in range(2, 64):
        lo, hi = bucket_of(n)
        assert lo <= n <= hi
    with pytest.raises(ValueError):
        bucket_of(64)


def test_lap_tile_heights_fit():
    assert channel_rows(14) == 48 and lap_electron_choices(14) == [1, 2, 3, 4]
    assert channel_rows(60) == 192 and lap_electron_choices(60) == [1]
    assert channel_rows(2) == 16 and lap_electron_choices(2) == [1, 2, 3, 4, 5, 6]


def test_table_round_trip_and_clamp():
    t = TuningTable({(9, 16): {"fwd_main_tile": 2, "fwd_mu_split": 4, "lap_electrons": 4}})
    t2 = TuningTable.from_json(t.to_json())
    assert t2.entries == t.entries
    assert t2.lookup(12) == {"fwd_main_tile": 2, "fwd_mu_split": 4, "lap_electrons": 4}
    assert t2.lookup(30) is None

    class M:  # a bucket member too large for the bucket's tile height gets the heuristic instead
        shape = ModelShape(16, 8, 2, 4)
        got = None

        def set_tuning(self, **kw):
            M.got = kw

    t2.apply(M())
    assert M.got["lap_electrons"] == 0 and M.got["fwd_main_tile"] == 2
    with pytest.raises(ValueError):
        TuningTable.from_json('{"5-9": {}}')


def test_unit_flops_match_published_counts():
    # per-walker algorithmic FLOPs of the LiH and N2 workloads: F_fwd 5.28 MF / 20.61 MF,
    # F_lap 0.0735 GF / 0.879 GF, unit (10 / 20 steps + E_loc) 0.126 GF / 1.291 GF
    lih, n2 = ModelShape(4, 2, 2, 4), ModelShape(14, 7, 2, 16)
    assert forward_flops(lih) == pytest.approx(5.28e6, rel=2e-3)
    assert forward_flops(n2) == pytest.approx(20.61e6, rel=2e-3)
    assert laplacian_flops(lih) == pytest.approx(0.0735e9, rel=3e-3)
    assert laplacian_flops(n2) == pytest.approx(0.879e9, rel=2e-3)
    assert unit_flops(lih, 10) == pytest.approx(0.126e9, rel=5e-3)
    assert unit_flops(n2, 20) == pytest.approx(1.291e9, rel=2e-3)


def test_executed_flops_follow_the_row_split():
    """The library multiplies the spin-mean blocks once per walker instead of once per electron: per layer
    the executed count saves 2 (n - 1) (2 d) W1 FLOPs per value pass, and C = 3 n + 2 times that per
    Laplacian pass (the 1-el part); everything else is counted the same in the value pass."""
    from deeperwin_amd.batching import executed_forward_flops, executed_laplacian_flops

    for shape in (ModelShape(4, 2, 2, 4), ModelShape(14, 7, 2, 16), ModelShape(26, 13, 6, 16)):
        n, I, W1 = shape.n_el, shape.n_ion, shape.width1
        d = [4 * I] + [W1] * (shape.n_iter - 1)
        saved = sum(2 * (n - 1) * 2 * dl * W1 for dl in d)
        assert forward_flops(shape) - executed_forward_flops(shape) == pytest.approx(saved, rel=1e-12)
        assert executed_laplacian_flops(shape) < laplacian_flops(shape)
