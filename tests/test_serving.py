"""Serving form: request generation and summaries (CPU) and continuous batching on the GPU (alone vs mix)."""
import numpy as np
import pytest

from deeperwin_amd.serving import Request, Result, poisson_stream, summarize


def test_poisson_stream_shape():
    reqs = poisson_stream(200, rate=50.0, seed=1)
    arr = np.array([q.arrival for q in reqs])
    assert np.all(np.diff(arr) > 0)
    assert abs(len(reqs) / arr[-1] - 50.0) < 12.0          # rate within a few sigma
    n = np.array([q.n_el for q in reqs])
    assert n.min() >= 2 and n.max() <= 60
    B = np.array([q.n_walkers for q in reqs])
    assert B.min() >= 256 - 32 and B.max() <= 4096 + 32 and np.all(B % 64 == 0)
    # log-uniform walker counts: the median sits near the geometric mean of the range
    assert 600 < np.median(B) < 1700


def test_summary_percentiles():
    res = [Result(k, 0.0, 0.0, 0.5, arrival=0.0, admitted=0.0, finished=float(k + 1)) for k in range(100)]
    s = summarize(res, isolated={k: 1.0 for k in range(100)})
    assert abs(s["p50_latency_s"] - 50.5) < 1e-9 and abs(s["p99_latency_s"] - 99.01) < 1e-9
    assert abs(s["p99_over_isolated_p99"] - 99.01) < 1e-9


@pytest.mark.gpu
def test_results_are_identical_alone_and_in_the_mix(cuda):
    from deeperwin_amd.serving import Server, SharedWeights

    w = SharedWeights(n_el_max=16, n_ion_max=8, n_det=4, seed=3, chunk=64)
    reqs = poisson_stream(6, rate=500.0, seed=7, n_el_range=(2, 16), walker_range=(64, 256), n_steps=6)
    mixed = Server(w, max_active=4, slice_steps=2).serve(reqs)
    alone = [Server(w, max_active=1, slice_steps=6).serve([q])[0] for q in reqs]
    for a, b in zip(alone, mixed):
        assert a.rid == b.rid
        assert a.energy_mean == b.energy_mean and a.energy_var == b.energy_var, (a, b)
        assert a.acceptance == b.acceptance
    assert all(np.isfinite(r.energy_mean) for r in mixed)
