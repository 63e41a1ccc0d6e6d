"""Pins for the oracle's local-stepsize and approximate-Langevin proposals (oracle/mcmc.py:propose, restating
mcmc.py:204-236, 255-284), CPU only:

* log q equals log N(r; r' + g' s'^2, s'^2) - log N(r'; r + g s^2, s^2), the log ratio of the reverse to the
  forward proposal density, with the densities from scipy.stats.norm and s, g from a scalar loop;
* the drift is minus the gradient of langevin_scale * sum_I Z_I |r - R_I| (central differences);
* with r_min = r_max and no drift the proposal is the symmetric Gaussian one (log q = 0 exactly);
* a Metropolis-Hastings chain with either proposal samples the hydrogen 1s density, and the same chain with
  log q dropped does not (the ratio is needed and the test can see its absence);
* planted mistakes in the step scale or the drift fail the density check.
"""
import numpy as np
import pytest
from scipy.stats import norm

from oracle import mcmc as om

KW = {"local": dict(r_min=0.1, r_max=1.0), "langevin": dict(r_min=0.2, r_max=2.0, langevin_scale=1.3)}


def _system(seed=0, B=48, n=5):
    rng = np.random.default_rng(seed)
    R = rng.uniform(-2.0, 2.0, (3, 3))
    Z = np.array([1.0, 3.0, 7.0])
    r = R[rng.integers(0, 3, (B, n))] + rng.standard_normal((B, n, 3)) * 0.8
    xi = rng.standard_normal((B, n, 3))
    return r, R, Z, xi


def _scale_drift_loop(x, R, Z, stepsize, r_min, r_max, scale):
    """s and g of one electron position x by a loop over the nuclei."""
    dmin, g = np.inf, np.zeros(3)
    for k in range(len(R)):
        d = np.sqrt(sum((x[c] - R[k][c]) ** 2 for c in range(3)))
        dmin = min(dmin, d)
        for c in range(3):
            g[c] -= scale * Z[k] * (x[c] - R[k][c]) / d
    return stepsize * min(max(dmin, r_min), r_max), g


def _density_log_q(kind, r, r_new, R, Z, stepsize, r_min, r_max, scale=1.0):
    """Sum over electrons of log N_reverse - log N_forward, from scipy.stats.norm per coordinate."""
    scale = scale if kind == "langevin" else 0.0
    out = np.zeros(r.shape[0])
    for b in range(r.shape[0]):
        for i in range(r.shape[1]):
            s, g = _scale_drift_loop(r[b, i], R, Z, stepsize, r_min, r_max, scale)
            s1, g1 = _scale_drift_loop(r_new[b, i], R, Z, stepsize, r_min, r_max, scale)
            fwd = norm.logpdf(r_new[b, i], loc=r[b, i] + g * s * s, scale=s).sum()
            rev = norm.logpdf(r[b, i], loc=r_new[b, i] + g1 * s1 * s1, scale=s1).sum()
            out[b] += rev - fwd
    return out


@pytest.mark.parametrize("kind", ["local", "langevin"])
@pytest.mark.parametrize("stepsize", [0.02, 0.4])
def test_log_q_is_reverse_over_forward_density(kind, stepsize):
    r, R, Z, xi = _system(1)
    kw = KW[kind]
    r_new, lq = om.propose(kind, r, R, Z, xi, stepsize, dtype=np.float64, **kw)
    # the move itself: r' - r - g s^2 = s xi per electron
    for b in range(3):
        for i in range(r.shape[1]):
            s, g = _scale_drift_loop(r[b, i], R, Z, stepsize, kw["r_min"], kw["r_max"],
                                     kw.get("langevin_scale", 0.0) if kind == "langevin" else 0.0)
            np.testing.assert_allclose(r_new[b, i] - r[b, i] - g * s * s, s * xi[b, i], rtol=1e-9, atol=1e-12)
    want = _density_log_q(kind, r, r_new, R, Z, stepsize, kw["r_min"], kw["r_max"], kw.get("langevin_scale", 1.0))
    np.testing.assert_allclose(lq, want, rtol=1e-9, atol=1e-9)
    assert np.abs(want).max() > 0.1  # the ratio is not trivially zero in this case


def test_drift_is_minus_gradient_of_nuclear_distance_sum():
    r, R, Z, _ = _system(2, B=4, n=3)
    scale = 1.3
    _, g = om.scale_and_drift(r, R, Z, 0.1, 0.1, 1.0, scale, dtype=np.float64)

    def phi(x):
        return scale * sum(Z[k] * np.linalg.norm(x - R[k]) for k in range(len(R)))

    h = 1e-6
    for b in range(r.shape[0]):
        for i in range(r.shape[1]):
            grad = np.zeros(3)
            for c in range(3):
                e = np.zeros(3)
                e[c] = h
                grad[c] = (phi(r[b, i] + e) - phi(r[b, i] - e)) / (2 * h)
            np.testing.assert_allclose(g[b, i], -grad, rtol=1e-6, atol=1e-7)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_constant_scale_without_drift_is_symmetric(dtype):
    r, R, Z, xi = _system(3)
    for kind in ("local", "langevin"):
        r_new, lq = om.propose(kind, r, R, Z, xi, 0.3, r_min=0.5, r_max=0.5, langevin_scale=0.0, dtype=dtype)
        assert np.all(lq == 0)
        tol = 1e-6 if dtype == np.float32 else 1e-14
        np.testing.assert_allclose(r_new, r + xi * dtype(0.3 * 0.5), rtol=tol, atol=tol)


@pytest.mark.parametrize("kind", ["local", "langevin"])
def test_stationary_distribution_hydrogen_1s(kind):
    # |psi|^2 = exp(-2 r), nucleus Z = 1 at the origin: <r> = 3/2, <r^2> = 3.  The walkers start far from
    # equilibrium (unit Gaussian); the local scale and the drift make the proposal asymmetric, so the
    # stationary law is right only with log q in the accept test.
    f = lambda r: -2.0 * np.linalg.norm(r[:, 0, :], axis=-1)
    R, Z = np.zeros((1, 3)), np.ones(1)
    kw = dict(KW[kind])
    if kind == "langevin":
        kw["langevin_scale"] = 1.0
    B, S = 6000, 400

    def run(drop_log_q):
        rng = np.random.default_rng(0)
        st = om.ChainState(r=rng.standard_normal((B, 1, 3)), log_psi_sqr=np.full(B, -1000.0),
                           walker_age=np.zeros(B, np.int64), stepsize=0.7)
        normals = rng.standard_normal((S, B, 1, 3))
        uniforms = rng.uniform(size=(S, B))

        def prop(r, xi, step, step_nr):
            r_new, lq = om.propose(kind, r, R, Z, xi, step, dtype=np.float64, **kw)
            return r_new, 0.0 * lq if drop_log_q else lq

        out, _, _ = om.mcmc_steps(f, st, normals, uniforms, max_age=1000, update_interval=10 ** 9, dtype=np.float64,
                                  proposal=prop)
        return np.linalg.norm(out.r[:, 0], axis=-1)

    band = 5 * np.sqrt(0.75 / B)
    rr = run(False)
    assert abs(rr.mean() - 1.5) < band
    assert abs((rr ** 2).mean() - 3.0) < 5 * np.sqrt(((rr ** 4).mean() - 9.0) / B)
    assert abs(run(True).mean() - 1.5) > band  # without log q the chain samples another law


@pytest.mark.parametrize("mistake", ["drift_sign", "farthest_nucleus", "constant_scale"])
def test_planted_mistakes_fail_the_density_check(mistake, monkeypatch):
    r, R, Z, xi = _system(4)
    kind = "langevin"
    kw = KW[kind]
    true_fn = om.scale_and_drift

    def mutated(r_, R_, Z_, stepsize, r_min, r_max, langevin_scale=1.0, dtype=np.float32):
        s, g = true_fn(r_, R_, Z_, stepsize, r_min, r_max, langevin_scale, dtype)
        if mistake == "drift_sign":
            g = -g
        elif mistake == "farthest_nucleus":
            d = np.linalg.norm(np.asarray(r_, dtype)[..., :, None, :] - R_, axis=-1)
            s = dtype(stepsize) * np.clip(d.max(axis=-1), r_min, r_max)
        else:
            s = np.full_like(s, dtype(stepsize) * dtype(r_min))
        return s, g

    monkeypatch.setattr(om, "scale_and_drift", mutated)
    r_new, lq = om.propose(kind, r, R, Z, xi, 0.4, dtype=np.float64, **kw)
    want = _density_log_q(kind, r, r_new, R, Z, 0.4, kw["r_min"], kw["r_max"], kw["langevin_scale"])
    assert np.abs(lq - want).max() > 1e-3


def test_one_electron_proposals():
    """normal_one_el / local_one_el (mcmc.py:183-192, 239-252): only electron step_nr mod n moves; the local
    form's log q is the reverse/forward density ratio of that electron alone (scipy.stats.norm)."""
    r, R, Z, xi_all = _system(5, B=16, n=4)
    xi = xi_all[:, 0]
    for step_nr in (0, 5):
        k = step_nr % 4
        r1, lq1 = om.propose("normal_one_el", r, R, Z, xi, 0.3, dtype=np.float64, step_nr=step_nr)
        assert np.all(lq1 == 0)
        np.testing.assert_array_equal(np.delete(r1, k, axis=1), np.delete(r, k, axis=1))
        np.testing.assert_allclose(r1[:, k], r[:, k] + 0.3 * xi, rtol=1e-12)
        r2, lq2 = om.propose("local_one_el", r, R, Z, xi, 0.3, r_min=0.1, r_max=1.0, dtype=np.float64,
                             step_nr=step_nr)
        np.testing.assert_array_equal(np.delete(r2, k, axis=1), np.delete(r, k, axis=1))
        for b in range(r.shape[0]):
            s, _ = _scale_drift_loop(r[b, k], R, Z, 0.3, 0.1, 1.0, 0.0)
            s1, _ = _scale_drift_loop(r2[b, k], R, Z, 0.3, 0.1, 1.0, 0.0)
            np.testing.assert_allclose(r2[b, k], r[b, k] + s * xi[b], rtol=1e-12)
            want = norm.logpdf(r[b, k], r2[b, k], s1).sum() - norm.logpdf(r2[b, k], r[b, k], s).sum()
            assert abs(lq2[b] - want) <= 1e-9 * max(1.0, abs(want))


def test_one_electron_chain_samples_two_independent_electrons():
    """A chain of local_one_el moves (electron step_nr mod 2 each step) on |psi|^2 = exp(-2|r_1| - 2|r_2|)
    samples <|r_i|> = 3/2 for both electrons."""
    f = lambda r: -2.0 * np.linalg.norm(r, axis=-1).sum(axis=-1)
    R, Z = np.zeros((1, 3)), np.ones(1)
    B, S = 4000, 600
    rng = np.random.default_rng(3)
    st = om.ChainState(r=rng.standard_normal((B, 2, 3)), log_psi_sqr=np.full(B, -1000.0),
                       walker_age=np.zeros(B, np.int64), stepsize=0.7)
    normals = rng.standard_normal((S, B, 3))
    uniforms = rng.uniform(size=(S, B))

    def prop(r, x, step, step_nr):
        return om.propose("local_one_el", r, R, Z, x, step, r_min=0.1, r_max=1.0, dtype=np.float64, step_nr=step_nr)

    out, _, _ = om.mcmc_steps(f, st, normals, uniforms, max_age=1000, update_interval=10 ** 9, dtype=np.float64,
                              proposal=prop)
    rr = np.linalg.norm(out.r, axis=-1)
    for e in range(2):
        assert abs(rr[:, e].mean() - 1.5) < 5 * np.sqrt(0.75 / B)
