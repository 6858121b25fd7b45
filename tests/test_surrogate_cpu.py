"""Surrogate and MCMC host logic on the CPU: the FP64 MLP oracle pinned by
finite differences, the vectorised Barlow-Beeston used by the surrogate
stage pinned against the oracle's per-bin restatement, and the delayed
acceptance sampler's exactness on a Gaussian target with a biased screen."""
import numpy as np
import pytest

from gundam_amd.mcmc import run_delayed, run_plain
from gundam_amd.surrogate import barlow_llh
from oracle import mlp_oracle, pyoracle


def _params(rng, i, h, o):
    return [rng.standard_normal((h, i)) / np.sqrt(i), 0.1 * rng.standard_normal(h),
            rng.standard_normal((h, h)) / np.sqrt(h), 0.1 * rng.standard_normal(h),
            rng.standard_normal((o, h)) / np.sqrt(h), 0.1 * rng.standard_normal(o)]


def test_mlp_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = _params(rng, 5, 7, 4)
    x = rng.standard_normal((6, 5))
    t = rng.standard_normal((6, 4))
    _, g = mlp_oracle.loss_and_grad(p, x, t)
    h = 1e-6
    for a in range(6):
        flat = p[a].reshape(-1)
        for j in range(0, flat.size, max(1, flat.size // 7)):
            old = flat[j]
            flat[j] = old + h
            lp, _ = mlp_oracle.loss_and_grad(p, x, t)
            flat[j] = old - h
            lm, _ = mlp_oracle.loss_and_grad(p, x, t)
            flat[j] = old
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[a].reshape(-1)[j]) < 1e-7 + 1e-6 * abs(fd)


def test_vectorised_barlow_matches_oracle():
    rng = np.random.default_rng(1)
    data = rng.poisson(20, 50).astype(float)
    data[:3] = 0.0
    pred = rng.uniform(0.5, 40, (4, 50))
    var = pred * rng.uniform(0.5, 2.0, (4, 50))
    var[0, :5] = 0.0
    got = barlow_llh(data, pred, var)
    cfg = pyoracle.llh_config("barlow")
    lib = pyoracle.lib()
    for r in range(4):
        want = sum(lib.oracle_llh_barlow_bin(d, p, np.sqrt(v), cfg)
                   for d, p, v in zip(data, pred[r], var[r]))
        assert abs(got[r] - want) <= 1e-9 * abs(want)


@pytest.mark.parametrize("n_sub", [1, 5])
def test_delayed_acceptance_targets_the_exact_distribution(n_sub):
    """Exact target N(m, S); the screen is a deliberately wrong Gaussian
    (shifted by 0.2, 1.3x wider).  The delayed-acceptance chains, with one
    screening step or a surrogate sub-chain of five, must still sample
    N(m, S): a missing second-stage correction would leave the screen's
    mean and width, which the tolerances below exclude."""
    rng = np.random.default_rng(2)
    P = 3
    m = np.array([0.5, -1.0, 2.0])
    s = np.array([0.3, 1.0, 0.6])

    def exact(th):
        th = np.atleast_2d(th)
        return -0.5 * (((th - m) / s) ** 2).sum(1)

    def screen(th):
        th = np.atleast_2d(th)
        return -0.5 * (((th - m - 0.2) / (1.3 * s)) ** 2).sum(1)

    cov = np.diag((2.38 * s) ** 2 / P)
    th0 = m + s * rng.standard_normal((64, P))
    r = run_delayed(exact, screen, th0, cov, 1500, rng, n_sub=n_sub)
    smp = r.samples[300:].reshape(-1, P)
    np.testing.assert_allclose(smp.mean(0), m, atol=0.05 * s.max() * 3)
    np.testing.assert_allclose(smp.std(0), s, rtol=0.06)
    assert r.exact_evals < r.proposals  # the screen saved exact evaluations
    p = run_plain(exact, m.copy(), cov, 20000, rng)
    np.testing.assert_allclose(p.samples[2000:, 0].mean(0), m, atol=0.1)
