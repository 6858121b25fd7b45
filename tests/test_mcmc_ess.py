"""Batch-means effective sample size (gundam_amd.mcmc.ess_batch_means)
against closed forms: independent draws have ESS = the number of draws; a
stationary AR(1) chain x_t = rho x_{t-1} + e_t has integrated autocorrelation
time (1 + rho) / (1 - rho), so ESS = n (1 - rho) / (1 + rho) per chain.  A
parameter that never moves has ESS 0.  An estimator that drops the batch
length factor (sigma^2 = var of the batch means) or that counts one chain
only misses these by factors far outside the tolerances."""
import numpy as np

from gundam_amd.mcmc import ess_batch_means


def _ar1(rng, n, chains, rho):
    x = np.empty((n, chains))
    x[0] = rng.standard_normal(chains) / np.sqrt(1 - rho * rho)
    e = rng.standard_normal((n, chains))
    for t in range(1, n):
        x[t] = rho * x[t - 1] + e[t]
    return x


def test_iid_draws():
    rng = np.random.default_rng(1)
    s = rng.standard_normal((10000, 8, 3))
    ess = ess_batch_means(s)
    kept = 9000 * 8
    assert np.all(np.abs(ess / kept - 1) < 0.15), ess / kept


def test_ar1_closed_form_and_frozen_parameter():
    rng = np.random.default_rng(2)
    n, C, rho = 60000, 4, 0.9
    s = np.stack([_ar1(rng, n, C, rho), _ar1(rng, n, C, 0.5),
                  np.full((n, C), 3.0)], axis=-1)
    ess = ess_batch_means(s)
    kept = (n - n // 10) * C
    for j, r in enumerate((rho, 0.5)):
        want = kept * (1 - r) / (1 + r)
        assert abs(ess[j] / want - 1) < 0.2, (r, ess[j], want)
    assert ess[2] == 0.0


def test_short_chain_has_no_batches():
    s = np.random.default_rng(3).standard_normal((1, 2, 2))
    assert np.all(ess_batch_means(s, burn_frac=0.0) == 0.0)
