"""The oracle's correlated-penalty restatement (oracle_penalty_set, after
LikelihoodInterface.cpp:186-213 and ParameterSet.cpp:441-463) pinned without
the reference: exact rational arithmetic on small integer cases, the
quadratic form d^T C^-1 d against an independent solve, the diagonal case
against sum (d/sigma)^2, and the out-of-bounds rule case by case."""
from fractions import Fraction

import numpy as np
import pytest

from oracle import pyoracle


def test_integer_case_is_exact():
    # d = (1, 2), A = [[2, 1], [1, 3]]: A d = (4, 7), d.(A d) = 18
    th = np.array([1.0, 2.0, 3.0])
    assert pyoracle.penalty_set(th, [0, 2], [0.0, 1.0], [[2.0, 1.0], [1.0, 3.0]]) == 18.0


def test_quadratic_form_matches_rational_brute_force():
    rng = np.random.default_rng(0)
    for n in (1, 3, 6):
        a = rng.integers(-4, 5, size=(n, n)).astype(float)
        a = a + a.T + 2 * n * np.eye(n)
        d = rng.integers(-5, 6, size=n).astype(float) / 4  # exact binary fractions
        th = d + 0.5
        want = sum(Fraction(d[i]) * Fraction(a[i, j]) * Fraction(d[j])
                   for i in range(n) for j in range(n))
        assert pyoracle.penalty_set(th, np.arange(n), np.full(n, 0.5), a) == float(want)


@pytest.mark.parametrize("n", [2, 10, 40])
def test_quadratic_form_matches_solve(n):
    """d^T C^-1 d with C^-1 given equals d . solve(C, d) (an independent
    route through the covariance itself)."""
    rng = np.random.default_rng(n)
    m = rng.standard_normal((n, n))
    cov = m @ m.T + n * np.eye(n)
    par = rng.permutation(n + 5)[:n]
    th = rng.standard_normal(n + 5)
    prior = 0.3 * rng.standard_normal(n)
    got = pyoracle.penalty_set(th, par, prior, np.linalg.inv(cov))
    d = th[par] - prior
    want = float(d @ np.linalg.solve(cov, d))
    assert abs(got - want) <= 1e-11 * want


def test_diagonal_covariance_is_the_diagonal_penalty():
    rng = np.random.default_rng(3)
    n = 12
    sig = rng.uniform(0.1, 2.0, n)
    th = rng.standard_normal(n)
    mean = rng.standard_normal(n)
    got = pyoracle.penalty_set(th, np.arange(n), mean, np.diag(1.0 / sig ** 2))
    want = pyoracle.llh("poisson", [], [], [], th, mean, sig)[2]
    assert abs(got - want) <= 1e-14 * want


def test_out_of_bounds_rule():
    th = np.array([0.5, -1.0, 2.0])
    par, prior = [0, 1, 2], [0.0, 0.0, 0.0]
    a = np.eye(3)
    lo, hi = np.array([-1.0, -1.0, -1.0]), np.array([1.0, 1.0, np.nan])
    # x2 = 2 has no upper bound; x1 = -1 sits on its lower bound
    assert pyoracle.penalty_set(th, par, prior, a, lo, hi) == 0.25 + 1.0 + 4.0
    assert pyoracle.penalty_set(th, par, prior, a, lo, np.array([1.0, 1.0, 1.5])) == np.inf
    assert pyoracle.penalty_set(th, par, prior, a, np.array([-1.0, -0.5, -1.0]), hi) == np.inf
    th_nan = th.copy()
    th_nan[0] = np.nan
    assert pyoracle.penalty_set(th_nan, par, prior, a, lo, hi) == np.inf
    # bounds only, no covariance
    assert pyoracle.penalty_set(th, par, prior, None, lo, hi) == 0.0
    assert pyoracle.penalty_set(th, [2], [0.0], None, [0.0], [1.0]) == np.inf


def test_sets_add_after_the_diagonal_term():
    th = np.array([1.0, 2.0, 3.0])
    sets = [{"parameters": [0, 2], "prior": [0.0, 1.0], "inv_cov": [[2.0, 1.0], [1.0, 3.0]]},
            {"parameters": [1], "prior": [0.0], "inv_cov": [[0.5]]}]
    diag = pyoracle.llh("poisson", [], [], [], th, np.zeros(3), np.ones(3))[2]
    assert pyoracle.penalty(th, sets, np.zeros(3), np.ones(3)) == (diag + 18.0) + 2.0
    assert pyoracle.penalty(th, sets) == (0.0 + 18.0) + 2.0
