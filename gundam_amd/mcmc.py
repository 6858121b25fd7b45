"""Metropolis chains on the exact likelihood of the cache, plain and with the
surrogate as the first stage of delayed acceptance.

Target: log pi(theta) = -LLH_total(theta) / 2, as GUNDAM's SimpleMcmc proxy
(SimpleMcmc.h:255-267: -0.5 x evalFit).  Proposals are Gaussian random walks
with a covariance adapted during burn-in (Haario, Saksman & Tamminen 2001,
Bernoulli 7:223) and frozen afterwards.

Delayed acceptance (Christen & Fox 2005, J. Comput. Graph. Stat. 14:795): a
proposal passes stage 1 with probability min(1, exp(s(theta') - s(theta)))
where s is the surrogate log target; only then is the exact log target L
evaluated, and the move is accepted with probability
min(1, exp((L(theta') - L(theta)) - (s(theta') - s(theta)))).  The chain's
stationary distribution is the exact target; the surrogate only decides
which proposals are worth an exact evaluation.  Many chains run in lockstep,
so each iteration is one batched MLP forward and one batched exact launch.

With n_sub > 1 the first stage is n_sub Metropolis steps on the surrogate
target (the surrogate transition method, Liu 2001, Monte Carlo Strategies in
Scientific Computing, Springer, sec. 9.4.3): that kernel is reversible with
respect to the surrogate, so its endpoint is a valid proposal with the same
second-stage ratio, and an exact evaluation buys up to n_sub surrogate moves.
"""
from __future__ import annotations

import time
from dataclasses import dataclass
from typing import Callable, Optional

import numpy as np

from . import _native as N
from .cache_manager import CacheManager
from .surrogate import Surrogate, barlow_llh


class ExactTarget:
    """-LLH/2 from the device path, K points per launch."""

    def __init__(self, cm: CacheManager, table_values: Optional[Callable] = None,
                 record: bool = False):
        self.cm = cm
        self.table_values = table_values
        self.calls = 0
        self.points = 0
        self.nan_batches = 0      # launches that reported a NaN event weight
        self.record = record      # keep (theta, bin contents) of every point
        self.rec_theta: list = []
        self.rec_sums: list = []

    def __call__(self, thetas: np.ndarray) -> np.ndarray:
        thetas = np.atleast_2d(thetas)
        out = np.empty(thetas.shape[0])
        K = self.cm.max_points
        for i in range(0, thetas.shape[0], K):
            th = thetas[i:i + K]
            self.cm.set_parameters(th)
            if self.table_values is not None and self.cm.table_len:
                self.cm.set_table_values(self.table_values(th))
            try:
                self.cm.fill().get()
            except N.GcmError as e:
                # a NaN event weight (CacheWeights.cpp:83-86): the points it
                # touches have a NaN likelihood, which the chains reject below
                if e.code != N.GCM_ERR_NUMERIC:
                    raise
                self.nan_batches += 1
            llh = self.cm.llh()[0][:th.shape[0]]
            out[i:i + th.shape[0]] = np.where(np.isnan(llh), -np.inf, -0.5 * llh)
            if self.record:
                self.rec_theta.append(th.copy())
                self.rec_sums.append(np.stack([self.cm.hist(k)[0]
                                               for k in range(th.shape[0])]))
            self.calls += 1
            self.points += th.shape[0]
        return out

    def recorded(self):
        """(theta, bin contents) of every recorded point, stacked."""
        return np.concatenate(self.rec_theta), np.concatenate(self.rec_sums)

    def sums(self, thetas: np.ndarray) -> np.ndarray:
        """Bin contents for each point (training pairs for the surrogate)."""
        thetas = np.atleast_2d(thetas)
        out = np.empty((thetas.shape[0], self.cm.n_bins))
        K = self.cm.max_points
        for i in range(0, thetas.shape[0], K):
            th = thetas[i:i + K]
            self.cm.set_parameters(th)
            if self.table_values is not None and self.cm.table_len:
                self.cm.set_table_values(self.table_values(th))
            self.cm.fill().get()
            for k in range(th.shape[0]):
                out[i + k] = self.cm.hist(k)[0]
        return out


class SurrogateTarget:
    """-LLH/2 with the surrogate's predicted bin contents: Barlow-Beeston
    against the data with MC variance = predicted content x the nominal
    (sum w^2 / sum w) ratio of each bin (evaluated on the device), plus the
    diagonal prior penalty."""

    def __init__(self, sur: Surrogate, data: np.ndarray, var_ratio: np.ndarray,
                 prior_mean: np.ndarray, prior_sigma: np.ndarray):
        self.sur = sur
        sur.set_head(data, var_ratio)
        self.mu = prior_mean
        self.sig = np.where(prior_sigma > 0, prior_sigma, np.inf)
        self.points = 0
        self.calls = 0

    def __call__(self, thetas: np.ndarray) -> np.ndarray:
        thetas = np.atleast_2d(thetas)
        stat = self.sur.chi2(thetas)
        pen = (((thetas - self.mu) / self.sig) ** 2).sum(1)
        self.points += thetas.shape[0]
        self.calls += 1
        return -0.5 * (stat + pen)


def var_ratio(sums: np.ndarray, sum_w2: np.ndarray) -> np.ndarray:
    """Per-bin (sum w^2 / sum w) at one point, 0 for empty bins; the inputs
    are the two outputs of CacheManager.hist (sum w and sum w^2)."""
    return np.where(sums > 0, sum_w2 / np.where(sums > 0, sums, 1.0), 0.0)


def laplace_scales(exact: ExactTarget, mode: np.ndarray, h: float = 1e-2) -> np.ndarray:
    """Per-parameter posterior widths from the diagonal curvature of the exact
    log target at the mode (one batched stencil of 2P + 1 points)."""
    P = mode.size
    pts = np.repeat(mode[None, :], 2 * P + 1, 0)
    for j in range(P):
        pts[1 + 2 * j, j] += h
        pts[2 + 2 * j, j] -= h
    L = exact(pts)
    curv = -(L[1::2] + L[2::2] - 2 * L[0]) / (h * h)
    return 1.0 / np.sqrt(np.maximum(curv, 1e-12))


@dataclass
class ChainResult:
    samples: np.ndarray          # [steps, chains, P]
    accepted: int                # accepted moves (all chains)
    proposals: int               # proposals (all chains)
    exact_evals: int             # exact likelihood evaluations
    stage1_passed: int
    seconds: float

    @property
    def accepted_per_s(self) -> float:
        return self.accepted / self.seconds

    def sq_jumps(self, scales: np.ndarray) -> float:
        """Sum over chains and steps of the squared jump ||d theta / scales||^2
        (the expected-squared-jumping-distance numerator, Pasarica & Gelman
        2010, Stat. Sinica 20:343)."""
        d = np.diff(self.samples, axis=0) / scales
        return float((d * d).sum())


def ess_batch_means(samples: np.ndarray, burn_frac: float = 0.1,
                    n_batches: Optional[int] = None) -> np.ndarray:
    """Effective sample size of each parameter by non-overlapping batch means
    (Flegal & Jones 2010, Ann. Statist. 38:1034).  samples: [steps, chains, P].
    The first burn_frac of the steps is dropped; each chain's remaining n
    steps are cut into a batches of b = n // a steps (a = floor(sqrt(n)) by
    default); sigma^2_BM = b * sum over all chains' batches of (batch mean -
    pooled mean)^2 / (a C - 1), and ESS = a b C var(x) / sigma^2_BM.  A
    parameter that never moves gets ESS 0."""
    x = samples[int(samples.shape[0] * burn_frac):]
    n, C_, P = x.shape
    a = n_batches if n_batches else max(2, int(np.sqrt(n)))
    b = n // a
    if b < 1:
        return np.zeros(P)
    x = x[n - a * b:]  # the last a b steps
    flat = x.reshape(-1, P)
    mu = flat.mean(0)
    var = flat.var(0, ddof=1)
    bm = x.reshape(a, b, C_, P).mean(1)
    s2 = b * ((bm - mu) ** 2).sum((0, 1)) / (a * C_ - 1)
    with np.errstate(divide="ignore", invalid="ignore"):
        ess = np.where(s2 > 0, a * b * C_ * var / s2, 0.0)
    return ess


def adapt_cov(samples: np.ndarray, scale0: np.ndarray) -> np.ndarray:
    d = samples.shape[-1]
    flat = samples.reshape(-1, d)
    cov = np.cov(flat.T) if flat.shape[0] > 2 * d else np.diag(scale0 ** 2)
    return (2.38 ** 2 / d) * cov + 1e-10 * np.eye(d)


def run_plain(exact: ExactTarget, theta0: np.ndarray, cov: np.ndarray, steps: int,
              rng: np.random.Generator, max_seconds: Optional[float] = None) -> ChainResult:
    """One chain, one exact evaluation per proposal (the plain GPU chain).
    Stops after `steps` or once max_seconds have passed."""
    P = theta0.size
    Lc = np.linalg.cholesky(cov)
    th = theta0.copy()
    L = exact(th[None, :])[0]
    out = []
    acc = 0
    t0 = time.perf_counter()
    n = 0
    for s in range(steps):
        prop = th + Lc @ rng.standard_normal(P)
        Lp = exact(prop[None, :])[0]
        if np.log(rng.uniform()) < Lp - L:
            th, L = prop, Lp
            acc += 1
        out.append(th[None, :].copy())
        n = s + 1
        if max_seconds is not None and time.perf_counter() - t0 > max_seconds:
            break
    return ChainResult(np.stack(out), acc, n, n, n, time.perf_counter() - t0)


def run_delayed(exact: ExactTarget, surr: SurrogateTarget, theta0: np.ndarray,
                cov: np.ndarray, steps: int, rng: np.random.Generator,
                n_sub: int = 1, max_seconds: Optional[float] = None) -> ChainResult:
    """C chains in lockstep with surrogate screening (theta0: [C, P]).  Each
    iteration runs n_sub surrogate Metropolis steps per chain (one batched
    surrogate call each) and sends the chains that moved to one batched exact
    launch; samples are recorded once per iteration."""
    C_, P = theta0.shape
    Lc = np.linalg.cholesky(cov)
    th = theta0.copy()
    L = exact(th)
    S = surr(th)
    out = []
    acc = passed = n_exact = n = 0
    t0 = time.perf_counter()
    for s in range(steps):
        y = th.copy()
        Sy = S.copy()
        for _ in range(n_sub):
            prop = y + rng.standard_normal((C_, P)) @ Lc.T
            Sp = surr(prop)
            a1 = np.log(rng.uniform(size=C_)) < (Sp - Sy)
            y[a1] = prop[a1]
            Sy[a1] = Sp[a1]
        idx = np.nonzero(np.any(y != th, axis=1))[0]
        passed += idx.size
        if idx.size:
            Lp = exact(y[idx])
            n_exact += idx.size
            a2 = np.log(rng.uniform(size=idx.size)) < (Lp - L[idx]) - (Sy[idx] - S[idx])
            take = idx[a2]
            th[take] = y[take]
            L[take] = Lp[a2]
            S[take] = Sy[take]
            acc += take.size
        out.append(th.copy())
        n = s + 1
        if max_seconds is not None and time.perf_counter() - t0 > max_seconds:
            break
    return ChainResult(np.stack(out), acc, n * C_, n_exact, passed, time.perf_counter() - t0)


def run_batched_exact(exact: ExactTarget, theta0: np.ndarray, cov: np.ndarray,
                      steps: int, rng: np.random.Generator,
                      max_seconds: Optional[float] = None) -> ChainResult:
    """C chains in lockstep, every proposal evaluated exactly (one batched
    launch per step) -- the reference for the surrogate's contribution."""
    C_, P = theta0.shape
    Lc = np.linalg.cholesky(cov)
    th = theta0.copy()
    L = exact(th)
    out = []
    acc = n = 0
    t0 = time.perf_counter()
    for s in range(steps):
        prop = th + rng.standard_normal((C_, P)) @ Lc.T
        Lp = exact(prop)
        a = np.log(rng.uniform(size=C_)) < (Lp - L)
        th[a] = prop[a]
        L[a] = Lp[a]
        acc += int(a.sum())
        out.append(th.copy())
        n = s + 1
        if max_seconds is not None and time.perf_counter() - t0 > max_seconds:
            break
    return ChainResult(np.stack(out), acc, n * C_, n * C_, n * C_, time.perf_counter() - t0)
