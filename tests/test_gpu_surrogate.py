"""The surrogate's BF16 MFMA kernels against the FP64 CPU restatement
(oracle/mlp_oracle.py), training, and what the trained network gains in
delayed-acceptance MCMC over the same chains without it.

Tolerances: BF16 operands (8-bit mantissa) with FP32 accumulation give
~0.4% relative error per rounding; the checks use 3% (normalised) on
gradients and outputs."""
import numpy as np
import pytest

from gundam_amd import CacheManager, Workload
from gundam_amd.mcmc import (ExactTarget, SurrogateTarget, laplace_scales,
                             run_batched_exact, run_delayed)
from gundam_amd.surrogate import MLP, Surrogate
from oracle import mlp_oracle

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)


@pytest.mark.parametrize("shape,rows", [((20, 64, 40), 100), ((37, 96, 130), 257)])
def test_mlp_forward_and_gradients_match_fp64(shape, rows):
    print("mlp grad check", shape, rows, flush=True)
    n_in, h, n_out = shape
    rng = np.random.default_rng(3)
    mlp = MLP(n_in, h, n_out, seed=5)
    flat = mlp.params().astype(np.float64)
    params = [a.copy() for a in mlp.split(flat)]
    params[1][:] = 0.1 * rng.standard_normal(h)
    params[3][:] = 0.1 * rng.standard_normal(h)
    params[5][:] = 0.1 * rng.standard_normal(n_out)
    mlp.set_params(np.concatenate([a.ravel() for a in params]).astype(np.float32))
    x = rng.standard_normal((rows, n_in)).astype(np.float32)
    t = rng.standard_normal((rows, n_out)).astype(np.float32)
    y_ref, _ = mlp_oracle.forward(params, x.astype(np.float64))
    y = mlp.predict(x)
    assert _rel(y, y_ref) < 3e-2
    loss, g = mlp.gradients(x, t)
    loss_ref, g_ref = mlp_oracle.loss_and_grad(params, x.astype(np.float64), t.astype(np.float64))
    assert abs(loss - loss_ref) < 3e-2 * loss_ref
    for a, b in zip(mlp.split(g.astype(np.float64)), g_ref):
        assert _rel(a, b) < 3e-2, _rel(a, b)


def test_mlp_training_reduces_loss():
    print("mlp training", flush=True)
    rng = np.random.default_rng(4)
    x = rng.standard_normal((4096, 8)).astype(np.float32)
    w = rng.standard_normal((8, 32))
    t = np.tanh(x @ w / 3).astype(np.float32)
    mlp = MLP(8, 128, 32, seed=1)
    mlp.set_data(x, t)
    losses = mlp.train(40, batch=128, lr=2e-3)
    assert losses[-1] < 0.1 * losses[0], losses


def test_surrogate_delayed_acceptance_gain():
    """On the 100k-event workload: delayed acceptance with the trained
    surrogate samples the same posterior as the exact chains while spending
    fewer exact evaluations per accepted move."""
    print("surrogate DA-MCMC workload=1", flush=True)
    wl = Workload(1, seed=1)
    t = wl.tables
    wl.bin_events_on_device(0)
    C_ = 32
    cm = CacheManager(0)
    cm.build(t, max_points=64)
    cm.propagate_parameters(wl.nominal())
    data, s2 = cm.hist(0)
    cm.set_likelihood("barlow", data.copy(), wl.prior_mean, wl.prior_sigma)
    exact = ExactTarget(cm)
    mode = wl.nominal()
    sc = laplace_scales(exact, mode)
    rng = np.random.default_rng(7)
    train = mode + 3.0 * sc * rng.standard_normal((12000, mode.size))
    sums = exact.sums(train)
    sur = Surrogate(mode.size, t.n_bins, mode, sc, hidden=256, seed=2)
    losses = sur.fit(train, sums, epochs=40, batch=256, lr=2e-3)
    assert losses[-1] < 0.05 * losses[0]
    ratio = np.where(data > 0, s2 / np.maximum(data, 1e-300), 1.0)
    surr = SurrogateTarget(sur, data, ratio, wl.prior_mean, wl.prior_sigma)
    d = mode.size
    cov = (2.38 ** 2 / d) * np.diag(sc ** 2)
    th0 = mode + sc * rng.standard_normal((C_, d))
    da = run_delayed(exact, surr, th0, cov, 400, rng)
    ex = run_batched_exact(exact, th0, cov, 400, rng)
    a = da.samples[100:].reshape(-1, d)
    b = ex.samples[100:].reshape(-1, d)
    # same posterior: means within 0.25 posterior sigma at these short lengths
    assert np.all(np.abs(a.mean(0) - b.mean(0)) < 0.25 * sc)
    gain = (ex.exact_evals / ex.accepted) / (da.exact_evals / da.accepted)
    print("exact evaluations per accepted move: exact", ex.exact_evals / ex.accepted,
          "delayed", da.exact_evals / da.accepted, "gain", gain)
    assert gain > 1.5
