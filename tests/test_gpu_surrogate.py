"""The surrogate's BF16 MFMA kernels against the FP64 CPU restatement
(oracle/mlp_oracle.py), training, and what the trained network gains in
delayed-acceptance MCMC over the same chains without it.

Tolerances: BF16 operands (8-bit mantissa) with FP32 accumulation give
~0.4% relative error per rounding; the checks use 3% (normalised) on
gradients and outputs."""
import numpy as np
import pytest

from gundam_amd import CacheManager, Workload
from gundam_amd.mcmc import (ExactTarget, SurrogateTarget, adapt_cov, laplace_scales,
                             run_batched_exact, run_delayed, var_ratio)
from gundam_amd.surrogate import MLP, Surrogate, barlow_llh
from oracle import mlp_oracle

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)


# (300, 512, 600) and (24, 64, 3000) take the split-K paths: at the first
# every GEMM of the step is split 2-5 ways along K; at the second the
# 3000-long reduction of the output layer's input gradient is split sixteen
# ways.  (300, 512, 6000) is the 5M-event model's network (300 parameters,
# 6000 bins) at the training minibatch of 256 rows and at 4096 rows, where
# the output layer's forward product takes the 128 x 128 tile.
# (37, 700, 1500) at 3000 rows puts that tile on edges that are not tile
# multiples in both dimensions, with a K tail (700 = 10 x 64 + 60), next to
# an unaligned input layer.
@pytest.mark.parametrize("shape,rows", [((20, 64, 40), 100), ((37, 96, 130), 257),
                                        ((300, 512, 600), 256), ((24, 64, 3000), 256),
                                        ((300, 512, 6000), 256), ((300, 512, 6000), 4096),
                                        ((37, 700, 1500), 3000)])
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


@pytest.mark.parametrize("shape,rows", [((300, 512, 6000), 4096), ((300, 512, 600), 2048)])
def test_split_k_sums_per_block_and_repeatable(shape, rows):
    """The split-K sums of the backward pass: every 64 x 64 block of every
    weight gradient within 10% of the FP64 restatement (a partial missing
    from a tile's sum, or read before it was written, leaves ~50% in its
    blocks; BF16 operands leave 0.3-0.4%, measured; a norm over the whole
    matrix would not tell a few such blocks apart), and four calls on the
    same rows equal bit for bit.  At 4096 rows dW3 (split 3) and the output
    layer's input gradient (split 4, GELU derivative in the sum's epilogue)
    take 128 x 128 tiles, dW2 and dW1 64 x 64 tiles split 16 ways."""
    print("split-K sums", shape, rows, flush=True)
    n_in, h, n_out = shape
    rng = np.random.default_rng(8)
    mlp = MLP(n_in, h, n_out, seed=6)
    flat = mlp.params().astype(np.float64)
    params = [a.copy() for a in mlp.split(flat)]
    for k in (1, 3, 5):
        params[k][:] = 0.1 * rng.standard_normal(params[k].size)
    mlp.set_params(np.concatenate([a.ravel() for a in params]).astype(np.float32))
    x = rng.standard_normal((rows, n_in)).astype(np.float32)
    t = rng.standard_normal((rows, n_out)).astype(np.float32)
    _, g0 = mlp.gradients(x, t)
    _, g_ref = mlp_oracle.loss_and_grad(params, x.astype(np.float64), t.astype(np.float64))
    worst = 0.0
    for a, b in zip(mlp.split(g0.astype(np.float64)), g_ref):
        if a.ndim != 2:
            continue
        for r0 in range(0, a.shape[0], 64):
            for c0 in range(0, a.shape[1], 64):
                blk, ref = a[r0:r0 + 64, c0:c0 + 64], b[r0:r0 + 64, c0:c0 + 64]
                worst = max(worst, _rel(blk, ref))
    print(f"  worst 64 x 64 block relative error {worst:.3e}", flush=True)
    assert worst < 0.1
    for _ in range(3):
        _, g = mlp.gradients(x, t)
        np.testing.assert_array_equal(g, g0)
    mlp.close()


def test_mlp_predict_row_chunks_and_row_counts():
    """predict over 1, 33 and 8193 rows against the FP64 restatement.  With
    1100 inputs the first layer's reduction is split along K for the small
    row counts (so the split workspace grows between calls) and not for the
    8192-row chunk; 8193 rows cross the chunk boundary.  Rows of the large
    call equal the same rows predicted alone (no mix-up of rows across
    chunks or splits; the two paths sum in different orders, hence the
    tolerance)."""
    print("mlp predict chunks", flush=True)
    rng = np.random.default_rng(6)
    n_in, h, n_out = 1100, 64, 96
    mlp = MLP(n_in, h, n_out, seed=9)
    params = [a.astype(np.float64) for a in mlp.split(mlp.params().astype(np.float64))]
    x = rng.standard_normal((8193, n_in)).astype(np.float32)
    for rows in (1, 33, 8193):
        y = mlp.predict(x[:rows])
        y_ref, _ = mlp_oracle.forward(params, x[:rows].astype(np.float64))
        assert _rel(y, y_ref) < 3e-2, (rows, _rel(y, y_ref))
    big = mlp.predict(x)
    for r in (0, 8191, 8192):
        one = mlp.predict(x[r:r + 1])
        np.testing.assert_allclose(big[r], one[0], rtol=1e-3, atol=1e-4)
    mlp.close()


@pytest.mark.parametrize("shape,rows", [((20, 64, 40), 256), ((21, 1024, 42), 256),
                                        ((19, 63, 37), 256), ((8, 512, 12000), 960)])
def test_mlp_train_step_matches_gradients(shape, rows):
    """One epoch whose minibatch is the whole training set is one Adam step on
    the whole set's gradient, whatever the shuffle: the parameters after it
    against gradients() on the rows in their own order and Adam's first step
    on the host.  The first step moves each parameter by about lr times the
    sign of its gradient, so rows gathered wrongly, or inputs and targets from
    different rows, move most parameters the wrong way; the two row orders
    sum in different orders, so a few parameters whose gradient is near zero
    may differ.  (20, 64, 40) reads the targets through the row order in the
    output layer's loss head; (21, 1024, 42), whose output layer is split
    along K, gathers them, and its row lengths are not multiples of 4.  The
    output layer's weights take their Adam step in the epilogue of their
    gradient's product where their range is 16-byte aligned: in a 64 x 64
    tile's epilogue here, in a 128 x 128 tile's at (8, 512, 12000) over 960
    rows (376 tiles, 15 K steps: not split), and in the Adam launch at
    (19, 63, 37), whose range is not aligned."""
    print("mlp train step", shape, rows, flush=True)
    n_in, h, n_out = shape
    lr = 1e-3
    rng = np.random.default_rng(12)
    x = rng.standard_normal((rows, n_in)).astype(np.float32)
    t = rng.standard_normal((rows, n_out)).astype(np.float32)
    mlp = MLP(n_in, h, n_out, seed=3)
    p0 = mlp.params().astype(np.float64)
    _, g = mlp.gradients(x, t)
    g = g.astype(np.float64)
    mlp.set_data(x, t)
    mlp.train(1, batch=rows, lr=lr, seed=7)
    p1 = mlp.params().astype(np.float64)
    b1, b2, eps = 0.9, 0.999, 1e-8
    m, v = (1 - b1) * g, (1 - b2) * g * g
    want = p0 - lr * (m / (1 - b1)) / (np.sqrt(v / (1 - b2)) + eps)
    off = np.abs(p1 - want) > 0.1 * lr
    print(f"  parameters off by more than lr/10: {off.sum()} of {off.size}", flush=True)
    assert off.mean() < 0.01
    mlp.close()


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
    """On the 100k-event workload: the surrogate, trained on pairs that the
    exact path generates during burn-in, screens proposals for delayed
    acceptance.  The chains sample the same posterior as the exact chains
    while spending fewer exact evaluations per accepted move, and surrogate
    sub-chains move farther per exact evaluation."""
    print("surrogate DA-MCMC workload=1", flush=True)
    wl = Workload(1, seed=1)
    t = wl.tables
    wl.bin_events_on_device(0)
    C_ = 32
    cm = CacheManager(0)
    cm.build(t, max_points=64)
    mode = wl.nominal()
    cm.propagate_parameters(mode)
    data, err = cm.hist(0)
    data = data.copy()
    ratio = var_ratio(data, err)
    # the surrogate's MC-variance model (content x nominal sum w^2 / sum w)
    # reproduces the device's Barlow-Beeston statistic at the nominal point
    off = 1.1 * data + 0.5
    cm.set_likelihood("barlow", off, wl.prior_mean, wl.prior_sigma)
    cm.propagate_parameters(mode)
    np.testing.assert_allclose(barlow_llh(off, data[None, :], (data * ratio)[None, :])[0],
                               cm.llh()[1][0], rtol=1e-10)
    cm.set_likelihood("barlow", data, wl.prior_mean, wl.prior_sigma)
    exact = ExactTarget(cm)
    sc = laplace_scales(exact, mode)
    d = mode.size
    rng = np.random.default_rng(7)
    # burn-in: exact chains from Laplace draws; every evaluated point is kept
    exact.record = True
    burn = run_batched_exact(exact, mode + sc * rng.standard_normal((C_, d)),
                             (2.38 ** 2 / d) * np.diag(sc ** 2), 150, rng)
    exact(mode + sc * rng.standard_normal((12000 - exact.points, d)))
    exact.record = False
    X, Y = exact.recorded()
    cov = adapt_cov(burn.samples[75:], sc)
    sur = Surrogate(d, t.n_bins, hidden=256, seed=2)
    losses = sur.fit(X, Y, epochs=150, batch=256, lr=2e-3)
    assert losses[-1] < 0.1 * losses[0], (losses[0], losses[-1])
    surr = SurrogateTarget(sur, data, ratio, wl.prior_mean, wl.prior_sigma)
    # the device likelihood head against the host Barlow-Beeston restatement
    test = mode + sc * rng.standard_normal((256, d))
    pred = sur.predict_sums(test)
    host = barlow_llh(data, pred, pred * ratio)
    np.testing.assert_allclose(sur.chi2(test), host, rtol=1e-9)
    start = burn.samples[-1]
    da = run_delayed(exact, surr, start.copy(), cov, 1000, rng)
    ex = run_batched_exact(exact, start.copy(), cov, 1000, rng)
    sub = run_delayed(exact, surr, start.copy(), cov, 250, rng, n_sub=4)
    a = da.samples[200:].reshape(-1, d)
    b = ex.samples[200:].reshape(-1, d)
    # same posterior: means within 0.3 posterior sigma at these lengths
    assert np.all(np.abs(a.mean(0) - b.mean(0)) < 0.3 * sc), np.abs(a.mean(0) - b.mean(0)) / sc
    gain = (ex.exact_evals / ex.accepted) / (da.exact_evals / da.accepted)
    jump_gain = (sub.sq_jumps(sc) / sub.exact_evals) / (ex.sq_jumps(sc) / ex.exact_evals)
    print("exact evaluations per accepted move: exact", ex.exact_evals / ex.accepted,
          "delayed", da.exact_evals / da.accepted, "gain", gain,
          "; squared jump per exact evaluation, 4-step sub-chains vs exact", jump_gain)
    assert gain > 1.5
    assert jump_gain > 2.0


def test_surrogate_delayed_acceptance_c2_model():
    """Delayed acceptance on the 5M-event model at its stated network
    (300 -> 512 -> 512 -> 6000, device table refresh): the surrogate, trained
    on 16k exact pairs, agrees with the exact chi2 up to a smooth offset; the
    device likelihood head equals the host Barlow-Beeston restatement; the
    delayed-acceptance chains (32, from the exact burn-in states) stay on the
    exact chains' posterior at these lengths and spend fewer exact
    evaluations per accepted move."""
    print("surrogate DA-MCMC workload=2 (5M events)", flush=True)
    from gundam_amd import _native as N
    wl = Workload(2, seed=3)
    t = wl.tables
    wl.bin_events_on_device(0)
    C_ = 32
    cm = CacheManager(0, block_threads=384, tile_words=3840, blocks_per_cu=2, ring_buffers=2)
    cm.build(t, max_points=64)
    cm.add_table_update(N.TABLE_UPDATE_COS, N.TABLE_UPDATE_COS_KERNEL, wl.table_jobs())
    mode = wl.nominal()
    cm.propagate_parameters(mode)
    data, err = (a.copy() for a in cm.hist(0))
    ratio = var_ratio(data, err)
    cm.set_likelihood("barlow", data, wl.prior_mean, wl.prior_sigma)
    exact = ExactTarget(cm)
    sc = laplace_scales(exact, mode)
    d = mode.size
    assert d == 300 and t.n_bins == 6000
    rng = np.random.default_rng(11)
    exact.record = True
    burn = run_batched_exact(exact, mode + sc * rng.standard_normal((C_, d)),
                             (2.38 ** 2 / d) * np.diag(sc ** 2), 120, rng)
    exact(mode + sc * rng.standard_normal((16000 - exact.points, d)))
    exact.record = False
    X, Y = exact.recorded()
    print(f"{X.shape[0]} training pairs", flush=True)
    cov = adapt_cov(burn.samples[60:], sc)
    sur = Surrogate(d, t.n_bins, hidden=512, seed=2)
    losses = sur.fit(X, Y, epochs=120, batch=256, lr=1e-3)
    assert losses[-1] < 0.5 * losses[0], (losses[0], losses[-1])
    surr = SurrogateTarget(sur, data, ratio, wl.prior_mean, wl.prior_sigma)
    test = mode + sc * rng.standard_normal((256, d))
    pred = sur.predict_sums(test)
    np.testing.assert_allclose(sur.chi2(test), barlow_llh(data, pred, pred * ratio), rtol=1e-9)
    # surrogate minus exact chi2 = an offset plus a spread well below the
    # spread of the exact statistic term over the same points (what a
    # surrogate that ignores theta would leave; the prior term is exact in
    # both)
    dchi2 = 2.0 * (surr(test) - exact(test))
    cen = np.abs(dchi2 - np.median(dchi2))
    stat = []
    for i in range(0, test.shape[0], 64):
        cm.set_parameters(test[i:i + 64])
        cm.fill()
        stat.append(cm.llh()[1].copy())
    stat = np.concatenate(stat)
    mad = np.median(np.abs(stat - np.median(stat)))
    print("surrogate - exact chi2: offset", np.median(dchi2), "|centred| p50", np.median(cen),
          "exact stat term spread (MAD)", mad, flush=True)
    assert np.median(cen) < 0.8 * mad
    start = burn.samples[-1]
    da = run_delayed(exact, surr, start.copy(), cov, 300, rng)
    print("delayed chains done", flush=True)
    ex = run_batched_exact(exact, start.copy(), cov, 300, rng)
    a = da.samples[100:].reshape(-1, d)
    b = ex.samples[100:].reshape(-1, d)
    diff = np.abs(a.mean(0) - b.mean(0)) / sc
    print("posterior mean difference / sigma: median", np.median(diff), "max", diff.max(),
          flush=True)
    assert np.median(diff) < 0.15 and diff.max() < 0.6
    gain = (ex.exact_evals / ex.accepted) / (da.exact_evals / da.accepted)
    print("exact evaluations per accepted move: exact", ex.exact_evals / ex.accepted,
          "delayed", da.exact_evals / da.accepted, "gain", gain, flush=True)
    assert gain > 1.5
    cm.close()
