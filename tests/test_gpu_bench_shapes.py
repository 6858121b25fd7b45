"""Parity at the shapes bench.py times, against the CPU oracle.

* The 64-point batched launch on the 100k-event set re-split into uneven
  samples (50k/30k/15k/5k), at the launch shape the autotuner picks for it,
  captured in a hipGraph and replayed: every point's bin sums, sqrt(sum w^2)
  and LLH against the oracle evaluated point by point.
* One full-size point of the 5M-event, 300-parameter, 6000-bin workload at the
  bench's launch shape with the device table refresh: event weights bit-exact,
  bin sums and LLH within 1e-10 relative (the oracle is fed the lookup tables
  the device computed, whose cos may differ from the host's in the last bit).
"""
import time

import numpy as np
import pytest

from gundam_amd import CacheManager, Workload
from gundam_amd import _native as N
from oracle import pyoracle

pytestmark = pytest.mark.gpu

RTOL = 1e-10


def _close(got, want):
    assert abs(got - want) <= RTOL * abs(want), (got, want)


def _c3_points(wl, seed):
    """32 chain points and a 32-point central-difference stencil (the
    bench's batch)."""
    rng = np.random.default_rng(seed)
    chains = wl.prior_mean + wl.prior_sigma * rng.standard_normal((32, wl.n_parameters))
    centre = wl.prior_mean + 0.3 * wl.prior_sigma * rng.standard_normal(wl.n_parameters)
    stencil = np.repeat(centre[None, :], 32, axis=0)
    for j in range(16):
        stencil[2 * j, j] += 1e-3
        stencil[2 * j + 1, j] -= 1e-3
    return np.concatenate([chains, stencil])


def test_batched_k64_graph_at_autotuned_shape():
    from gundam_amd.autotune import autotune
    print("bench shape: K=64 graph, workload 3, autotuned", flush=True)
    wl = Workload(3, seed=0)
    t = wl.tables
    wl.bin_events_on_device(0)
    assert np.array_equal(t.bin, pyoracle.bin_events(wl.vars, wl.edges))
    pts = _c3_points(wl, 3)
    tuned = autotune(t, pts, None, n_events=t.n_events, reps=2)
    shape = {k: tuned[k] for k in ("block_threads", "tile_words", "blocks_per_cu",
                                   "ring_buffers")}
    print(f"autotuned shape {shape}", flush=True)
    cm = CacheManager(0, **shape)
    cm.build(t, max_points=64)
    orc = pyoracle.Oracle(t, n_threads=8)
    _, data, _ = orc.propagate(wl.nominal(), want_weights=False)
    cm.set_likelihood("barlow", data, wl.prior_mean, wl.prior_sigma)
    cm.enable_graph(True)
    for rep in range(3):  # capture, then two replays with other points
        p = pts if rep == 0 else _c3_points(wl, 100 + rep)
        cm.set_parameters(p)
        cm.fill().get()
        tot = cm.llh()[0].copy()
        for k in range(64):
            _, s_ref, e_ref = orc.propagate(p[k], want_weights=False)
            s, s2 = cm.hist(k)
            np.testing.assert_allclose(s, s_ref, rtol=RTOL, atol=0)
            np.testing.assert_allclose(np.sqrt(s2), e_ref, rtol=RTOL, atol=0)
            tot_ref, _, _ = pyoracle.llh("barlow", data, s_ref, e_ref, p[k],
                                         wl.prior_mean, wl.prior_sigma)
            _close(tot[k], tot_ref)
    cm.close()
    orc.close()


def test_full_size_plain_point():
    print("bench shape: workload 2 full size, one point", flush=True)
    t0 = time.time()
    wl = Workload(2, seed=0)
    t = wl.tables
    wl.bin_events_on_device(0)
    print(f"generated and binned {t.n_events} events in {time.time() - t0:.1f}s", flush=True)
    cm = CacheManager(0, block_threads=384, tile_words=3840, blocks_per_cu=2, ring_buffers=2)
    cm.build(t)
    cm.set_event_weight_copy(True)
    cm.add_table_update(N.TABLE_UPDATE_COS, N.TABLE_UPDATE_COS_KERNEL, wl.table_jobs())
    orc = pyoracle.Oracle(t, n_threads=16)
    print(f"device and oracle built at {time.time() - t0:.1f}s", flush=True)
    nominal = wl.nominal()
    cm.propagate_parameters(nominal)
    orc.set_table_values(cm.device_table_values(0))
    _, data, _ = orc.propagate(nominal, want_weights=False)
    cm.set_likelihood("barlow", data, wl.prior_mean, wl.prior_sigma)
    theta = wl.random_points(1, 0.5, seed=17)[0]
    cm.propagate_parameters(theta)
    orc.set_table_values(cm.device_table_values(0))
    w_ref, s_ref, e_ref = orc.propagate(theta)
    w = cm.copy_event_weights(np.full(t.n_events, np.nan))
    binned = t.bin >= 0
    assert np.array_equal(w[binned], w_ref[binned])
    s, e = cm.copy_histogram_contents()
    np.testing.assert_allclose(s, s_ref, rtol=RTOL, atol=0)
    np.testing.assert_allclose(e, e_ref, rtol=RTOL, atol=0)
    tot_ref, st_ref, pen_ref = pyoracle.llh("barlow", data, s_ref, e_ref, theta,
                                            wl.prior_mean, wl.prior_sigma)
    tot, st, pen = cm.llh()
    _close(st[0], st_ref)
    assert pen[0] == pen_ref
    _close(tot[0], tot_ref)
    print(f"done at {time.time() - t0:.1f}s", flush=True)
    cm.close()
    orc.close()
