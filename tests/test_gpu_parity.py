"""Parity of the HIP reweight-and-fill path against the CPU oracle on
identical seeded inputs (the reference CPU path: EventDialCache.cpp:156-171,
Histogram.cpp:107-135, PoissonLogLikelihood.h, BarlowBeestonBanff2022.h).

Bars: event -> bin indices bit-exact; per-event weights bit-exact in FP64
mode (the device rounds every operation as the host does; the required bar
is 1e-12 relative); bin sums, sqrt(sum w^2) and LLH within 1e-10 relative.
"""
import numpy as np
import pytest

from gundam_amd import CacheManager, Workload
from oracle import pyoracle

pytestmark = pytest.mark.gpu

RTOL_SUM = 1e-10
RTOL_LLH = 1e-10


def _setup(workload, n_events, seed=1):
    wl = Workload(workload, n_events=n_events, seed=seed)
    t = wl.tables
    ref_bins = pyoracle.bin_events(wl.vars, wl.edges)
    wl.bin_events_on_device(0)
    assert np.array_equal(t.bin, ref_bins), "event->bin indexing differs"
    return wl, t


def _check_point(wl, t, cm, orc, theta, kind="barlow"):
    tab = wl.table_values(theta)
    if t.table_len:
        orc.set_table_values(tab)
    w_ref, s_ref, e_ref = orc.propagate(theta)
    cm.propagate_parameters(theta, tab if t.table_len else None)
    s, e = cm.copy_histogram_contents()
    w = cm.copy_event_weights(np.full(t.n_events, np.nan))
    binned = t.bin >= 0
    np.testing.assert_array_equal(w[binned], w_ref[binned])
    np.testing.assert_allclose(s, s_ref, rtol=RTOL_SUM, atol=0)
    np.testing.assert_allclose(e, e_ref, rtol=RTOL_SUM, atol=0)
    return s_ref, e_ref


@pytest.mark.parametrize("workload,n_events", [(0, 0), (1, 0), (2, 60000), (3, 0)])
def test_fp64_parity(workload, n_events):
    print(f"parity workload={workload} n_events={n_events}", flush=True)
    wl, t = _setup(workload, n_events)
    orc = pyoracle.Oracle(t, n_threads=8)
    cm = CacheManager(0)
    cm.build(t)
    cm.set_event_weight_copy(True)
    nominal = wl.nominal()
    s_nom, _ = _check_point(wl, t, cm, orc, nominal)
    data = s_nom.copy()  # Asimov data at nominal
    kind = "poisson" if workload in (1, 3) else "barlow"
    cm.set_likelihood(kind, data, wl.prior_mean, wl.prior_sigma)
    for i, theta in enumerate(wl.random_points(3, 0.5, seed=11)):
        s_ref, e_ref = _check_point(wl, t, cm, orc, theta)
        tot_ref, st_ref, pen_ref = pyoracle.llh(kind, data, s_ref, e_ref, theta,
                                                wl.prior_mean, wl.prior_sigma)
        tot, st, pen = cm.llh()
        _close(st[0], st_ref)
        assert pen[0] == pen_ref
        _close(tot[0], tot_ref)


def _close(got, want):
    # +inf (a bin predicted <= 0 under Poisson) must match exactly
    if not np.isfinite(want):
        assert got == want, (got, want)
    else:
        assert abs(got - want) <= RTOL_LLH * abs(want), (got, want)


def test_batched_points_match_single():
    print("batched K=8 workload=2", flush=True)
    wl, t = _setup(2, 20000)
    pts = wl.random_points(8, 0.5, seed=5)
    tabs = wl.table_values(pts)
    cm = CacheManager(0)
    cm.build(t, max_points=8)
    cm.set_likelihood("barlow", np.full(t.n_bins, 3.0), wl.prior_mean, wl.prior_sigma)
    cm.set_parameters(pts)
    cm.set_table_values(tabs)
    cm.fill().get()
    hk = [cm.hist(k) for k in range(8)]
    llh_k = cm.llh()[0].copy()
    for k in range(8):
        cm.set_parameters(pts[k])
        cm.set_table_values(tabs[k])
        cm.fill().get()
        s, s2 = cm.hist(0)
        np.testing.assert_array_equal(s, hk[k][0])
        np.testing.assert_array_equal(s2, hk[k][1])
        assert cm.llh()[0][0] == llh_k[k]


@pytest.mark.parametrize("shape,K,prec", [
    # (tile words, evaluator threads, blocks per CU, LDS buffers)
    ((2048, 256, 2, 3), 3, "fp64"),  # three buffers: two producers flag alternate tiles
    ((3840, 128, 2, 2), 2, "fp64"),  # one table wave, ~90 table dials a tile: extra rounds
    ((3840, 384, 2, 2), 1, "fp64"),  # the 5M-event shape: two table waves
    ((3840, 64, 2, 2), 2, "fp64"),   # one evaluator wave: no table waves
    ((3584, 384, 2, 2), 3, "fp32"),
])
def test_table_wave_shapes_match_oracle(shape, K, prec):
    """The Tabulated/Kriged dials go to table waves that load their table
    values one step ahead (the tile's table part flagged in LDS by the
    producer that copied it); every launch shape and point count gives the
    oracle's bin sums (FP64: 1e-10, weights bit-exact; FP32: 1e-4)."""
    tw, bt, bpc, rb = shape
    print(f"table-wave shape {shape} K={K} {prec}", flush=True)
    wl, t = _setup(2, 30000)
    orc = pyoracle.Oracle(t, n_threads=8)
    cm = CacheManager(0, precision=prec, tile_words=tw, block_threads=bt,
                      blocks_per_cu=bpc, ring_buffers=rb)
    cm.build(t, max_points=K)
    cm.set_event_weight_copy(True)
    pts = wl.random_points(K, 0.5, seed=13)
    tabs = wl.table_values(pts)
    cm.set_parameters(pts)
    cm.set_table_values(tabs)
    cm.fill().get()
    w = cm.copy_event_weights(np.full(t.n_events, np.nan))  # point 0
    binned = t.bin >= 0
    for k in range(K):
        orc.set_table_values(tabs[k])
        w_ref, s_ref, e_ref = orc.propagate(pts[k])
        s, e = cm.copy_histogram_contents(k)
        if prec == "fp64":
            np.testing.assert_allclose(s, s_ref, rtol=RTOL_SUM, atol=0)
            np.testing.assert_allclose(e, e_ref, rtol=RTOL_SUM, atol=0)
            if k == 0:
                np.testing.assert_array_equal(w[binned], w_ref[binned])
        else:
            nz = s_ref > 0
            assert (np.abs(s[nz] - s_ref[nz]) / s_ref[nz]).max() < 1e-4
            assert np.all(s[~nz] == 0)


def test_fp32_mode_error_bounds():
    print("fp32 workload=2", flush=True)
    wl, t = _setup(2, 40000)
    orc = pyoracle.Oracle(t, n_threads=8)
    cm = CacheManager(0, precision="fp32")
    cm.build(t)
    theta = wl.random_points(1, 1.0, seed=9)[0]
    tab = wl.table_values(theta)
    orc.set_table_values(tab)
    _, s_ref, _ = orc.propagate(theta)
    cm.propagate_parameters(theta, tab)
    s, _ = cm.copy_histogram_contents()
    nz = s_ref > 0
    rel = np.abs(s[nz] - s_ref[nz]) / s_ref[nz]
    print("fp32 max per-bin rel err", rel.max())
    assert rel.max() < 1e-4


def test_cpp_facade_demo_runs():
    """The C++ Cache::Manager-shaped facade drives the library end to end."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "gundam_amd", "facade_demo")
    print("facade demo", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    vals = [float(l.split("LLH")[1]) for l in r.stdout.splitlines() if "LLH" in l and "point" in l]
    assert len(vals) == 4 and all(np.isfinite(vals)) and all(v > 0 for v in vals)


def test_graph_replays_match_eager():
    """hipGraph mode: several replays with new parameters, interleaved with
    eager work and synchronisations, each equal to the eager result."""
    print("graph replays workload=3 K=4", flush=True)
    wl, t = _setup(3, 0)
    pts = wl.random_points(12, 0.4, seed=21)
    eager = CacheManager(0)
    eager.build(t, max_points=4)
    graph = CacheManager(0)
    graph.build(t, max_points=4)
    data = np.full(t.n_bins, 40.0)
    for cm in (eager, graph):
        cm.set_likelihood("poisson", data, wl.prior_mean, wl.prior_sigma)
    graph.enable_graph(True)
    for r in range(3):
        batch = pts[4 * r:4 * r + 4]
        eager.set_parameters(batch)
        eager.fill().get()
        graph.set_parameters(batch)
        graph.fill().get()
        for k in range(4):
            np.testing.assert_array_equal(graph.hist(k)[0], eager.hist(k)[0])
            np.testing.assert_array_equal(graph.hist(k)[1], eager.hist(k)[1])
        np.testing.assert_array_equal(graph.llh()[0], eager.llh()[0])
        # eager single-point work on the graph context between replays
        graph.enable_graph(False)
        graph.set_parameters(pts[r])
        graph.fill().get()
        eager.set_parameters(pts[r])
        eager.fill().get()
        np.testing.assert_array_equal(graph.hist(0)[0], eager.hist(0)[0])
        graph.enable_graph(True)


def _search_tables(rng):
    """One event per dial (each in its own bin), all dials on parameter 0:
    general splines and graphs with sorted knots (including more than 33
    knots, where the reference's binary search stops at segment 31, and
    repeated abscissae) and unsorted knots.  (A NaN knot also takes the
    binary search, but its NaN weights are an error on both sides.)"""
    from gundam_amd import _native as N
    from gundam_amd.tables import INPUT_DTYPE, Tables

    dials = []
    for n in (2, 3, 5, 9, 15, 33, 34, 40, 60):
        xs = np.sort(rng.uniform(-3, 3, n))
        dials.append((N.DIAL_GENERAL, xs))
        dials.append((N.DIAL_GRAPH, xs))
    dup = np.array([-2.0, -1.0, -1.0, 0.0, 0.0, 0.0, 1.5, 2.5])
    dials += [(N.DIAL_GENERAL, dup), (N.DIAL_GRAPH, dup)]
    uns = np.array([-2.0, 1.0, -0.5, 0.5, 2.0, 1.8, 3.0])
    dials += [(N.DIAL_GENERAL, uns), (N.DIAL_GRAPH, uns)]
    dials.append((N.DIAL_GRAPH, np.array([0.3])))
    data, begin, types = [], [0], []
    for typ, xs in dials:
        n = xs.size
        y = rng.uniform(0.5, 1.5, n)
        if typ == N.DIAL_GENERAL:
            rec = [xs[0], 1.0]
            for j in range(n):
                rec += [y[j], rng.uniform(-0.5, 0.5), xs[j]]
        else:
            rec = []
            for j in range(n):
                rec += [y[j], xs[j]]
        data += rec
        begin.append(len(data))
        types.append(typ)
    nd = len(types)
    inputs = np.zeros(1, dtype=INPUT_DTYPE)
    inputs["mirror_lower"] = np.nan
    inputs["mirror_range"] = np.nan
    inputs["min_response"] = 0.0
    inputs["max_response"] = np.nan
    t = Tables(base_weight=rng.uniform(0.5, 2.0, nd), bin=np.arange(nd, dtype=np.int32),
               dial_begin=np.arange(nd + 1, dtype=np.int64),
               dial_type=np.array(types, np.uint8), dial_input=np.zeros(nd, np.int32),
               data_begin=np.array(begin, np.int64), dial_data=np.array(data),
               inputs=inputs, table_begin=np.zeros(1, np.int64), n_parameters=1,
               n_bins=nd)
    knots = np.concatenate([xs for _, xs in dials])
    return t, knots[np.isfinite(knots)]


def test_segment_search_edge_cases():
    """The device's counting segment search (sorted knots) and its binary
    search (unsorted or NaN knots) return the reference's segment: weights
    bit-exact against the oracle at parameter values on, between and beyond
    the knots."""
    print("segment search edge cases", flush=True)
    rng = np.random.default_rng(12)
    t, knots = _search_tables(rng)
    xs = np.concatenate([knots, np.nextafter(knots, np.inf), np.nextafter(knots, -np.inf),
                         rng.uniform(-4, 4, 40), [-1e30, 1e30, 0.0]])
    orc = pyoracle.Oracle(t)
    for prec in ("fp64",):
        cm = CacheManager(0, precision=prec)
        cm.build(t)
        cm.set_event_weight_copy(True)
        for x in xs:
            theta = np.array([x])
            w_ref, s_ref, _ = orc.propagate(theta)
            cm.propagate_parameters(theta)
            w = cm.copy_event_weights(np.full(t.n_events, np.nan))
            np.testing.assert_array_equal(w, w_ref, err_msg=f"x={x!r}")
        cm.close()
    orc.close()


@pytest.mark.parametrize("kind,opts", [
    ("poisson", {}), ("barlow", {}), ("barlow", {"nominal": True}),
    ("barlow", {"use_poisson_likelihood": True}), ("barlow_classic", {}),
    ("barlow_oa2020", {}), ("barlow_oa2021_sfgd", {"extra": True}), ("chi2", {}),
    ("least_squares", {}), ("least_squares", {"lsq_poissonian_approximation": True})])
def test_llh_variants_match_oracle(kind, opts):
    """Every likelihood type and option on the device against the oracle on
    the same histogram (stat within 1e-10 relative), with bins that have no
    prediction masked as ignoreBinsWithZeroPredictionAtPrior does."""
    print("llh variant", kind, opts, flush=True)
    wl, t = _setup(1, 0)
    orc = pyoracle.Oracle(t)
    theta = wl.random_points(1, 0.3, seed=9)[0]
    _, s_ref, e_ref = orc.propagate(theta)
    rng = np.random.default_rng(4)
    data = rng.poisson(np.maximum(s_ref, 0) * 1.05 + 0.2).astype(float)
    o = dict(opts)
    kw = {"ignore_bins": (s_ref == 0).astype(np.uint8)}
    if o.pop("nominal", False):
        kw["nominal_err"] = 1.1 * e_ref + 0.01
    if o.pop("extra", False):
        kw["extra_fraction"] = np.where(np.arange(t.n_bins) % 3 == 0, 0.04, 0.0)
    kw.update(o)
    cm = CacheManager(0)
    cm.build(t)
    cm.set_likelihood(kind, data, wl.prior_mean, wl.prior_sigma, **kw)
    cm.propagate_parameters(theta)
    _, st, pen = cm.llh()
    _, st_ref, pen_ref = pyoracle.llh(kind, data, s_ref, e_ref, theta, wl.prior_mean,
                                      wl.prior_sigma, **kw)
    assert np.isfinite(st_ref)
    assert abs(st[0] - st_ref) <= RTOL_LLH * abs(st_ref), (st[0], st_ref)
    assert pen[0] == pen_ref
    cm.close()
    orc.close()


def test_batched_gradient_matches_single_points_and_oracle():
    """The batched central-difference gradient (2m+1 points in launches of up
    to max_points) equals the one formed from single-point evaluations bit
    for bit, and the oracle's to the LLH tolerance."""
    print("batched gradient workload=2", flush=True)
    wl, t = _setup(2, 20000)
    cm = CacheManager(0)
    cm.build(t, max_points=8)
    nominal = wl.nominal()
    cm.propagate_parameters(nominal, wl.table_values(nominal[None])[0])
    data = cm.hist(0)[0] * 1.02 + 0.5
    cm.set_likelihood("barlow", data, wl.prior_mean, wl.prior_sigma)
    theta = wl.random_points(1, 0.3, seed=11)[0]
    params = [0, 5, 40, 120, 299]
    h = 1e-3
    l0, g = cm.gradient(theta, h, params, wl.table_values)
    orc = pyoracle.Oracle(t)
    single, ref = [], []
    for p in params:
        vals, rvals = [], []
        for sgn in (1, -1):
            th = theta.copy()
            th[p] += sgn * h
            tab = wl.table_values(th[None])[0]
            cm.propagate_parameters(th, tab)
            vals.append(cm.llh()[0][0])
            orc.set_table_values(tab)
            _, s, e = orc.propagate(th, want_weights=False)
            rvals.append(pyoracle.llh("barlow", data, s, e, th, wl.prior_mean,
                                      wl.prior_sigma)[0])
        single.append((vals[0] - vals[1]) / (2 * h))
        ref.append((rvals[0] - rvals[1]) / (2 * h))
    np.testing.assert_array_equal(g, np.array(single))
    tol = 2 * RTOL_LLH * abs(l0) / (2 * h)
    assert np.all(np.abs(g - np.array(ref)) <= tol), (g, ref, tol)
    cm.close()
    orc.close()


def test_many_clamp_classes_use_the_escape_path():
    """More distinct response supervisors (lo, hi, flags) than the dial
    descriptors' 15 input classes: the inputs past them read their full
    record per dial.  Every input gets its own upper clamp here (and every
    third one no extrapolation); bit-exact weights and 1e-10 sums against the
    oracle on the feature workload, single point and batched."""
    import dataclasses
    print("escape classes workload=0", flush=True)
    wl, t = _setup(0, 0)
    inp = t.inputs.copy()
    inp["max_response"] = 1.2 + 0.01 * np.arange(inp.shape[0])
    inp["flags"][::3] |= 1
    t2 = dataclasses.replace(t, inputs=inp, _keep=[t])
    assert len({(a, b, c) for a, b, c in zip(inp["min_response"], inp["max_response"],
                                             inp["flags"])}) > 15
    orc = pyoracle.Oracle(t2, n_threads=8)
    cm = CacheManager(0)
    cm.build(t2, max_points=4)
    cm.set_event_weight_copy(True)
    pts = wl.random_points(4, 0.6, seed=13)
    for theta in pts[:2]:
        _check_point(wl, t2, cm, orc, theta)
    tabs = wl.table_values(pts) if t2.table_len else None
    cm.set_parameters(pts)
    if tabs is not None:
        cm.set_table_values(tabs)
    cm.fill().get()
    for k in range(4):
        if tabs is not None:
            orc.set_table_values(tabs[k])
        _, s_ref, e_ref = orc.propagate(pts[k])
        s, s2 = cm.hist(k)
        np.testing.assert_allclose(s, s_ref, rtol=RTOL_SUM, atol=0)
        np.testing.assert_allclose(np.sqrt(s2), e_ref, rtol=RTOL_SUM, atol=0)


def test_nan_weight_is_reported_once():
    """A NaN event weight makes the next sync fail with GCM_ERR_NUMERIC
    (the reference throws, CacheWeights.cpp:83-86); the flag is cleared, so
    a later fill with finite parameters syncs cleanly.  Norm dials only: the
    NaN parameter flows through no index computation."""
    from gundam_amd import INPUT_DTYPE, Tables
    from gundam_amd import _native as N
    print("nan flag", flush=True)
    n = 200
    inputs = np.zeros(2, dtype=INPUT_DTYPE)
    inputs["parameter"] = [0, 1]
    inputs["mirror_lower"] = np.nan
    inputs["mirror_range"] = np.nan
    inputs["min_response"] = np.nan
    inputs["max_response"] = np.nan
    t = Tables(base_weight=np.ones(n), bin=(np.arange(n) % 4).astype(np.int32),
               dial_begin=np.arange(n + 1, dtype=np.int64),
               dial_type=np.full(n, N.DIAL_NORM, np.uint8),
               dial_input=(np.arange(n) % 2).astype(np.int32),
               data_begin=np.zeros(n + 1, np.int64), dial_data=np.zeros(0),
               inputs=inputs, table_begin=np.zeros(1, np.int64), n_parameters=2, n_bins=4)
    cm = CacheManager(0)
    cm.build(t)
    cm.set_parameters(np.array([np.nan, 1.0]))
    with pytest.raises(N.GcmError) as e:
        cm.fill().get()
    assert e.value.code == N.GCM_ERR_NUMERIC
    cm.set_parameters(np.array([2.0, 1.0]))
    cm.fill().get()
    np.testing.assert_array_equal(cm.hist(0)[0], [100.0, 50.0, 100.0, 50.0])
    # a NaN fill that is never synchronised leaves nothing behind: the next
    # evaluation clears the flag in-stream (eager and graph replay alike)
    for graph in (False, True):
        cm.enable_graph(graph)
        cm.set_parameters(np.array([np.nan, 1.0]))
        cm.fill()
        cm.set_parameters(np.array([3.0, 1.0]))
        cm.fill().get()
        np.testing.assert_array_equal(cm.hist(0)[0], [150.0, 50.0, 150.0, 50.0])
        cm.set_parameters(np.array([np.nan, 1.0]))
        with pytest.raises(N.GcmError) as e:
            cm.fill().get()
        assert e.value.code == N.GCM_ERR_NUMERIC
    cm.close()


def test_validate_histogram_contents():
    """Cache::Manager::ValidateHistogramContents: the device histogram matches
    the oracle's (CPU-filled) one under the reference's almostEqual rule, and
    a bin moved by one part in 1e9 is reported."""
    print("validate histograms workload=3", flush=True)
    wl, t = _setup(3, 0)
    orc = pyoracle.Oracle(t, n_threads=8)
    cm = CacheManager(0)
    cm.build(t)
    theta = wl.random_points(1, 0.5, seed=19)[0]
    _, s_ref, e_ref = orc.propagate(theta, want_weights=False)
    cm.propagate_parameters(theta)
    assert cm.validate_histogram_contents(s_ref, e_ref)
    bad = s_ref.copy()
    j = int(np.argmax(bad))
    bad[j] *= 1 + 1e-9
    assert not cm.validate_histogram_contents(bad, e_ref, verbose=True)
    assert cm.last_validation_failures == 1
    err = e_ref.copy()
    err[0] = np.nan
    assert not cm.validate_histogram_contents(s_ref, err)
    cm.close()
    orc.close()


@pytest.mark.parametrize("workload,n_events,K", [(0, 0, 16), (2, 20000, 64), (3, 0, 70)])
def test_points_kernel_matches_oracle(workload, n_events, K):
    """K >= 16 points per launch run one lane per point (fill_points_kernel):
    every point's bin sums, sqrt(sum w^2) and LLH against the oracle evaluated
    point by point (1e-10), point 0's event weights bit-exact.  Workload 0
    holds every dial type (2-D dials, mirrors, caps, shifts, second inputs);
    K = 70 takes two chunks of lanes (64 + 6)."""
    print(f"points kernel workload={workload} n_events={n_events} K={K}", flush=True)
    wl, t = _setup(workload, n_events)
    orc = pyoracle.Oracle(t, n_threads=8)
    cm = CacheManager(0)
    cm.build(t, max_points=K)
    cm.set_event_weight_copy(True)
    _, data, _ = orc.propagate(wl.nominal(), want_weights=False) if not t.table_len else \
        (None, np.full(t.n_bins, 3.0), None)
    kind = "poisson" if workload in (1, 3) else "barlow"
    cm.set_likelihood(kind, data, wl.prior_mean, wl.prior_sigma)
    pts = wl.random_points(K, 0.5, seed=23)
    tabs = wl.table_values(pts) if t.table_len else None
    cm.set_parameters(pts)
    if tabs is not None:
        cm.set_table_values(tabs)
    cm.fill().get()
    w = cm.copy_event_weights(np.full(t.n_events, np.nan))
    tot = cm.llh()[0].copy()
    binned = t.bin >= 0
    for k in range(K):
        if tabs is not None:
            orc.set_table_values(tabs[k])
        w_ref, s_ref, e_ref = orc.propagate(pts[k], want_weights=(k == 0))
        if k == 0:
            np.testing.assert_array_equal(w[binned], w_ref[binned])
        s, s2 = cm.hist(k)
        np.testing.assert_allclose(s, s_ref, rtol=RTOL_SUM, atol=0)
        np.testing.assert_allclose(np.sqrt(s2), e_ref, rtol=RTOL_SUM, atol=0)
        tot_ref, _, _ = pyoracle.llh(kind, data, s_ref, e_ref, pts[k], wl.prior_mean,
                                     wl.prior_sigma)
        _close(tot[k], tot_ref)
    cm.close()
    orc.close()


def test_points_kernel_against_single_points_and_fp32():
    """The one-lane-per-point kernel against one-point launches of the same
    tables: event weights equal bit for bit (same products in the same
    order), bin sums within 1e-12 relative (the two kernels add a bin's
    events in different orders: serial per (tile, range) slot here, a tree
    per tile there); and the FP32 mode's 32-point launch within the FP32
    bounds of test_fp32_mode_error_bounds against the FP64 one."""
    print("points kernel vs single, fp32", flush=True)
    wl, t = _setup(2, 30000)
    pts = wl.random_points(32, 0.5, seed=29)
    tabs = wl.table_values(pts)
    cms = {}
    for prec in ("fp64", "fp32"):
        cms[prec] = CacheManager(0, precision=prec)
        cms[prec].build(t, max_points=32)
        cms[prec].set_event_weight_copy(True)
        cms[prec].set_likelihood("barlow", np.full(t.n_bins, 3.0), wl.prior_mean,
                                 wl.prior_sigma)
        cms[prec].set_parameters(pts)
        cms[prec].set_table_values(tabs)
        cms[prec].fill().get()
    cm = cms["fp64"]
    hk = [cm.hist(k) for k in range(32)]
    h32 = [cms["fp32"].hist(k) for k in range(32)]
    w_batch = cm.copy_event_weights(np.full(t.n_events, np.nan))
    binned = t.bin >= 0
    for k in range(32):
        s64, _ = hk[k]
        nz = s64 != 0
        assert np.max(np.abs(h32[k][0][nz] - s64[nz]) / s64[nz]) < 1e-4
    for k in (0, 7, 31):
        cm.set_parameters(pts[k])
        cm.set_table_values(tabs[k])
        cm.fill().get()
        s, s2 = cm.hist(0)
        np.testing.assert_allclose(hk[k][0], s, rtol=1e-12, atol=0)
        np.testing.assert_allclose(hk[k][1], s2, rtol=1e-12, atol=0)
        if k == 0:
            w = cm.copy_event_weights(np.full(t.n_events, np.nan))
            np.testing.assert_array_equal(w_batch[binned], w[binned])
    for c in cms.values():
        c.close()
