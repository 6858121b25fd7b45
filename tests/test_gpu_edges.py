"""Edge cases of the tables against the oracle: no events, every event
unbinned, events without dials, bins without events, tiles closed by the
64-event limit (norm-only events carry no record words), ragged dial counts
and spline sizes including the largest knot counts, uniform-knot splines
with power-of-two and other knot steps, the reweight cap, one event, and an
event too large for a tile (a clean build error).

Bars as in the parity tests: weights bit-exact, sums within 1e-10 relative,
LLH within 1e-10 relative.  Every case runs one point per launch, a 4-point
launch and a 20-point launch (the one-lane-per-point kernel)."""
import numpy as np
import pytest

from gundam_amd import INPUT_DTYPE, CacheManager, Tables
from gundam_amd import _native as N
from oracle import pyoracle

pytestmark = pytest.mark.gpu


def _inputs(n_par):
    inp = np.zeros(n_par, dtype=INPUT_DTYPE)
    inp["parameter"] = np.arange(n_par)
    inp["mirror_lower"] = np.nan
    inp["mirror_range"] = np.nan
    inp["min_response"] = 0.0
    inp["max_response"] = np.nan
    return inp


def _mono(rng, n):
    return np.concatenate([[-3.0, 6.0 / (n - 1)], 1 + 0.1 * rng.standard_normal(n)])


def _general(rng, n):
    x = np.sort(np.concatenate([[-3.0, 3.0], rng.uniform(-2.9, 2.9, n - 2)]))
    y = 1 + 0.1 * rng.standard_normal(n)
    m = np.gradient(y, x)
    out = np.empty(3 * n)
    out[0::3], out[1::3], out[2::3] = y, m, x
    return np.concatenate([[x[0], 0.0], out])


def _graph(rng, n):
    x = np.sort(rng.uniform(-3, 3, n))
    out = np.empty(2 * n)
    out[0::2], out[1::2] = 1 + 0.1 * rng.standard_normal(n), x
    return out


def _tables(events, n_par, n_bins, cap=np.nan):
    """events: list of (base, bin, [(type, input, payload), ...])."""
    base = np.array([e[0] for e in events], np.float64)
    bins = np.array([e[1] for e in events], np.int32)
    dials = [d for e in events for d in e[2]]
    dial_begin = np.concatenate([[0], np.cumsum([len(e[2]) for e in events])]).astype(np.int64)
    lens = [len(d[2]) for d in dials]
    return Tables(base_weight=base, bin=bins, dial_begin=dial_begin,
                  dial_type=np.array([d[0] for d in dials], np.uint8),
                  dial_input=np.array([d[1] for d in dials], np.int32),
                  data_begin=np.concatenate([[0], np.cumsum(lens)]).astype(np.int64),
                  dial_data=(np.concatenate([np.asarray(d[2], np.float64) for d in dials])
                             if dials and sum(lens) else np.zeros(0)),
                  inputs=_inputs(n_par), table_begin=np.zeros(1, np.int64),
                  n_parameters=n_par, n_bins=n_bins, reweight_cap=cap)


def _check(t, thetas, max_points=4):
    orc = pyoracle.Oracle(t, n_threads=4)
    cm = CacheManager(0)
    cm.build(t, max_points=max_points)
    cm.set_event_weight_copy(True)
    data = np.full(t.n_bins, 3.0)
    cm.set_likelihood("poisson", data)
    for th in thetas:
        w_ref, s_ref, e_ref = orc.propagate(th)
        cm.set_parameters(th)
        cm.fill().get()
        w = cm.copy_event_weights(np.full(t.n_events, np.nan))
        binned = t.bin >= 0
        np.testing.assert_array_equal(w[binned], w_ref[binned])
        s, e = cm.copy_histogram_contents()
        np.testing.assert_allclose(s, s_ref, rtol=1e-10, atol=0)
        np.testing.assert_allclose(e, e_ref, rtol=1e-10, atol=0)
        _, st_ref, _ = pyoracle.llh("poisson", data, s_ref, e_ref)
        st = cm.llh()[1][0]
        if np.isfinite(st_ref):
            assert abs(st - st_ref) <= 1e-10 * max(1.0, abs(st_ref))
        else:
            assert st == st_ref
    # batched points equal the single points
    pts = np.array(thetas[:max_points])
    cm.set_parameters(pts)
    cm.fill().get()
    for k in range(pts.shape[0]):
        _, s_ref, _ = orc.propagate(pts[k], want_weights=False)
        np.testing.assert_allclose(cm.hist(k)[0], s_ref, rtol=1e-10, atol=0)
    cm.close()
    # 20 points per launch: the one-lane-per-point kernel (lanes 20-63 idle);
    # point 0's event weights bit-exact, every point's sums and LLH to 1e-10
    base = np.array(thetas)
    rng = np.random.default_rng(7)
    pts = base[np.arange(20) % len(base)] + (np.arange(20) >= len(base))[:, None] * \
        0.1 * rng.standard_normal((20, base.shape[1]))
    cmp = CacheManager(0)
    cmp.build(t, max_points=20)
    cmp.set_event_weight_copy(True)
    cmp.set_likelihood("poisson", data)
    cmp.set_parameters(pts)
    cmp.fill().get()
    w = cmp.copy_event_weights(np.full(t.n_events, np.nan))
    st = cmp.llh()[1]
    for k in range(20):
        w_ref, s_ref, e_ref = orc.propagate(pts[k])
        if k == 0:
            binned = t.bin >= 0
            np.testing.assert_array_equal(w[binned], w_ref[binned])
        s, s2 = cmp.hist(k)
        np.testing.assert_allclose(s, s_ref, rtol=1e-10, atol=0)
        np.testing.assert_allclose(np.sqrt(s2), e_ref, rtol=1e-10, atol=0)
        _, st_ref, _ = pyoracle.llh("poisson", data, s_ref, e_ref)
        if np.isfinite(st_ref):
            assert abs(st[k] - st_ref) <= 1e-10 * max(1.0, abs(st_ref))
        else:
            assert st[k] == st_ref
    cmp.close()
    orc.close()


def test_no_events_and_all_unbinned():
    print("edges: empty", flush=True)
    t = _tables([], 2, 3)
    _check(t, [np.array([1.0, 2.0])])
    rng = np.random.default_rng(1)
    ev = [(1.0, -1, [(N.DIAL_MONOTONIC, 0, _mono(rng, 5))]) for _ in range(50)]
    _check(_tables(ev, 2, 3), [np.array([0.5, 1.0]), np.array([-1.0, 2.0])])


def test_dialless_events_empty_bins_and_event_limit():
    print("edges: dialless / empty bins / 64-event tiles", flush=True)
    rng = np.random.default_rng(2)
    ev = []
    for i in range(3000):
        b = 3 if i < 2000 else 7  # bins 0-2, 4-6, 8-9 stay empty
        kind = i % 3
        if kind == 0:
            dials = []
        elif kind == 1:
            dials = [(N.DIAL_NORM, 0, [])]
        else:
            dials = [(N.DIAL_NORM, 0, []), (N.DIAL_NORM, 1, [])]
        ev.append((rng.uniform(0.5, 2.0), b, dials))
    t = _tables(ev, 2, 10)
    _check(t, [np.array([1.0, 1.0]), np.array([2.5, 0.3]), np.array([-0.5, 1.5]),
               np.array([0.0, 4.0])])


def test_ragged_dials_largest_knots_and_cap():
    print("edges: ragged / knots / cap", flush=True)
    rng = np.random.default_rng(3)
    n_par = 6
    ev = []
    for i in range(4000):
        nd = int(rng.integers(0, 31))
        dials = []
        for j in range(nd):
            k = int(rng.integers(0, 4))
            inp = int(rng.integers(0, n_par))
            if k == 0:
                dials.append((N.DIAL_MONOTONIC, inp, _mono(rng, int(rng.integers(5, 16)))))
            elif k == 1:
                dials.append((N.DIAL_GENERAL, inp, _general(rng, int(rng.choice([2, 15, 30])))))
            elif k == 2:
                dials.append((N.DIAL_GRAPH, inp, _graph(rng, int(rng.choice([1, 2, 15, 40])))))
            else:
                dials.append((N.DIAL_NORM, inp, []))
        ev.append((rng.uniform(0.1, 1.0), int(rng.integers(-1, 25)), dials))
    t = _tables(ev, n_par, 25, cap=2.0)
    th = [np.zeros(n_par), rng.normal(0, 1, n_par), rng.normal(0, 2.5, n_par),
          np.full(n_par, 1.7)]
    _check(t, th)


def test_uniform_knot_steps():
    """Uniform-knot splines (monotonic, compact, uniform) with knot steps
    2^-10 .. 2^8, -1/2 and steps that are not powers of two (the knot index
    is the correctly rounded quotient (x - x0) / step in every kernel), x
    inside and far outside the knots.  Each rotation of the points puts
    another point first, so that the K-point launch's event weights are
    compared bit-exact for every point."""
    print("edges: uniform knot steps", flush=True)
    rng = np.random.default_rng(5)
    n_par = 5
    steps = [2.0 ** k for k in (-10, -3, -1, 0, 1, 3, 8)] + [-0.5, 0.75, 1.0 / 3.0, 3.0, 0.1]
    ev = []
    for i in range(1500):
        dials = []
        for j in range(int(rng.integers(1, 6))):
            st = float(rng.choice(steps))
            n = int(rng.integers(3, 12))
            low = float(rng.uniform(-3.0, 0.0))
            kind = int(rng.integers(0, 3))
            y = 1 + 0.1 * rng.standard_normal(n)
            if kind == 2:
                pairs = np.empty(2 * n)
                pairs[0::2], pairs[1::2] = y, 0.1 * rng.standard_normal(n)
                dials.append((N.DIAL_UNIFORM, int(rng.integers(0, n_par)),
                              np.concatenate([[low, st], pairs])))
            else:
                dials.append((N.DIAL_MONOTONIC if kind == 0 else N.DIAL_COMPACT,
                              int(rng.integers(0, n_par)), np.concatenate([[low, st], y])))
        ev.append((rng.uniform(0.1, 1.0), int(rng.integers(0, 10)), dials))
    t = _tables(ev, n_par, 10)
    th = [np.zeros(n_par), rng.normal(0, 1, n_par), rng.normal(0, 300, n_par),
          np.full(n_par, -0.02)]
    for r in range(len(th)):
        _check(t, th[r:] + th[:r])


def test_single_event_and_oversized_event():
    print("edges: one event / oversized", flush=True)
    rng = np.random.default_rng(4)
    t = _tables([(2.0, 0, [(N.DIAL_MONOTONIC, 0, _mono(rng, 7)), (N.DIAL_NORM, 1, [])])], 2, 1)
    _check(t, [np.array([0.3, 1.2]), np.array([-2.0, 0.5])])
    # 2000 dials in one event exceed every tile's dial capacity
    big = _tables([(1.0, 0, [(N.DIAL_NORM, 0, [])] * 2000)], 1, 1)
    cm = CacheManager(0)
    with pytest.raises(N.GcmError) as e:
        cm.build(big)
    assert e.value.code == N.GCM_ERR_UNSUPPORTED
    cm.close()
