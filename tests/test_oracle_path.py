"""The oracle's event path (mirror, supervisor clamp, product order, cap,
histogram refill, binning) and likelihoods against independent restatements,
known answers and closed forms."""
import math

import numpy as np
import pytest
from scipy.optimize import minimize_scalar

from gundam_amd import INPUT_DTYPE, Tables, Workload
from gundam_amd import _native as N
from oracle import pyoracle


def _workload(wid, n_events=0, seed=1):
    wl = Workload(wid, n_events=n_events, seed=seed)
    wl.tables.bin[:] = pyoracle.bin_events(wl.vars, wl.edges)
    return wl


def test_nominal_weights_are_exactly_one():
    wl = _workload(1, 5000)
    orc = pyoracle.Oracle(wl.tables, n_threads=4)
    w, s, e = orc.propagate(wl.nominal())
    assert np.all(w == 1.0)
    b = wl.tables.bin
    np.testing.assert_array_equal(s, np.bincount(b[b >= 0], minlength=wl.tables.n_bins))


def _python_weights(wl, theta, tab):
    """Independent pure-Python restatement of EventDialCache::reweightEntry
    over DialInputBuffer::update and DialResponseSupervisor::process."""
    t = wl.tables
    xin = []
    for inp in t.inputs:
        v = theta[inp["parameter"]]
        lo = inp["mirror_lower"]
        if not math.isnan(lo):
            r = inp["mirror_range"]
            v = abs(math.fmod(v - lo, 2 * r))
            if v > r:
                v = -(v - 2 * r)
            v += lo
        xin.append(v)
    cap = t.reweight_cap
    w = np.empty(t.n_events)
    for e in range(t.n_events):
        prod = 1.0
        for d in range(t.dial_begin[e], t.dial_begin[e + 1]):
            typ = int(t.dial_type[d])
            inp = t.inputs[t.dial_input[d]]
            p = t.dial_data[t.data_begin[d]:t.data_begin[d + 1]]
            x = xin[t.dial_input[d]]
            noext = inp["flags"] & N.INPUT_NO_EXTRAPOLATION
            if typ == N.DIAL_NORM:
                r = x
            elif typ == N.DIAL_SHIFT:
                r = p[0]
            elif typ == N.DIAL_TABULATED:
                base = t.table_begin[int(p[0])]
                f = float(np.float32(p[2]))
                i = base + int(p[1])
                r = f * tab[i] + (1.0 - f) * tab[i + 1]
            elif typ == N.DIAL_KRIGED:
                base = t.table_begin[int(p[0])]
                r = 0.0
                for k in range(1, p.size, 2):
                    r += float(np.float32(p[k + 1])) * tab[base + int(p[k])]
            elif typ in (N.DIAL_BILINEAR, N.DIAL_BICUBIC):
                # Bilinear/Bicubic::evalResponse: both coordinates clamped to
                # the knot range when extrapolation is off
                y = xin[t.dial_input2[d]]
                nx, ny = int(p[0]), int(p[1])
                xk, yk = p[2:2 + nx], p[2 + nx:2 + nx + ny]
                if noext:
                    x = min(max(x, xk[0]), xk[-1])
                    y = min(max(y, yk[0]), yk[-1])
                pc = np.ascontiguousarray(p)
                r = pyoracle.lib().oracle_calc2d(typ, x, y, -1e20, 1e20, pc.ctypes.data)
            elif typ == N.DIAL_GRAPH:
                if noext and x <= p[1]:
                    r = p[0]
                elif noext and x >= p[-1]:
                    r = p[-2]
                else:
                    r = pyoracle.calc(typ, x, -1e20, 1e20, p, p.size)
            else:
                n = {N.DIAL_MONOTONIC: p.size - 2, N.DIAL_COMPACT: p.size - 2,
                     N.DIAL_UNIFORM: (p.size - 2) // 2,
                     N.DIAL_GENERAL: (p.size - 2) // 3}[typ]
                if noext:
                    hi = p[2 + 3 * (n - 1) + 2] if typ == N.DIAL_GENERAL else p[0] + (n - 1) * p[1]
                    x = p[0] if x <= p[0] else (hi if x >= hi else x)
                dim = n if typ in (N.DIAL_MONOTONIC, N.DIAL_COMPACT) else p.size
                r = pyoracle.calc(typ, x, -1e20, 1e20, p, dim)
            mn, mx = inp["min_response"], inp["max_response"]
            if not math.isnan(mn) and r < mn:
                r = mn
            elif not math.isnan(mx) and r > mx:
                r = mx
            prod *= r
        if not math.isnan(cap) and prod > cap:
            prod = cap
        w[e] = t.base_weight[e] * prod
    return w


def test_event_assembly_matches_python_restatement():
    wl = _workload(0, 400, seed=3)
    t = wl.tables
    orc = pyoracle.Oracle(t, n_threads=2)
    for th in wl.random_points(2, 0.8, seed=4):
        tab = wl.table_values(th)[0]
        orc.set_table_values(tab)
        w, _, _ = orc.propagate(th)
        np.testing.assert_array_equal(w, _python_weights(wl, th, tab))


def test_mirror_fmod_matches_iterative_reflection():
    """DialInputBuffer.cpp:56-70 (fmod) and the reference GPU host reflection
    (CacheParameters.cpp:141-154) give the same mirrored value."""
    lo, hi = -1.5, 1.5
    rng = np.random.default_rng(0)
    for v in rng.uniform(-10, 10, 2000):
        r = hi - lo
        a = abs(math.fmod(v - lo, 2 * r))
        if a > r:
            a = -(a - 2 * r)
        a += lo
        b = v
        while b < lo or b > hi:
            if b < lo:
                b = lo + (lo - b)
            if b > hi:
                b = hi - (b - hi)
        assert abs(a - b) < 1e-12


def _norm_tables(bins, theta_par=0):
    n = bins.size
    inputs = np.zeros(1, dtype=INPUT_DTYPE)
    inputs["parameter"] = theta_par
    inputs["mirror_lower"] = np.nan
    inputs["mirror_range"] = np.nan
    inputs["min_response"] = 0.0
    inputs["max_response"] = np.nan
    return Tables(
        base_weight=np.ones(n), bin=bins.astype(np.int32),
        dial_begin=np.arange(n + 1, dtype=np.int64),
        dial_type=np.full(n, N.DIAL_NORM, np.uint8),
        dial_input=np.zeros(n, np.int32),
        data_begin=np.zeros(n + 1, np.int64), dial_data=np.zeros(1),
        inputs=inputs, table_begin=np.zeros(1, np.int64), n_parameters=1,
        n_bins=int(bins.max()) + 1)


def test_histogram_known_answer():
    # cachedSumsTest.cpp:23-85: 10 bins x 100 unit weights -> 100 per bin
    bins = np.repeat(np.arange(10), 100)
    orc = pyoracle.Oracle(_norm_tables(bins), n_threads=3)
    w, s, e = orc.propagate(np.array([1.0]))
    assert np.all(s == 100.0)
    assert np.all(e == 10.0)


def test_histogram_sums_equal_weight_sums():
    wl = _workload(2, 3000)
    t = wl.tables
    orc = pyoracle.Oracle(t, n_threads=4)
    th = wl.random_points(1, 0.5, seed=1)[0]
    orc.set_table_values(wl.table_values(th))
    w, s, e = orc.propagate(th)
    b = t.bin >= 0
    np.testing.assert_allclose(s.sum(), w[b].sum(), rtol=1e-12)
    np.testing.assert_allclose(s, np.bincount(t.bin[b], w[b], t.n_bins), rtol=1e-12)
    np.testing.assert_allclose(e ** 2, np.bincount(t.bin[b], w[b] ** 2, t.n_bins),
                               rtol=1e-12)


def test_binning_first_match_bruteforce():
    wl = Workload(2, n_events=5000, seed=2)
    got = pyoracle.bin_events(wl.vars, wl.edges)
    v, E = wl.vars, wl.edges
    lo, hi = E[None, :, :, 0], E[None, :, :, 1]
    x = v[:, None, :]
    inside = np.where(lo == hi, x == lo, (lo <= x) & (x < hi)).all(axis=2)
    want = np.where(inside.any(1), inside.argmax(1), -1)
    np.testing.assert_array_equal(got, want)


def test_binning_edge_semantics():
    # [min, max) edges; min == max is an equality condition (Bin.cpp:167-176)
    edges = np.array([[[0, 0], [0.0, 2.5], [-1.0, 0.0]],
                      [[0, 0], [0.0, 2.5], [0.0, 1.0]],
                      [[0, 0], [2.5, 5.0], [-1.0, 1.0]],
                      [[1, 1], [0.0, 5.0], [-1.0, 1.0]]], dtype=float)
    vals = np.array([[0, 5.0, 0.0], [0, 2.5, -1.0], [0, 0.0, -1.0], [0, 1.0, 1.0],
                     [1, 4.99, 0.99], [2, 1.0, 0.0], [0, 0.0, 0.0]])
    np.testing.assert_array_equal(pyoracle.bin_events(vals, edges),
                                  [-1, 2, 0, -1, 3, -1, 1])


# ---- likelihoods ------------------------------------------------------------

def test_poisson_known_answers():
    lib = pyoracle.lib()
    assert lib.oracle_llh_poisson_bin(5.0, 5.0) == 0.0
    assert lib.oracle_llh_poisson_bin(0.0, 3.0) == 6.0
    assert lib.oracle_llh_poisson_bin(2.0, 0.0) == math.inf
    d, p = 7.0, 4.0
    assert abs(lib.oracle_llh_poisson_bin(d, p) - 2 * (p - d + d * math.log(d / p))) < 1e-15


@pytest.mark.parametrize("data,pred,err", [(10.0, 8.0, 1.5), (3.0, 9.0, 4.0),
                                           (100.0, 101.0, 0.3), (1.0, 0.5, 0.4)])
def test_barlow_beta_is_the_profile_minimum(data, pred, err):
    """BB OA2021: chi2 = 2(b p - d + d ln(d/(b p))) + (b-1)^2/f^2 with f =
    err/p, profiled over b; the oracle's closed-form beta must be the
    numerical minimiser and give the same chi2."""
    cfg = pyoracle.llh_config("barlow")
    got = pyoracle.lib().oracle_llh_barlow_bin(data, pred, err, cfg)
    f = err / pred

    def obj(b):
        mc = b * pred
        return 2 * (mc - data + data * (math.log(data) - math.log(mc))) + (b - 1) ** 2 / f ** 2

    res = minimize_scalar(obj, bounds=(1e-6, 50), method="bounded",
                          options={"xatol": 1e-12})
    assert abs(got - res.fun) <= 1e-8 * max(1.0, abs(res.fun))


def test_barlow_options():
    lib = pyoracle.lib()
    base = lib.oracle_llh_barlow_bin(10.0, 8.0, 1.5, pyoracle.llh_config("barlow"))
    nopen = lib.oracle_llh_barlow_bin(10.0, 8.0, 1.5, pyoracle.llh_config(
        "barlow", use_poisson_likelihood=True))
    assert nopen < base
    # no MC uncertainty: the plain Poisson value
    zero = lib.oracle_llh_barlow_bin(7.0, 4.0, 0.0, pyoracle.llh_config("barlow"))
    assert abs(zero - lib.oracle_llh_poisson_bin(7.0, 4.0)) < 1e-14
    # empty data bin: stat = newmc
    assert lib.oracle_llh_barlow_bin(0.0, 4.0, 0.0, pyoracle.llh_config("barlow")) == 8.0
    # expected_value_minimum reimplementation (BarlowBeestonBanff2022.h:157-166)
    cfg = pyoracle.llh_config("barlow", expected_value_minimum=1.0)
    assert lib.oracle_llh_barlow_bin(3.0, 0.5, 0.1, cfg) == 0.0
    assert lib.oracle_llh_barlow_bin(3.0, 0.0, 0.1, cfg) == math.inf


def test_llh_monotone_in_prediction_above_data():
    # unitTest_JointProbability.cpp property: larger excess -> larger LLH
    lib = pyoracle.lib()
    cfg = pyoracle.llh_config("barlow")
    vals = [lib.oracle_llh_barlow_bin(10.0, p, 0.1 * p, cfg) for p in (10, 12, 15, 20, 40)]
    assert all(a < b for a, b in zip(vals, vals[1:]))
    pv = [lib.oracle_llh_poisson_bin(10.0, p) for p in (10, 12, 15, 20, 40)]
    assert all(a < b for a, b in zip(pv, pv[1:]))


def test_total_is_stat_plus_penalty():
    data = np.array([3.0, 4.0, 0.0])
    sums = np.array([2.5, 5.0, 1.0])
    err = np.sqrt(sums)
    theta = np.array([0.5, 1.2, -0.3])
    mean = np.array([0.0, 1.0, 0.0])
    sig = np.array([1.0, 0.1, 0.0])  # sigma 0: no penalty term
    tot, st, pen = pyoracle.llh("poisson", data, sums, err, theta, mean, sig)
    assert pen == 0.5 ** 2 + ((1.2 - 1.0) / 0.1) ** 2
    want = sum(pyoracle.lib().oracle_llh_poisson_bin(d, p) for d, p in zip(data, sums))
    assert st == want
    assert tot == st + pen
