"""The oracle's dial math against golden vectors produced by the reference's
own Calculate*.h headers (tests/golden/make_golden.py), plus the invariants
the reference's unit tests check (two-point linearity, mirror symmetry)."""
import os

import numpy as np
import pytest

from oracle import pyoracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "calc_golden.npz")
MONO, UNIF, GEN, COMP, GRAPH = 1, 2, 3, 4, 5


def _records():
    g = np.load(GOLDEN)
    for i in range(len(g["type"])):
        d = g["data"][g["data_off"][i]:g["data_off"][i + 1]]
        sl = slice(g["x_off"][i], g["x_off"][i + 1])
        yield (int(g["type"][i]), d, int(g["dim"][i]), g["lo"][i], g["hi"][i],
               g["x"][sl], g["expected"][sl])


def test_golden_fixture_covers_every_type():
    types = {r[0] for r in _records()}
    assert types == {MONO, UNIF, GEN, COMP, GRAPH}


def test_calc_bit_exact_against_reference_headers():
    n = 0
    for t, d, dim, lo, hi, xs, exp in _records():
        got = np.array([pyoracle.calc(t, x, lo, hi, d, dim) for x in xs])
        same = (got == exp) | (np.isnan(got) & np.isnan(exp))
        assert same.all(), (t, dim, xs[~same][:3], got[~same][:3], exp[~same][:3])
        n += xs.size
    assert n > 10000


def test_fixture_distinguishes_limiter():
    """The fixture must catch a dropped Fritsch-Carlson limiter: evaluating the
    monotonic records with the unlimited (compact) formula differs somewhere."""
    differs = 0
    for t, d, dim, lo, hi, xs, exp in _records():
        if t != MONO:
            continue
        alt = np.array([pyoracle.calc(COMP, x, lo, hi, d, dim) for x in xs])
        differs += int((alt != exp).sum())
    assert differs > 100


def test_live_reference_library_if_built():
    ref = pyoracle.ref_lib()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = np.random.default_rng(5)
    for t, d, dim, lo, hi, xs, exp in list(_records())[::7]:
        for x in rng.uniform(-6, 6, 16):
            a = pyoracle.calc(t, x, lo, hi, d, dim)
            b = ref.ref_calc(t, float(x), float(lo), float(hi), d.ctypes.data, dim)
            assert a == b or (np.isnan(a) and np.isnan(b))


@pytest.mark.parametrize("t", [MONO, COMP])
def test_two_point_linear(t):
    # unitTest_CalculateMonotonicSpline.cpp:15-37
    d = np.array([0.0, 1.0, 0.0, 1.0])
    for x in np.arange(-1.0, 2.0, 0.1):
        assert abs(pyoracle.calc(t, x, -10, 10, d, 2) - x) < 1e-6


@pytest.mark.parametrize("t", [MONO, COMP])
def test_symmetric_knots_give_symmetric_spline(t):
    # unitTest_CalculateMonotonicSpline.cpp:39-60 and :116-144
    n = 17
    xk = -1.0 + np.arange(n) * 2.0 / (n - 1)
    d = np.concatenate([[-1.0, 2.0 / (n - 1)], np.cos(4.0 * xk)])
    for x in np.arange(-1.5, 1.5, 0.01):
        assert abs(pyoracle.calc(t, x, -10, 10, d, n)
                   - pyoracle.calc(t, -x, -10, 10, d, n)) < 1e-6


def test_graph_antisymmetric():
    # unitTest_CalculateGraph.cpp:139-165 style: an odd graph
    xk = np.array([-2.0, -1.0, 0.0, 1.0, 2.0])
    yk = xk ** 3
    d = np.stack([yk, xk], 1).ravel()
    for x in np.arange(-3.0, 3.0, 0.05):
        assert abs(pyoracle.calc(GRAPH, x, -50, 50, d, 10)
                   + pyoracle.calc(GRAPH, -x, -50, 50, d, 10)) < 1e-12


def test_uniform_and_general_agree_on_uniform_knots():
    """A general spline on uniformly spaced knots is the same Hermite cubic as
    the uniform spline with the same slopes (closed form)."""
    rng = np.random.default_rng(2)
    n = 7
    xk = -3.0 + np.arange(n) * 1.0
    yk = 1 + 0.2 * rng.standard_normal(n)
    sk = rng.standard_normal(n)
    du = np.concatenate([[-3.0, 1.0], np.stack([yk, sk], 1).ravel()])
    dg = np.concatenate([[-3.0, 1.0], np.stack([yk, sk, xk], 1).ravel()])
    for x in rng.uniform(-2.99, 2.99, 200):
        a = pyoracle.calc(UNIF, x, -1e20, 1e20, du, du.size)
        b = pyoracle.calc(GEN, x, -1e20, 1e20, dg, dg.size)
        assert abs(a - b) < 1e-12


GOLDEN2D = os.path.join(os.path.dirname(__file__), "golden", "calc2d_golden.npz")
BILIN, BICUB = 9, 10


def _grids():
    g = np.load(GOLDEN2D)
    for i in range(len(g["type"])):
        d = np.ascontiguousarray(g["data"][g["data_off"][i]:g["data_off"][i + 1]])
        sl = slice(g["p_off"][i], g["p_off"][i + 1])
        yield (int(g["type"][i]), d, g["lo"][i], g["hi"][i], g["x"][sl], g["y"][sl],
               g["expected"][sl])


def _calc2d(t, d, lo, hi, x, y):
    return pyoracle.lib().oracle_calc2d(t, float(x), float(y), lo, hi, d.ctypes.data)


def test_calc2d_bit_exact_against_reference_headers():
    """Bilinear and bicubic grids (2x2 up to 20x7, points inside, on knots and
    outside, with clamps) against CalculateBilinearInterpolation.h and
    CalculateBicubicSpline.h compiled in place."""
    n, types = 0, set()
    for t, d, lo, hi, xs, ys, exp in _grids():
        got = np.array([_calc2d(t, d, lo, hi, x, y) for x, y in zip(xs, ys)])
        same = (got == exp) | (np.isnan(got) & np.isnan(exp))
        assert same.all(), (t, d[:2], xs[~same][:3], ys[~same][:3], got[~same][:3], exp[~same][:3])
        n += xs.size
        types.add(t)
    assert types == {BILIN, BICUB} and n > 3000


def test_calc2d_planted_mistakes_are_caught():
    """A bicubic record read with the bilinear formula, or with its value
    array transposed (y-major instead of x-major), differs from the fixture:
    the vectors pin both."""
    swapped = wrong = 0
    for t, d, lo, hi, xs, ys, exp in _grids():
        if t != BICUB or lo > -1:
            continue
        alt = np.array([_calc2d(BILIN, d, lo, hi, x, y) for x, y in zip(xs, ys)])
        wrong += int((alt != exp).sum())
        nx, ny = int(d[0]), int(d[1])
        if nx == ny:
            continue
        v = d[2 + nx + ny:].reshape(nx, ny)
        dt = np.concatenate([d[:2 + nx + ny], v.T.ravel()])
        alt2 = np.array([_calc2d(t, dt, lo, hi, x, y) for x, y in zip(xs, ys)])
        swapped += int((alt2 != exp).sum())
    assert wrong > 100
    assert swapped > 100


def test_bilinear_reproduces_planes():
    """A plane v = a + b x + c y is reproduced exactly (to rounding) inside
    the grid by both methods (closed form)."""
    rng = np.random.default_rng(5)
    xk = np.sort(rng.uniform(-3, 3, 6))
    yk = np.sort(rng.uniform(-2, 2, 5))
    v = 0.7 + 0.3 * xk[:, None] - 0.2 * yk[None, :]
    d = np.concatenate([[6, 5], xk, yk, v.ravel()])
    for t in (BILIN, BICUB):
        for x, y in rng.uniform([xk[0], yk[0]], [xk[-1], yk[-1]], (50, 2)):
            assert abs(_calc2d(t, d, -1e20, 1e20, x, y) - (0.7 + 0.3 * x - 0.2 * y)) < 1e-12
