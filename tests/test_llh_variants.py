"""The oracle's likelihood variants (oracle/oracle.cpp) pinned without the
reference's code: each Barlow-Beeston form against the numerical minimiser of
its profiled chi2, Chi2 and LeastSquares against their closed forms, the
continuity/monotonicity properties of the reference's
tests/GTests/unitTest_JointProbability.cpp, and the bin mask and
BBNoUpdateWeights switches against their definitions."""
import math

import numpy as np
import pytest
from scipy.optimize import minimize_scalar

from oracle import pyoracle

CASES = [(10.0, 8.0, 1.5), (3.0, 9.0, 4.0), (100.0, 101.0, 0.3), (1.0, 0.5, 0.4),
         (0.0, 4.0, 1.0)]


def _profile_min(data, pred, f2):
    """min over beta of 2(beta p - d + d ln(d/(beta p))) + (beta-1)^2 / f2
    (d = 0: 2 beta p + (beta-1)^2 / f2)."""
    def obj(b):
        mc = b * pred
        stat = mc if data == 0 else mc - data + data * (math.log(data) - math.log(mc))
        return 2 * stat + (b - 1) ** 2 / f2
    res = minimize_scalar(obj, bounds=(1e-9, 60), method="bounded",
                          options={"xatol": 1e-13})
    return res.fun


@pytest.mark.parametrize("data,pred,err", CASES)
def test_classic_barlow_is_the_profile_minimum(data, pred, err):
    """BarlowBeeston.h: relative variance err / pred^2 (err used as given)."""
    got = pyoracle.llh_bin("barlow_classic", 0, data, pred, err)
    want = _profile_min(data, pred, err / pred ** 2)
    assert abs(got - want) <= 1e-8 * max(1.0, abs(want))


@pytest.mark.parametrize("data,pred,err", CASES)
def test_oa2020_is_the_profile_minimum(data, pred, err):
    """BarlowBeestonBanff2020.h: fractional error sqrt(err) / pred."""
    got = pyoracle.llh_bin("barlow_oa2020", 0, data, pred, err)
    want = _profile_min(data, pred, err / pred ** 2)
    assert abs(got - want) <= 1e-8 * max(1.0, abs(want))


@pytest.mark.parametrize("data,pred,err", CASES)
@pytest.mark.parametrize("frac", [0.0, 0.02, 0.1])
def test_sfgd_is_the_profile_minimum(data, pred, err, frac):
    """BarlowBeestonBanff2022Sfgd.h: fractional error (err + pred frac) / pred."""
    got = pyoracle.llh_bin("barlow_oa2021_sfgd", 0, data, pred, err, extra_fraction=[frac])
    want = _profile_min(data, pred, ((err + pred * frac) / pred) ** 2)
    assert abs(got - want) <= 1e-8 * max(1.0, abs(want))


def test_classic_and_oa2020_agree():
    """Both use the variance err / pred^2: the same profiled chi2."""
    for d, p, e in CASES:
        a = pyoracle.llh_bin("barlow_classic", 0, d, p, e)
        b = pyoracle.llh_bin("barlow_oa2020", 0, d, p, e)
        assert abs(a - b) <= 1e-12 * max(1.0, abs(a))


def test_chi2_and_least_squares_closed_forms():
    assert pyoracle.llh_bin("chi2", 0, 3.0, 2.0, 1.0) == 0.5
    assert pyoracle.llh_bin("chi2", 0, 3.0, 0.0, 1.0) == math.inf
    assert pyoracle.llh_bin("least_squares", 0, 3.0, 2.0, 1.0) == 1.0
    # Poissonian approximation divides by data / 2 only when data > 1
    assert pyoracle.llh_bin("least_squares", 0, 4.0, 2.0, 1.0,
                            lsq_poissonian_approximation=True) == 2.0
    assert pyoracle.llh_bin("least_squares", 0, 0.5, 2.0, 1.0,
                            lsq_poissonian_approximation=True) == 2.25


def _sweep(kind, err_of, data=20.0, **opts):
    """The reference unit test's sweeps: the prediction moves from the data
    down by x0.9 until it stops changing, and up by x1.01 to 20x; the value
    must never be NaN and never decrease (0.99999 slack for the OA2021
    downward sweep, as there)."""
    slack = 0.99999 if kind == "barlow" else 1.0
    last, pred, lastpred = -1.0, data, 0.0
    while pred != lastpred:
        v = pyoracle.llh_bin(kind, 0, data, pred, err_of(pred), **opts)
        assert not math.isnan(v), (kind, pred)
        assert v >= slack * last, (kind, pred, v, last)
        last, lastpred, pred = v, pred, pred * 0.9
    last, pred = -1.0, data
    while pred < 20 * data:
        v = pyoracle.llh_bin(kind, 0, data, pred, err_of(pred), **opts)
        assert not math.isnan(v), (kind, pred)
        assert v >= last, (kind, pred, v, last)
        last, pred = v, pred * 1.01


@pytest.mark.parametrize("kind", ["least_squares", "chi2", "poisson", "barlow",
                                  "barlow_classic"])
def test_continuity_sweeps(kind):
    """The kinds the reference's unit test sweeps, with its error model."""
    frac = 1.0 / math.sqrt(20.0)
    _sweep(kind, lambda p: frac * p)


@pytest.mark.parametrize("kind", ["barlow_oa2020", "barlow_oa2021_sfgd"])
def test_banff_forms_monotone_away_from_data(kind):
    """OA2020 and SFGD are not swept by the reference (they throw on the
    underflowing predictions of the full sweep): a moderate range here."""
    frac = 1.0 / math.sqrt(20.0)
    for preds in (20.0 * 0.9 ** np.arange(60), 20.0 * 1.01 ** np.arange(300)):
        v = [pyoracle.llh_bin(kind, 0, 20.0, p, frac * p, extra_fraction=[0.04]) for p in preds]
        assert not np.any(np.isnan(v))
        assert all(b >= a for a, b in zip(v, v[1:]))


def test_oa2021_fixed_error_sweep():
    """unitTest_JointProbability.cpp OA2021_NoUpdateErr: the error stays at its
    value for the data, and a zero-data sweep from a zero prediction."""
    err0 = 20.0 / math.sqrt(10.0 * 20.0)
    _sweep("barlow", lambda p: err0)
    last, pred = -1.0, 0.0
    while pred < 20.0:
        v = pyoracle.llh_bin("barlow", 0, 0.0, pred, err0)
        assert not math.isnan(v) and v >= last
        last, pred = v, pred + np.finfo(float).tiny + 1.01 * pred


def test_bin_mask_and_nominal_errors():
    data = np.array([3.0, 4.0, 0.0, 7.0])
    sums = np.array([2.5, 5.0, 0.0, 6.0])
    err = np.array([0.5, 1.0, 0.0, 2.0])
    mask = (sums == 0).astype(np.uint8)
    _, st, _ = pyoracle.llh("poisson", data, sums, err, ignore_bins=mask)
    want = sum(pyoracle.llh_bin("poisson", 0, d, p, e)
               for d, p, e, m in zip(data, sums, err, mask) if not m)
    assert st == want
    _, unmasked, _ = pyoracle.llh("poisson", data, sums, err)
    assert unmasked == math.inf and math.isfinite(st)
    # BBNoUpdateWeights: the nominal sqrt(sum w^2) replaces the current one
    nominal = np.array([0.7, 1.1, 0.2, 1.5])
    _, a, _ = pyoracle.llh("barlow", data, sums + 1, err, nominal_err=nominal)
    _, b, _ = pyoracle.llh("barlow", data, sums + 1, nominal)
    assert a == b
