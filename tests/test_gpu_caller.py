"""The YAML-driven caller end to end on the device: a configuration with the
reference's keys (tests/caller_fixture.py) loaded, propagated and evaluated,
against the oracle fed the same assembled tables: stat likelihood (1e-10),
penalty (exact, per-set covariance form and the out-of-bounds rule), event
weights copied back per sample (bit-exact), and the data histogram against
an independent binning of the data store."""
import numpy as np
import pytest

from gundam_amd.propagator import load_config
from oracle import pyoracle

import caller_fixture

pytestmark = pytest.mark.gpu


def _oracle_check(prop, li, theta, kind, **opts):
    orc = pyoracle.Oracle(prop.tables, n_threads=8)
    w_ref, s_ref, e_ref = orc.propagate(theta)
    prop.set_parameter_values(theta)
    total = li.propagate_and_eval_likelihood()
    s, e = prop.cache.copy_histogram_contents()
    np.testing.assert_allclose(s, s_ref, rtol=1e-10, atol=0)
    np.testing.assert_allclose(e, e_ref, rtol=1e-10, atol=0)
    _, st_ref, _ = pyoracle.llh(kind, li.data, s_ref, e_ref, **opts)
    assert abs(li.buffer["stat"] - st_ref) <= 1e-10 * abs(st_ref), (li.buffer, st_ref)
    assert li.buffer["penalty"] == pyoracle.penalty(theta, li.penalty_sets)
    assert total == li.buffer["stat"] + li.buffer["penalty"]
    # event weights per sample, in store order, equal the oracle's
    ev = prop.event_weights()
    pos = 0
    for s in prop.samples:
        for _, _, keep, w in ev[s.name]:
            np.testing.assert_array_equal(w, w_ref[pos:pos + keep.size])
            pos += keep.size
    assert pos == prop.tables.n_events
    orc.close()


def test_least_squares_config_with_data_entry(tmp_path):
    print("caller config LeastSquares / TestData", flush=True)
    path = caller_fixture.write(str(tmp_path))
    prop, li = load_config(path)
    # the data histogram: the data store binned independently per sample
    from gundam_amd.eventstore import EventStore
    st = EventStore(str(tmp_path / "data.store"))
    want = np.zeros(prop.n_bins)
    for s in prop.samples:
        vals = np.stack([st.column(v) for v in s.binning.variables()], 1)
        sel = np.ones(st.n_events, bool) if s.selection is None else st.column("C") < 0
        b = pyoracle.bin_events(vals[sel], s.binning.edge_array(s.binning.variables()))
        ok = b >= 0
        want += np.bincount(s.bin_offset + b[ok], weights=np.asarray(st.weight)[sel][ok],
                            minlength=prop.n_bins)
    np.testing.assert_array_equal(li.data, want)
    rng = np.random.default_rng(5)
    prior = prop.prior_values()
    opts = {"lsq_poissonian_approximation": True}
    _oracle_check(prop, li, prior, "least_squares", **opts)
    for _ in range(3):
        th = prior + 0.3 * rng.standard_normal(prior.size)
        th[1] = abs(th[1])  # norm_B inside its limits
        _oracle_check(prop, li, th, "least_squares", **opts)
    # out of bounds: norm_B below its lower limit -> +inf penalty
    th = prior.copy()
    th[1] = -0.1
    prop.set_parameter_values(th)
    li.propagate_and_eval_likelihood()
    assert li.buffer["penalty"] == np.inf


def test_barlow_asimov_config_with_zero_prediction_mask(tmp_path):
    print("caller config BarlowLLH_BANFF_OA2021 / Asimov", flush=True)
    path = caller_fixture.write(str(tmp_path), llh="BarlowLLH_BANFF_OA2021",
                                llh_opts="ignoreBinsWithZeroPredictionAtPrior: true",
                                data_entry="Asimov")
    prop, li = load_config(path)
    # Asimov data = the prediction at the prior
    prop.set_parameter_values(prop.prior_values())
    prop.apply_parameters().get()
    np.testing.assert_array_equal(li.data, prop.cache.hist(0)[0])
    mask = (prop.cache.hist(0)[0] == 0).astype(np.uint8)
    assert li.eval_likelihood() == li.buffer["total"]
    rng = np.random.default_rng(9)
    prior = prop.prior_values()
    for _ in range(3):
        th = prior + 0.2 * rng.standard_normal(prior.size)
        th[1] = abs(th[1])
        _oracle_check(prop, li, th, "barlow", ignore_bins=mask)
