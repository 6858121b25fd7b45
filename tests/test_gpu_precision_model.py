"""The precision-study model (6 samples of 60x50 bins, 500 parameters, 15
dials per event of the mixed types; the generator's 2M-event scaling of the
20M-event workload, which keeps the test within seconds on the oracle):
the FP64 path against the oracle at 10 points drawn from the priors, as the
precision study validates its reference, and the FP32 mode against the FP64
path on the same points.

Bars: FP64 bin sums and sqrt(sum w^2) within 1e-10 relative of the oracle,
point 0's event weights bit-exact; FP32 per-bin relative error below 2e-4
(measured at these seeded points: at most 1.4e-5; the precision study's
20M-event run 5.3e-5 over 128 points) and the Barlow-Beeston LLH within 1e-5
relative."""
import numpy as np
import pytest

from gundam_amd import CacheManager, Workload
from oracle import pyoracle

pytestmark = pytest.mark.gpu


def test_precision_model_fp64_matches_oracle_fp32_bounded():
    print("precision model 2M events K=10", flush=True)
    wl = Workload(4, n_events=2_000_000, seed=0)
    t = wl.tables
    wl.bin_events_on_device(0)
    assert t.n_bins == 18000 and wl.n_parameters == 500
    K = 10
    pts = wl.prior_points(K, seed=4)
    tabs = wl.table_values(pts)
    cms = {p: CacheManager(0, precision=p) for p in ("fp64", "fp32")}
    nominal = wl.nominal()
    for cm in cms.values():
        cm.build(t, max_points=K)
    cms["fp64"].propagate_parameters(nominal, wl.table_values(nominal))
    data = cms["fp64"].hist(0)[0].copy()
    for cm in cms.values():
        cm.set_likelihood("barlow", data, wl.prior_mean, wl.prior_sigma)
        cm.set_parameters(pts)
        cm.set_table_values(tabs)
        cm.fill().get()
    orc = pyoracle.Oracle(t, n_threads=16)
    cm64, cm32 = cms["fp64"], cms["fp32"]
    llh64, llh32 = cm64.llh()[0], cm32.llh()[0]
    for k in range(K):
        orc.set_table_values(tabs[k])
        w_ref, s_ref, e_ref = orc.propagate(pts[k], want_weights=(k == 0))
        s, s2 = cm64.hist(k)
        np.testing.assert_allclose(s, s_ref, rtol=1e-10, atol=0)
        np.testing.assert_allclose(np.sqrt(s2), e_ref, rtol=1e-10, atol=0)
        f, _ = cm32.hist(k)
        nz = s != 0
        rel = np.max(np.abs(f[nz] - s[nz]) / np.abs(s[nz]))
        print(f"point {k}: FP32 max per-bin rel err {rel:.3g}, "
              f"|dLLH| {abs(llh32[k] - llh64[k]):.3g}", flush=True)
        assert rel < 2e-4
        assert np.all(f[~nz] == 0)
        assert abs(llh32[k] - llh64[k]) <= 1e-5 * abs(llh64[k])
        if k == 0:
            cm64.set_event_weight_copy(True)
            cm64.set_parameters(pts[0])
            cm64.set_table_values(tabs[0])
            cm64.fill().get()
            w = cm64.copy_event_weights(np.full(t.n_events, np.nan))
            binned = t.bin >= 0
            np.testing.assert_array_equal(w[binned], w_ref[binned])
            cm64.set_parameters(pts)
            cm64.set_table_values(tabs)
            cm64.fill().get()
    orc.close()
    for cm in cms.values():
        cm.close()
