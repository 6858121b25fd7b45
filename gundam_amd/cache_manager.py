"""Python mirror of the reference's ``Cache::Manager`` contract
(src/CacheManager/include/CacheManager.h:56-410) over the C ABI.

    Build(tables)                       -> CacheManager.build
    Fill() -> deferred future           -> CacheManager.fill (returns Future)
    PropagateParameters()               -> CacheManager.propagate_parameters
    CopyHistogramContents()             -> CacheManager.copy_histogram_contents
    CopyEventWeights()                  -> CacheManager.copy_event_weights
    GetResidentMemory()                 -> CacheManager.resident_memory
    SetIsEventWeightCopyEnabled(bool)   -> CacheManager.set_event_weight_copy

The likelihood that the reference evaluates on the host after the copy-back
(JointProbabilityBase.h:46-54, LikelihoodInterface.cpp:165-215) is evaluated
on the device and fetched with ``llh()``.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _native as N
from .tables import Tables


class Future:
    """A deferred result, like the std::launch::deferred future that
    Cache::Manager::Fill returns (CacheManager.cpp:804-809)."""

    def __init__(self, mgr: "CacheManager"):
        self._mgr = mgr

    def valid(self) -> bool:
        return self._mgr is not None

    def get(self) -> bool:
        m, self._mgr = self._mgr, None
        if m is None:
            raise RuntimeError("future already consumed")
        m.sync()
        return True


class CacheManager:
    def __init__(self, device: int = 0, precision: str = "fp64",
                 block_threads: int = 0, tile_words: int = 0,
                 blocks_per_cu: int = 0, ring_buffers: int = 0):
        self._lib = N.lib()
        h = C.c_void_p()
        N.check(self._lib.gcm_create(int(device), C.byref(h)))
        self._h = h
        self.device = device
        prec = {"fp64": N.PRECISION_FP64, "fp32": N.PRECISION_FP32}[precision]
        N.check(self._lib.gcm_set_precision(self._h, prec))
        self.precision = precision
        if block_threads or tile_words or blocks_per_cu or ring_buffers:
            self.set_launch(block_threads, tile_words, blocks_per_cu, ring_buffers)
        self.n_bins = 0
        self.n_parameters = 0
        self.n_events = 0
        self.table_len = 0
        self.k_points = 0
        self.max_points = 1

    # -- configuration ----------------------------------------------------
    def set_launch(self, block_threads=0, tile_words=0, blocks_per_cu=0,
                   ring_buffers=0):
        cfg = N.GcmLaunchConfig(int(block_threads), int(tile_words),
                                int(blocks_per_cu), int(ring_buffers))
        N.check(self._lib.gcm_set_launch_config(self._h, C.byref(cfg)))

    def launch_config(self) -> dict:
        cfg = N.GcmLaunchConfig()
        N.check(self._lib.gcm_get_launch_config(self._h, C.byref(cfg)))
        return {"block_threads": cfg.block_threads, "tile_words": cfg.tile_words,
                "blocks_per_cu": cfg.blocks_per_cu, "ring_buffers": cfg.ring_buffers}

    # -- Build --------------------------------------------------------------
    def build(self, tables: Tables, max_points: int = 1) -> None:
        t = tables.as_c()
        N.check(self._lib.gcm_build(self._h, C.byref(t), int(max_points)))
        self.n_bins = tables.n_bins
        self.n_parameters = tables.n_parameters
        self.n_events = tables.n_events
        self.table_len = tables.table_len
        self.max_points = max_points

    def counts(self) -> dict:
        e, d, t = C.c_int64(), C.c_int64(), C.c_int64()
        N.check(self._lib.gcm_counts(self._h, C.byref(e), C.byref(d), C.byref(t)))
        return {"events": e.value, "dials": d.value, "tiles": t.value}

    def resident_memory(self) -> int:
        b = C.c_int64()
        N.check(self._lib.gcm_resident_bytes(self._h, C.byref(b)))
        return b.value

    def set_event_weight_copy(self, enable: bool) -> None:
        N.check(self._lib.gcm_enable_weights(self._h, int(bool(enable))))

    def set_likelihood(self, kind: str, data: np.ndarray,
                       prior_mean: Optional[np.ndarray] = None,
                       prior_sigma: Optional[np.ndarray] = None,
                       use_poisson_likelihood: bool = False,
                       fractional_error_limit: float = 1.0e150,
                       expected_value_minimum: float = -1.0,
                       lsq_poissonian_approximation: bool = False,
                       nominal_err: Optional[np.ndarray] = None,
                       extra_fraction: Optional[np.ndarray] = None,
                       ignore_bins: Optional[np.ndarray] = None) -> None:
        """kind: one of _native.LLH_TYPES ("poisson", "barlow" (OA2021),
        "barlow_classic", "barlow_oa2020", "barlow_oa2021_sfgd", "chi2",
        "least_squares").  nominal_err switches on BBNoUpdateWeights (OA2021);
        ignore_bins is the ignoreBinsWithZeroPredictionAtPrior mask
        (zero_prediction_mask() computes it)."""
        if kind not in N.LLH_TYPES:
            raise ValueError(f"unknown likelihood {kind!r}")
        f64 = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
        nominal_err, extra_fraction = f64(nominal_err), f64(extra_fraction)
        if ignore_bins is not None:
            ignore_bins = np.ascontiguousarray(ignore_bins, dtype=np.uint8)
        for a in (nominal_err, extra_fraction, ignore_bins):
            if a is not None and a.shape != (self.n_bins,):
                raise ValueError("per-bin array has the wrong size")
        ptr = lambda a: None if a is None else a.ctypes.data
        cfg = N.GcmLlhConfig(
            N.LLH_TYPES[kind], int(use_poisson_likelihood), float(fractional_error_limit),
            float(expected_value_minimum), int(lsq_poissonian_approximation),
            int(nominal_err is not None), ptr(nominal_err), ptr(extra_fraction),
            ptr(ignore_bins))
        data = np.ascontiguousarray(data, dtype=np.float64)
        if data.shape != (self.n_bins,):
            raise ValueError("data histogram has the wrong size")
        pm = ps = None
        if prior_mean is not None:
            pm = np.ascontiguousarray(prior_mean, dtype=np.float64)
            ps = np.ascontiguousarray(prior_sigma, dtype=np.float64)
        self._keep_llh = (data, pm, ps, nominal_err, extra_fraction, ignore_bins)
        N.check(self._lib.gcm_set_likelihood(
            self._h, C.byref(cfg), data.ctypes.data,
            None if pm is None else pm.ctypes.data,
            None if ps is None else ps.ctypes.data))

    def gradient(self, theta: np.ndarray, h: float = 1e-3, params=None,
                 table_values=None):
        """Batched central-difference gradient of the total -2 ln L for a
        minimiser: theta and theta +- h e_j for every j in `params` (default:
        all) go out in launches of up to max_points points (the reference's
        RootMinimizer gets Minuit's numerical gradient, one evalFit per point,
        RootMinimizer.cpp:125-126).  table_values(points) -> [k, table_len]
        when the lookup tables depend on the parameters.  Returns
        (llh(theta), grad)."""
        theta = np.asarray(theta, dtype=np.float64)
        params = np.arange(theta.size) if params is None else np.asarray(params, dtype=int)
        pts = np.repeat(theta[None, :], 2 * params.size + 1, axis=0)
        pts[1 + 2 * np.arange(params.size), params] += h
        pts[2 + 2 * np.arange(params.size), params] -= h
        llh = np.empty(pts.shape[0])
        for i in range(0, pts.shape[0], self.max_points):
            chunk = pts[i:i + self.max_points]
            self.set_parameters(chunk)
            if table_values is not None and self.table_len:
                self.set_table_values(table_values(chunk))
            self.fill().get()
            llh[i:i + chunk.shape[0]] = self.llh()[0][:chunk.shape[0]]
        return llh[0], (llh[1::2] - llh[2::2]) / (2 * h)

    def zero_prediction_mask(self, k: int = 0) -> np.ndarray:
        """Bins whose current prediction (point k) is exactly zero: the
        ignoreBinsWithZeroPredictionAtPrior flags when called at the prior
        (JointProbabilityBase.cpp:20-45)."""
        return (self.hist(k)[0] == 0.0).astype(np.uint8)

    # -- per evaluation -------------------------------------------------------
    def set_parameters(self, theta: np.ndarray) -> None:
        th = np.ascontiguousarray(np.atleast_2d(theta), dtype=np.float64)
        N.check(self._lib.gcm_set_parameters(self._h, th.ctypes.data,
                                             th.shape[1], th.shape[0]))
        self.k_points = th.shape[0]

    def set_table_values(self, values: np.ndarray) -> None:
        v = np.ascontiguousarray(np.atleast_2d(values), dtype=np.float64)
        N.check(self._lib.gcm_set_table_values(self._h, v.ctypes.data,
                                               v.shape[1], v.shape[0]))

    def add_table_update(self, code_object: str, kernel: str, jobs) -> None:
        """Refresh lookup tables on the device at every fill with a gfx950
        update function (include/gcm_table_update.h) instead of a host
        callback + upload.  jobs: iterable of (table, [input indices],
        user_id)."""
        jobs = list(jobs)
        keep = [np.ascontiguousarray(j[1], dtype=np.int32) for j in jobs]
        arr = (N.GcmTableJob * len(jobs))()
        for a, (t, _, uid), ins in zip(arr, jobs, keep):
            a.table, a.n_inputs, a.inputs, a.user_id = int(t), ins.size, ins.ctypes.data, int(uid)
        N.check(self._lib.gcm_add_table_update(self._h, code_object.encode(), kernel.encode(),
                                               len(jobs), arr))

    def clear_table_updates(self) -> None:
        N.check(self._lib.gcm_clear_table_updates(self._h))

    def device_table_values(self, k: int = 0) -> np.ndarray:
        """The lookup tables point k was last evaluated with."""
        out = np.zeros(max(1, self.table_len))
        N.check(self._lib.gcm_get_table_values(self._h, int(k), out.ctypes.data))
        return out[:self.table_len]

    def set_penalty_sets(self, sets) -> None:
        """Correlated prior penalties (gcm_penalty_set): each set is a dict
        with 'parameters', 'prior' and optional 'inv_cov' ([n][n]), 'lower',
        'upper'; they are added after the diagonal prior term in set order."""
        sets = list(sets)
        keep, arr = [], (N.GcmPenaltySet * max(1, len(sets)))()

        def ptr(a, dt):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=dt)
            keep.append(a)
            return a.ctypes.data

        for a, st in zip(arr, sets):
            par = np.asarray(st["parameters"])
            a.n = par.size
            a.parameters = ptr(par, np.int32)
            a.prior = ptr(st["prior"], np.float64)
            a.inv_cov = ptr(st.get("inv_cov"), np.float64)
            a.lower = ptr(st.get("lower"), np.float64)
            a.upper = ptr(st.get("upper"), np.float64)
        N.check(self._lib.gcm_set_penalty_sets(self._h, len(sets), arr))

    def fill(self, stream: Optional[int] = None) -> Future:
        N.check(self._lib.gcm_fill(self._h, stream))
        return Future(self)

    def sync(self) -> None:
        N.check(self._lib.gcm_sync(self._h))

    def propagate_parameters(self, theta, table_values=None) -> bool:
        self.set_parameters(theta)
        if table_values is not None and self.table_len:
            self.set_table_values(table_values)
        return self.fill().get()

    # -- fetch ----------------------------------------------------------------
    def hist(self, k: int = 0):
        """(sum w, sum w^2) per bin for point k."""
        s = np.empty(self.n_bins)
        s2 = np.empty(self.n_bins)
        N.check(self._lib.gcm_get_hist(self._h, int(k), s.ctypes.data,
                                       s2.ctypes.data))
        return s, s2

    def copy_histogram_contents(self, k: int = 0):
        """(sumWeights, sqrtSumSqWeights) per bin, as
        CacheSampleHistFiller::copyHistogram fills them
        (CacheSampleHistFiller.cpp:9-19)."""
        s, s2 = self.hist(k)
        return s, np.sqrt(s2)

    def validate_histogram_contents(self, sums: np.ndarray, sqrt_sum_w2: np.ndarray,
                                    k: int = 0, fraction_of_bits_lost: float = 0.3,
                                    verbose: bool = False) -> bool:
        """Cache::Manager::ValidateHistogramContents: the device's bin sums and
        sqrt(sum w^2) of point k against CPU-filled ones (GundamUtils::almostEqual
        rule); True when every bin matches."""
        s = np.ascontiguousarray(sums, dtype=np.float64)
        e = np.ascontiguousarray(sqrt_sum_w2, dtype=np.float64)
        if s.shape != (self.n_bins,) or e.shape != (self.n_bins,):
            raise ValueError("per-bin array has the wrong size")
        bad = C.c_int32()
        N.check(self._lib.gcm_validate_hist(self._h, int(k), s.ctypes.data, e.ctypes.data,
                                            float(fraction_of_bits_lost), int(bool(verbose)),
                                            C.byref(bad)))
        self.last_validation_failures = bad.value
        return bad.value == 0

    def copy_event_weights(self, out: Optional[np.ndarray] = None):
        if out is None:
            out = np.full(self.n_events, np.nan)
        N.check(self._lib.gcm_get_weights(self._h, out.ctypes.data))
        return out

    def llh(self):
        k = max(1, self.k_points)
        tot, st, pen = np.empty(k), np.empty(k), np.empty(k)
        N.check(self._lib.gcm_get_llh(self._h, tot.ctypes.data, st.ctypes.data,
                                      pen.ctypes.data))
        return tot, st, pen

    def enable_graph(self, enable: bool = True) -> None:
        """Capture each point count's evaluation in a hipGraph and replay it."""
        N.check(self._lib.gcm_enable_graph(self._h, int(bool(enable))))

    def enable_timing(self, enable: bool = True) -> None:
        N.check(self._lib.gcm_enable_timing(self._h, int(bool(enable))))

    def timing(self) -> dict:
        f, z, n = C.c_double(), C.c_double(), C.c_int64()
        N.check(self._lib.gcm_get_timing(self._h, C.byref(f), C.byref(z),
                                         C.byref(n)))
        return {"fill_ms": f.value, "finalize_ms": z.value, "count": n.value}

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.gcm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
