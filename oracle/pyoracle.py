"""TEST INFRASTRUCTURE: ctypes binding of the CPU oracle (liboracle.so).

Only tests/, the smoke check and the bench's CPU-baseline leg import this
module; the product path never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "liboracle.so")
REF_LIB = os.path.join(_HERE, "_ref", "libref_calc.so")

_P, _I32, _I64, _D = C.c_void_p, C.c_int32, C.c_int64, C.c_double


class LlhConfig(C.Structure):
    _fields_ = [("type", C.c_int32), ("use_poisson_likelihood", C.c_int32),
                ("fractional_error_limit", C.c_double),
                ("expected_value_minimum", C.c_double),
                ("lsq_poissonian_approximation", C.c_int32),
                ("bb_no_update_weights", C.c_int32),
                ("nominal_err", C.c_void_p), ("extra_fraction", C.c_void_p),
                ("ignore_bins", C.c_void_p)]


LLH_KINDS = {"poisson": 0, "barlow": 1, "barlow_oa2021": 1, "barlow_classic": 2,
             "barlow_oa2020": 3, "barlow_oa2021_sfgd": 4, "chi2": 5, "least_squares": 6}


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        l = C.CDLL(LIB)
        l.oracle_last_error.restype = C.c_char_p
        l.oracle_create.restype = _P
        l.oracle_create.argtypes = [_P, C.c_int]
        l.oracle_destroy.argtypes = [_P]
        l.oracle_set_table_values.argtypes = [_P, _P, _I64]
        l.oracle_propagate.argtypes = [_P, _P, _P, _P, _P]
        l.oracle_llh_poisson_bin.restype = _D
        l.oracle_llh_poisson_bin.argtypes = [_D, _D]
        l.oracle_llh_barlow_bin.restype = _D
        l.oracle_llh_barlow_bin.argtypes = [_D, _D, _D, C.POINTER(LlhConfig)]
        l.oracle_llh_bin.restype = _D
        l.oracle_llh_bin.argtypes = [C.POINTER(LlhConfig), C.c_int, _D, _D, _D]
        l.oracle_llh.argtypes = [C.POINTER(LlhConfig), C.c_int, _P, _P, _P,
                                 C.c_int, _P, _P, _P, _P, _P, _P]
        l.oracle_bin_events.argtypes = [_I64, C.c_int, _P, C.c_int, _P, _P]
        l.oracle_penalty_set.restype = _D
        l.oracle_penalty_set.argtypes = [C.c_int, _P, _P, _P, _P, _P, _P]
        l.oracle_calc.restype = _D
        l.oracle_calc.argtypes = [C.c_int, _D, _D, _D, _P, C.c_int]
        l.oracle_calc2d.restype = _D
        l.oracle_calc2d.argtypes = [C.c_int, _D, _D, _D, _D, _P]
        _lib = l
    return _lib


def llh_config(kind: str = "poisson", use_poisson_likelihood=False,
               fractional_error_limit=1.0e150, expected_value_minimum=-1.0,
               lsq_poissonian_approximation=False, nominal_err=None,
               extra_fraction=None, ignore_bins=None):
    """The C configuration; per-bin arrays are kept alive on the returned
    object (attribute _keep)."""
    keep = []

    def ptr(a, dt):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=dt)
        keep.append(a)
        return a.ctypes.data
    cfg = LlhConfig(LLH_KINDS[kind], int(use_poisson_likelihood),
                    float(fractional_error_limit), float(expected_value_minimum),
                    int(lsq_poissonian_approximation), int(nominal_err is not None),
                    ptr(nominal_err, np.float64), ptr(extra_fraction, np.float64),
                    ptr(ignore_bins, np.uint8))
    cfg._keep = keep
    return cfg


def llh_bin(kind, b, data, pred, err, **opts):
    """One bin of the likelihood `kind` (see LLH_KINDS)."""
    cfg = llh_config(kind, **opts)
    return lib().oracle_llh_bin(C.byref(cfg), int(b), float(data), float(pred), float(err))


class Oracle:
    """The reference CPU path on a gundam_amd.tables.Tables instance."""

    def __init__(self, tables, n_threads: int = 1):
        self.tables = tables
        t = tables.as_c()
        self._h = lib().oracle_create(C.byref(t), int(n_threads))
        if not self._h:
            raise RuntimeError(lib().oracle_last_error().decode())

    def set_table_values(self, values: np.ndarray) -> None:
        v = np.ascontiguousarray(values, dtype=np.float64).ravel()
        if lib().oracle_set_table_values(self._h, v.ctypes.data, v.size):
            raise ValueError("table values too short")

    def propagate(self, theta: np.ndarray, want_weights: bool = True):
        th = np.ascontiguousarray(theta, dtype=np.float64)
        nb = self.tables.n_bins
        w = np.empty(self.tables.n_events) if want_weights else None
        s = np.empty(nb)
        e = np.empty(nb)
        lib().oracle_propagate(self._h, th.ctypes.data,
                               None if w is None else w.ctypes.data,
                               s.ctypes.data, e.ctypes.data)
        return w, s, e

    def close(self):
        if getattr(self, "_h", None):
            lib().oracle_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def llh(kind, data, sums, sqrt_sum_sq, theta=None, prior_mean=None,
        prior_sigma=None, **opts):
    """(total, stat, penalty) of the likelihood `kind` (see LLH_KINDS)."""
    cfg = llh_config(kind, **opts)
    data = np.ascontiguousarray(data, dtype=np.float64)
    sums = np.ascontiguousarray(sums, dtype=np.float64)
    e = np.ascontiguousarray(sqrt_sum_sq, dtype=np.float64)
    tot, st, pen = C.c_double(), C.c_double(), C.c_double()
    if theta is None:
        n_par, th, pm, ps = 0, None, None, None
    else:
        th = np.ascontiguousarray(theta, dtype=np.float64)
        pm = np.ascontiguousarray(prior_mean, dtype=np.float64)
        ps = np.ascontiguousarray(prior_sigma, dtype=np.float64)
        n_par = th.size
    lib().oracle_llh(C.byref(cfg), data.size, data.ctypes.data, sums.ctypes.data,
                     e.ctypes.data, n_par,
                     None if th is None else th.ctypes.data,
                     None if pm is None else pm.ctypes.data,
                     None if ps is None else ps.ctypes.data,
                     C.byref(tot), C.byref(st), C.byref(pen))
    return tot.value, st.value, pen.value


def penalty_set(theta, parameters, prior, inv_cov=None, lower=None, upper=None) -> float:
    """One parameter set's penalty (oracle_penalty_set)."""
    keep = []

    def ptr(a, dt):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=dt)
        keep.append(a)
        return a.ctypes.data

    par = np.asarray(parameters)
    return lib().oracle_penalty_set(par.size, ptr(theta, np.float64), ptr(par, np.int32),
                                    ptr(prior, np.float64), ptr(inv_cov, np.float64),
                                    ptr(lower, np.float64), ptr(upper, np.float64))


def penalty(theta, sets, prior_mean=None, prior_sigma=None) -> float:
    """The diagonal prior term (as in llh) followed by the sets in order."""
    pen = 0.0
    if prior_sigma is not None:
        pen = llh("poisson", [], [], [], theta, prior_mean, prior_sigma)[2]
    for st in sets:
        pen += penalty_set(theta, st["parameters"], st["prior"], st.get("inv_cov"),
                           st.get("lower"), st.get("upper"))
    return pen


def bin_events(values: np.ndarray, edges: np.ndarray) -> np.ndarray:
    values = np.ascontiguousarray(values, dtype=np.float64)
    edges = np.ascontiguousarray(edges, dtype=np.float64)
    n, nv = values.shape
    out = np.empty(n, dtype=np.int32)
    lib().oracle_bin_events(n, nv, values.ctypes.data, edges.shape[0],
                            edges.ctypes.data, out.ctypes.data)
    return out


def calc(dial_type: int, x: float, lo: float, hi: float, data, dim: int) -> float:
    d = np.ascontiguousarray(data, dtype=np.float64)
    return lib().oracle_calc(int(dial_type), float(x), float(lo), float(hi),
                             d.ctypes.data, int(dim))


_ref = None


def ref_lib():
    """oracle/_ref/libref_calc.so (the reference headers compiled here), or
    None when it has not been built."""
    global _ref
    if _ref is None and os.path.exists(REF_LIB):
        l = C.CDLL(REF_LIB)
        l.ref_calc.restype = _D
        l.ref_calc.argtypes = [C.c_int, _D, _D, _D, _P, C.c_int]
        if hasattr(l, "ref_calc2d"):
            l.ref_calc2d.restype = _D
            l.ref_calc2d.argtypes = [C.c_int, _D, _D, _D, _D, _P]
        _ref = l
    return _ref
