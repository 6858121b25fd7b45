"""Host-side dial tables (numpy) in the layout of ``gcm_tables`` and the
seeded synthetic workloads that fill them.

``Tables`` is the Python twin of the reference's EventDialCache content as the
Cache::Manager reads it in Update (CacheManager.cpp:503-680): per event a base
weight, a bin and a dial list; per dial a type, an input (parameter reference
plus response supervisor) and the reference's record payload.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import _native as N

INPUT_DTYPE = np.dtype([("parameter", "<i4"), ("flags", "<i4"),
                        ("mirror_lower", "<f8"), ("mirror_range", "<f8"),
                        ("min_response", "<f8"), ("max_response", "<f8")],
                       align=True)
assert INPUT_DTYPE.itemsize == C.sizeof(N.GcmInput)


def _ptr(a):
    return None if a is None else a.ctypes.data


@dataclass
class Tables:
    base_weight: np.ndarray
    bin: np.ndarray
    dial_begin: np.ndarray
    dial_type: np.ndarray
    dial_input: np.ndarray
    data_begin: np.ndarray
    dial_data: np.ndarray
    inputs: np.ndarray
    table_begin: np.ndarray
    n_parameters: int
    n_bins: int
    reweight_cap: float = float("nan")
    dial_input2: Optional[np.ndarray] = None  # [n_dials] y input of 2-D dials
    _keep: list = field(default_factory=list, repr=False)

    @property
    def n_events(self) -> int:
        return int(self.base_weight.shape[0])

    @property
    def n_dials(self) -> int:
        return int(self.dial_type.shape[0])

    @property
    def n_tables(self) -> int:
        return int(self.table_begin.shape[0]) - 1

    @property
    def table_len(self) -> int:
        return int(self.table_begin[-1]) if self.n_tables > 0 else 0

    def check_dtypes(self) -> None:
        for name, dt in [("base_weight", np.float64), ("bin", np.int32),
                         ("dial_begin", np.int64), ("dial_type", np.uint8),
                         ("dial_input", np.int32), ("data_begin", np.int64),
                         ("dial_data", np.float64), ("table_begin", np.int64)]:
            a = getattr(self, name)
            if a.dtype != dt or not a.flags.c_contiguous:
                raise TypeError(f"{name} must be contiguous {np.dtype(dt)}")
        if self.inputs.dtype != INPUT_DTYPE:
            raise TypeError("inputs must use INPUT_DTYPE")
        if self.dial_input2 is not None and (self.dial_input2.dtype != np.int32
                                             or not self.dial_input2.flags.c_contiguous):
            raise TypeError("dial_input2 must be contiguous int32")

    def as_c(self) -> N.GcmTables:
        """The gcm_tables struct; valid while self is alive."""
        self.check_dtypes()
        t = N.GcmTables()
        t.n_events = self.n_events
        t.base_weight = _ptr(self.base_weight)
        t.bin = _ptr(self.bin)
        t.dial_begin = _ptr(self.dial_begin)
        t.n_dials = self.n_dials
        t.dial_type = _ptr(self.dial_type)
        t.dial_input = _ptr(self.dial_input)
        t.data_begin = _ptr(self.data_begin)
        t.dial_data = _ptr(self.dial_data)
        t.n_inputs = int(self.inputs.shape[0])
        t.inputs = _ptr(self.inputs)
        t.n_tables = self.n_tables
        t.table_begin = _ptr(self.table_begin)
        t.n_parameters = int(self.n_parameters)
        t.n_bins = int(self.n_bins)
        t.reweight_cap = float(self.reweight_cap)
        t.dial_input2 = _ptr(self.dial_input2)
        return t

    # -- reference record sizes, for the algorithmic byte count ------------
    def algorithmic_bytes(self, word_bytes: int = 8) -> int:
        """Bytes one LLH evaluation must move at minimum:
        sum over dials of (w*L_d + 2) + 16 per binned event + 8 per table
        entry + 8 per parameter + 16 per bin, with L_d the reference record
        payload (monotonic/compact n+2, uniform 2n+2, general 3n+2, graph 2n,
        norm 0, shift 1 word, tabulated 12 B (8 B in FP32), kriged 8k+4 B,
        bilinear/bicubic 2+nx+ny+nx*ny words plus a second 2-byte parameter
        index)."""
        t = self.dial_type
        ln = np.diff(self.data_begin)
        b = np.zeros(t.shape[0], dtype=np.int64)
        spl = np.isin(t, [N.DIAL_MONOTONIC, N.DIAL_COMPACT, N.DIAL_UNIFORM,
                          N.DIAL_GENERAL, N.DIAL_GRAPH, N.DIAL_SHIFT])
        b[spl] = ln[spl] * word_bytes
        grid = np.isin(t, [N.DIAL_BILINEAR, N.DIAL_BICUBIC])
        b[grid] = ln[grid] * word_bytes + 2
        tab = t == N.DIAL_TABULATED
        b[tab] = 12 if word_bytes == 8 else 8
        kr = t == N.DIAL_KRIGED
        k = (ln[kr] - 1) // 2
        b[kr] = 8 * k + 4
        b += 2
        nb = int((self.bin >= 0).sum())
        return int(b.sum() + 16 * nb + 8 * self.table_len
                   + 8 * self.n_parameters + 16 * self.n_bins)


# generator parameter kinds that own a lookup table (synth.cpp ParType)
PAR_TABULATED, PAR_KRIGED = 6, 7


class Workload:
    """A seeded synthetic workload from libgcm_synth (see synth.cpp)."""

    def __init__(self, workload: int, n_events: int = 0, seed: int = 0,
                 threads: int = 16):
        s = N.synth()
        self._h = s.gcm_synth_create(int(workload), int(n_events), int(seed),
                                     int(threads))
        if not self._h:
            raise RuntimeError(s.gcm_synth_last_error().decode())
        v = N.GcmSynthView()
        s.gcm_synth_view_get(self._h, C.byref(v))
        self.view = v
        self.workload = workload
        ne, nd, ndata = v.n_events, v.n_dials, v.n_data

        def arr(ptr, ctype, n, shape=None):
            if n == 0:
                return np.zeros(0 if shape is None else shape,
                                dtype=np.dtype(ctype))
            a = np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ctype)), shape=(n,))
            return a if shape is None else a.reshape(shape)

        self.vars = arr(v.vars, C.c_double, ne * v.n_vars, (ne, v.n_vars))
        self.edges = arr(v.edges, C.c_double, v.n_bins * v.n_vars * 2,
                         (v.n_bins, v.n_vars, 2))
        self.prior_mean = arr(v.prior_mean, C.c_double, v.n_parameters)
        self.prior_sigma = arr(v.prior_sigma, C.c_double, v.n_parameters)
        self.par_type = arr(v.par_type, C.c_int32, v.n_parameters)
        inputs = np.ctypeslib.as_array(
            C.cast(v.inputs, C.POINTER(C.c_uint8)),
            shape=(v.n_inputs * INPUT_DTYPE.itemsize,)).view(INPUT_DTYPE)
        self.tables = Tables(
            base_weight=arr(v.base_weight, C.c_double, ne),
            bin=arr(v.bin, C.c_int32, ne),
            dial_begin=arr(v.dial_begin, C.c_int64, ne + 1),
            dial_type=arr(v.dial_type, C.c_uint8, nd),
            dial_input=arr(v.dial_input, C.c_int32, nd),
            data_begin=arr(v.data_begin, C.c_int64, nd + 1),
            dial_data=arr(v.dial_data, C.c_double, ndata),
            inputs=inputs,
            table_begin=arr(v.table_begin, C.c_int64, v.n_tables + 1),
            n_parameters=v.n_parameters, n_bins=v.n_bins,
            reweight_cap=v.reweight_cap,
            dial_input2=arr(v.dial_input2, C.c_int32, nd))
        self.tables._keep.append(self)

    @property
    def n_parameters(self) -> int:
        return self.view.n_parameters

    def table_values(self, theta: np.ndarray) -> np.ndarray:
        """Host factory callback: lookup tables for each row of theta."""
        theta = np.ascontiguousarray(np.atleast_2d(theta), dtype=np.float64)
        out = np.zeros((theta.shape[0], max(1, self.tables.table_len)))
        for k in range(theta.shape[0]):
            N.synth().gcm_synth_table_values(self._h, theta[k].ctypes.data,
                                             out[k].ctypes.data)
        return out[:, :self.tables.table_len].copy()

    def table_jobs(self) -> list:
        """Device table-update jobs equivalent to table_values: table t is the
        t-th tabulated/kriged parameter in parameter order, read through the
        first input of that parameter; one job per table for the code object
        table_update_cos.hsaco (kernel gcm_synth_cos_table)."""
        pars = [p for p in range(self.n_parameters)
                if self.par_type[p] in (PAR_TABULATED, PAR_KRIGED)]
        assert len(pars) == self.tables.n_tables
        first_input = {}
        for i, p in enumerate(self.tables.inputs["parameter"]):
            first_input.setdefault(int(p), i)
        return [(t, [first_input[p]], 0) for t, p in enumerate(pars)]

    def nominal(self) -> np.ndarray:
        return self.prior_mean.copy()

    def random_points(self, k: int, scale: float, seed: int) -> np.ndarray:
        """Parameter points: prior mean + scale*N(0,1) for shape parameters and
        prior mean + scale*N(0,1) for norm parameters as well."""
        rng = np.random.default_rng(seed)
        return self.prior_mean[None, :] + scale * rng.standard_normal(
            (k, self.n_parameters))

    def prior_points(self, k: int, seed: int) -> np.ndarray:
        """Parameter points drawn from the priors: prior mean + prior sigma *
        N(0,1) (shape parameters N(0,1), norm parameters 1 + 0.1 N(0,1))."""
        rng = np.random.default_rng(seed)
        return self.prior_mean[None, :] + self.prior_sigma[None, :] * rng.standard_normal(
            (k, self.n_parameters))

    def bin_events_on_device(self, device: int = 0) -> None:
        v = self.view
        N.check(N.lib().gcm_bin_events(device, v.n_events, v.n_vars,
                                       self.vars.ctypes.data, v.n_bins,
                                       self.edges.ctypes.data,
                                       self.tables.bin.ctypes.data))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                N.synth().gcm_synth_destroy(h)
            except Exception:
                pass
            self._h = None
