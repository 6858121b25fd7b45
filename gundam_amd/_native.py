"""ctypes bindings of the C ABI in include/gcm.h (libgcm.so) and of the host
workload generator (libgcm_synth.so).

The HIP library is the product path: if it is missing, loading raises
immediately -- there is no CPU fallback behind this module.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgcm.so")
# device table function of the synthetic workloads (gcm_add_table_update)
TABLE_UPDATE_COS = os.path.join(_HERE, "table_update_cos.hsaco")
TABLE_UPDATE_COS_KERNEL = "gcm_synth_cos_table"
SYNTH_PATH = os.path.join(_HERE, "libgcm_synth.so")

# status codes
GCM_OK = 0
GCM_ERR_ARG = 1
GCM_ERR_STATE = 2
GCM_ERR_HIP = 3
GCM_ERR_UNSUPPORTED = 4
GCM_ERR_NUMERIC = 5
GCM_ERR_NOMEM = 6

# dial types
DIAL_NORM = 0
DIAL_MONOTONIC = 1
DIAL_UNIFORM = 2
DIAL_GENERAL = 3
DIAL_COMPACT = 4
DIAL_GRAPH = 5
DIAL_TABULATED = 6
DIAL_KRIGED = 7
DIAL_SHIFT = 8
DIAL_BILINEAR = 9
DIAL_BICUBIC = 10
DIAL_NAMES = ["norm", "monotonic", "uniform", "general", "compact", "graph",
              "tabulated", "kriged", "shift", "bilinear", "bicubic"]

INPUT_NO_EXTRAPOLATION = 1
PRECISION_FP64 = 0
PRECISION_FP32 = 1
LLH_POISSON = 0
LLH_BARLOW_OA2021 = 1
LLH_BARLOW_BEESTON = 2
LLH_BARLOW_OA2020 = 3
LLH_BARLOW_OA2021_SFGD = 4
LLH_CHI_SQUARED = 5
LLH_LEAST_SQUARES = 6
# names of the reference's JointProbability plugins (JointProbability.cpp)
LLH_TYPES = {"poisson": LLH_POISSON, "barlow": LLH_BARLOW_OA2021,
             "barlow_oa2021": LLH_BARLOW_OA2021, "barlow_classic": LLH_BARLOW_BEESTON,
             "barlow_oa2020": LLH_BARLOW_OA2020, "barlow_oa2021_sfgd": LLH_BARLOW_OA2021_SFGD,
             "chi2": LLH_CHI_SQUARED, "least_squares": LLH_LEAST_SQUARES}


class GcmInput(C.Structure):
    _fields_ = [("parameter", C.c_int32), ("flags", C.c_int32),
                ("mirror_lower", C.c_double), ("mirror_range", C.c_double),
                ("min_response", C.c_double), ("max_response", C.c_double)]


class GcmTables(C.Structure):
    _fields_ = [("n_events", C.c_int64),
                ("base_weight", C.c_void_p), ("bin", C.c_void_p),
                ("dial_begin", C.c_void_p),
                ("n_dials", C.c_int64),
                ("dial_type", C.c_void_p), ("dial_input", C.c_void_p),
                ("data_begin", C.c_void_p), ("dial_data", C.c_void_p),
                ("n_inputs", C.c_int32), ("inputs", C.c_void_p),
                ("n_tables", C.c_int32), ("table_begin", C.c_void_p),
                ("n_parameters", C.c_int32), ("n_bins", C.c_int32),
                ("reweight_cap", C.c_double), ("dial_input2", C.c_void_p)]


class GcmLlhConfig(C.Structure):
    _fields_ = [("type", C.c_int32), ("use_poisson_likelihood", C.c_int32),
                ("fractional_error_limit", C.c_double),
                ("expected_value_minimum", C.c_double),
                ("lsq_poissonian_approximation", C.c_int32),
                ("bb_no_update_weights", C.c_int32),
                ("nominal_err", C.c_void_p), ("extra_fraction", C.c_void_p),
                ("ignore_bins", C.c_void_p)]


class GcmPenaltySet(C.Structure):
    _fields_ = [("n", C.c_int32), ("parameters", C.c_void_p), ("prior", C.c_void_p),
                ("inv_cov", C.c_void_p), ("lower", C.c_void_p), ("upper", C.c_void_p)]


class GcmTableJob(C.Structure):
    _fields_ = [("table", C.c_int32), ("n_inputs", C.c_int32), ("inputs", C.c_void_p),
                ("user_id", C.c_int32)]


class GcmLaunchConfig(C.Structure):
    _fields_ = [("block_threads", C.c_int32), ("tile_words", C.c_int32),
                ("blocks_per_cu", C.c_int32), ("ring_buffers", C.c_int32)]


class GcmSynthView(C.Structure):
    _fields_ = [("n_events", C.c_int64), ("n_vars", C.c_int32),
                ("n_bins", C.c_int32), ("n_parameters", C.c_int32),
                ("n_inputs", C.c_int32), ("n_tables", C.c_int32),
                ("n_samples", C.c_int32),
                ("n_dials", C.c_int64), ("n_data", C.c_int64),
                ("table_len", C.c_int64), ("reweight_cap", C.c_double),
                ("vars", C.c_void_p), ("edges", C.c_void_p),
                ("base_weight", C.c_void_p), ("bin", C.c_void_p),
                ("dial_begin", C.c_void_p), ("dial_type", C.c_void_p),
                ("dial_input", C.c_void_p), ("data_begin", C.c_void_p),
                ("dial_data", C.c_void_p), ("inputs", C.c_void_p),
                ("table_begin", C.c_void_p), ("prior_mean", C.c_void_p),
                ("prior_sigma", C.c_void_p), ("par_type", C.c_void_p),
                ("dial_input2", C.c_void_p)]


_P = C.c_void_p
_I32 = C.c_int32
_I64 = C.c_int64
_D = C.c_double

# Every symbol declared in include/gcm.h with its (restype, argtypes).
GCM_SIGNATURES = {
    "gcm_last_error": (C.c_char_p, []),
    "gcm_version": (C.c_int, []),
    "gcm_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "gcm_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "gcm_destroy": (C.c_int, [_P]),
    "gcm_set_precision": (C.c_int, [_P, C.c_int]),
    "gcm_set_launch_config": (C.c_int, [_P, C.POINTER(GcmLaunchConfig)]),
    "gcm_get_launch_config": (C.c_int, [_P, C.POINTER(GcmLaunchConfig)]),
    "gcm_build": (C.c_int, [_P, C.POINTER(GcmTables), _I32]),
    "gcm_resident_bytes": (C.c_int, [_P, C.POINTER(_I64)]),
    "gcm_counts": (C.c_int, [_P, C.POINTER(_I64), C.POINTER(_I64), C.POINTER(_I64)]),
    "gcm_set_parameters": (C.c_int, [_P, _P, _I32, _I32]),
    "gcm_set_table_values": (C.c_int, [_P, _P, _I64, _I32]),
    "gcm_set_likelihood": (C.c_int, [_P, C.POINTER(GcmLlhConfig), _P, _P, _P]),
    "gcm_set_penalty_sets": (C.c_int, [_P, _I32, C.POINTER(GcmPenaltySet)]),
    "gcm_add_table_update": (C.c_int, [_P, C.c_char_p, C.c_char_p, _I32, C.POINTER(GcmTableJob)]),
    "gcm_clear_table_updates": (C.c_int, [_P]),
    "gcm_get_table_values": (C.c_int, [_P, _I32, _P]),
    "gcm_enable_weights": (C.c_int, [_P, C.c_int]),
    "gcm_fill": (C.c_int, [_P, _P]),
    "gcm_sync": (C.c_int, [_P]),
    "gcm_enable_graph": (C.c_int, [_P, C.c_int]),
    "gcm_get_hist": (C.c_int, [_P, _I32, _P, _P]),
    "gcm_validate_hist": (C.c_int, [_P, _I32, _P, _P, _D, C.c_int, C.POINTER(_I32)]),
    "gcm_get_weights": (C.c_int, [_P, _P]),
    "gcm_get_llh": (C.c_int, [_P, _P, _P, _P]),
    "gcm_enable_timing": (C.c_int, [_P, C.c_int]),
    "gcm_get_timing": (C.c_int, [_P, C.POINTER(_D), C.POINTER(_D), C.POINTER(_I64)]),
    "gcm_bin_events": (C.c_int, [C.c_int, _I64, _I32, _P, _I32, _P, _P]),
    "gcm_mlp_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64, C.POINTER(_P)]),
    "gcm_mlp_destroy": (C.c_int, [_P]),
    "gcm_mlp_param_count": (C.c_int, [_P, C.POINTER(_I64)]),
    "gcm_mlp_get_params": (C.c_int, [_P, _P]),
    "gcm_mlp_set_params": (C.c_int, [_P, _P]),
    "gcm_mlp_predict": (C.c_int, [_P, _P, _I64, _P]),
    "gcm_mlp_gradients": (C.c_int, [_P, _P, _P, C.c_int, _P, C.POINTER(_D)]),
    "gcm_mlp_set_data": (C.c_int, [_P, _P, _P, _I64]),
    "gcm_mlp_train": (C.c_int, [_P, C.c_int, C.c_int, C.c_float, C.c_uint64, _P]),
    "gcm_mlp_set_head": (C.c_int, [_P, _P, _P, _P, _P]),
    "gcm_mlp_predict_llh": (C.c_int, [_P, _P, _I64, _P]),
    "gcm_mlp_predict_llh_noalloc": (C.c_int, [_P, _P, _I64, _P, _P]),
}

SYNTH_SIGNATURES = {
    "gcm_synth_last_error": (C.c_char_p, []),
    "gcm_synth_create": (_P, [C.c_int, _I64, C.c_uint64, C.c_int]),
    "gcm_synth_destroy": (None, [_P]),
    "gcm_synth_view_get": (C.c_int, [_P, C.POINTER(GcmSynthView)]),
    "gcm_synth_table_values": (C.c_int, [_P, _P, _P]),
}

_lib = None
_synth = None


class GcmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"gcm error {code}: {msg}")
        self.code = code


def _bind(lib, sigs):
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def lib() -> C.CDLL:
    """The HIP library.  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build()"
                " (the device path has no CPU fallback)")
        l = C.CDLL(LIB_PATH)
        _bind(l, GCM_SIGNATURES)
        _lib = l
    return _lib


def synth() -> C.CDLL:
    global _synth
    if _synth is None:
        if not os.path.exists(SYNTH_PATH):
            raise RuntimeError(f"{SYNTH_PATH} is missing: run the build first")
        l = C.CDLL(SYNTH_PATH)
        _bind(l, SYNTH_SIGNATURES)
        _synth = l
    return _synth


def check(rc: int) -> None:
    if rc != GCM_OK:
        raise GcmError(rc, lib().gcm_last_error().decode())


def device_count() -> int:
    n = C.c_int(0)
    check(lib().gcm_device_count(C.byref(n)))
    return n.value
