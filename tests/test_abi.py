"""The C ABI library loads and exports every entry point declared in
include/gcm.h; host-side error behaviour; the workload generator is
deterministic.  No device compute here."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from gundam_amd import INPUT_DTYPE, Tables, Workload
from gundam_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    inc = os.path.join(ROOT, "include")
    text = "\n".join(open(os.path.join(inc, f)).read() for f in sorted(os.listdir(inc))
                     if f.endswith(".h"))
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gcm_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported_and_bound():
    lib = N.lib()
    names = _declared()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), f"{name} not exported"
        assert name in N.GCM_SIGNATURES, f"{name} has no ctypes binding"
    assert set(N.GCM_SIGNATURES) == set(names)


def test_version_and_error_paths():
    lib = N.lib()
    assert lib.gcm_version() >= 10000
    assert lib.gcm_set_precision(None, 0) == N.GCM_ERR_ARG
    assert b"NULL" in lib.gcm_last_error()
    assert lib.gcm_fill(None, None) == N.GCM_ERR_ARG
    assert lib.gcm_bin_events(0, 5, 0, None, 1, None, None) == N.GCM_ERR_ARG


def test_create_without_device_fails_cleanly():
    if N.device_count() > 0:
        pytest.skip("a device is visible")
    h = C.c_void_p()
    assert N.lib().gcm_create(0, C.byref(h)) == N.GCM_ERR_HIP
    assert not h.value


def test_workload_is_deterministic_across_thread_counts():
    a = Workload(2, n_events=9000, seed=5, threads=1)
    b = Workload(2, n_events=9000, seed=5, threads=7)
    for name in ("base_weight", "dial_begin", "dial_type", "dial_input",
                 "data_begin", "dial_data"):
        np.testing.assert_array_equal(getattr(a.tables, name), getattr(b.tables, name))
    np.testing.assert_array_equal(a.vars, b.vars)
    c = Workload(2, n_events=9000, seed=6)
    assert not np.array_equal(a.vars, c.vars)


def test_workload_shapes():
    wl = Workload(2, n_events=20000, seed=1)
    t = wl.tables
    assert t.n_bins == 4 * 1500 and t.n_parameters == 300
    assert np.all(np.diff(t.dial_begin) == 12)
    assert t.n_tables == 90 and t.table_len == 45 * 64 + 45 * 256
    w1 = Workload(1, seed=1)
    assert w1.tables.n_events == 100000 and w1.tables.n_bins == 400
    assert np.all(np.diff(w1.tables.dial_begin) == 5)


def test_algorithmic_bytes_hand_count():
    inputs = np.zeros(1, dtype=INPUT_DTYPE)
    t = Tables(base_weight=np.ones(2), bin=np.array([0, -1], np.int32),
               dial_begin=np.array([0, 2, 3], np.int64),
               dial_type=np.array([N.DIAL_MONOTONIC, N.DIAL_NORM, N.DIAL_KRIGED], np.uint8),
               dial_input=np.zeros(3, np.int32),
               data_begin=np.array([0, 9, 9, 14], np.int64), dial_data=np.zeros(14),
               inputs=inputs, table_begin=np.array([0, 16], np.int64),
               n_parameters=3, n_bins=4)
    # monotonic 9 words*8 + 2, norm 2, kriged (k=2) 8*2+4+2, one binned event
    want = (72 + 2) + 2 + (20 + 2) + 16 + 8 * 16 + 8 * 3 + 16 * 4
    assert t.algorithmic_bytes(8) == want


def test_cpp_facade_compiles_and_links():
    """gcm_cache_manager.hpp compiles as C++17 and links against libgcm."""
    import subprocess
    import tempfile
    csrc = os.path.join(ROOT, "gundam_amd", "csrc")
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, "demo")
        r = subprocess.run(["g++", "-std=c++17", "-O0", "-o", out,
                            os.path.join(csrc, "facade_demo.cpp"),
                            "-L" + os.path.join(ROOT, "gundam_amd"), "-lgcm", "-lgcm_synth",
                            "-lpthread"], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr


def test_struct_layouts_match_the_header(tmp_path):
    """The ctypes mirrors in _native.py have the sizes and field offsets the
    C compiler gives the structs of include/gcm.h."""
    structs = {"gcm_input": N.GcmInput, "gcm_tables": N.GcmTables,
               "gcm_llh_config": N.GcmLlhConfig, "gcm_launch_config": N.GcmLaunchConfig,
               "gcm_penalty_set": N.GcmPenaltySet, "gcm_table_job": N.GcmTableJob}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "gcm.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'  printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    subprocess.run(["gcc", "-std=c11", "-I", inc, "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    for line in out.splitlines():
        cname, field, val = line.split()
        py = structs[cname]
        want = C.sizeof(py) if field == "size" else getattr(py, field).offset
        assert int(val) == want, (cname, field, int(val), want)
