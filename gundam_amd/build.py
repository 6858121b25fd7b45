"""Build the native libraries in-tree (no JIT cache): the HIP library for
gfx950 with hipcc and the host workload generator with g++."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# -ffp-contract=off: every FP64 operation rounds as the reference's x86-64
# build does (no fused multiply-add), which the parity tests rely on.
# packed-fp32-ops is disabled for all HIP code.
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
             "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-Wall", "-Wno-unused-function"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "is not a recognized feature" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("build failed: " + " ".join(cmd))
    return out


def build_gcm() -> str:
    out = os.path.join(HERE, "libgcm.so")
    tmp = out + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-o", tmp, os.path.join(CSRC, "gcm.hip"),
          os.path.join(CSRC, "gcm_points.hip"), os.path.join(CSRC, "mlp.hip")])
    os.replace(tmp, out)
    return out


def build_table_update() -> str:
    """The synthetic workloads' device table function as a code object for
    gcm_add_table_update."""
    out = os.path.join(HERE, "table_update_cos.hsaco")
    tmp = out + ".tmp"
    flags = [f for f in HIP_FLAGS if f not in ("-shared", "-fPIC")]
    _run([HIPCC, "--genco", *flags, "-o", tmp, os.path.join(CSRC, "table_update_cos.hip")])
    os.replace(tmp, out)
    return out


def build_synth() -> str:
    out = os.path.join(HERE, "libgcm_synth.so")
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-o", tmp,
          os.path.join(CSRC, "synth.cpp"), "-lpthread"])
    os.replace(tmp, out)
    return out


def build_hbm_copy() -> str:
    """The copy kernel of known size whose FETCH_SIZE / WRITE_SIZE readings
    calibrate the fill kernel's measured HBM traffic (bench.py's
    roofline.traffic)."""
    out = os.path.join(HERE, "hbm_copy")
    tmp = out + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang",
          "-target-feature", "-Xclang", "-packed-fp32-ops", "-o", tmp,
          os.path.join(CSRC, "hbm_copy.hip")])
    os.replace(tmp, out)
    return out


def build_facade_demo() -> str:
    """The C++ facade example, linked against the in-tree libraries."""
    out = os.path.join(HERE, "facade_demo")
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-o", tmp,
          os.path.join(CSRC, "facade_demo.cpp"), "-L" + HERE, "-lgcm",
          "-lgcm_synth", "-Wl,-rpath,$ORIGIN", "-lpthread"])
    os.replace(tmp, out)
    return out


def build_all() -> None:
    build_synth()
    build_gcm()
    build_table_update()
    build_hbm_copy()
    build_facade_demo()


if __name__ == "__main__":
    build_all()
