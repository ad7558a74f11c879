"""Build the gfx950 propagator library in-tree: _lib/libtdse.so.

Every translation unit is compiled by hipcc for gfx950 only, in parallel, then
linked into one shared library that exports the C ABI of include/tdse.h.
Packed FP32 instructions are disabled for all device code (they are formed
from ordinary float code otherwise).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_lib")
OBJ_DIR = os.path.join(OUT_DIR, "obj")
LIB = os.path.join(OUT_DIR, "libtdse.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = ["-std=c++17", "-O3", "-fPIC", "-Wall", "-Wno-unused-parameter"]
DEVICE = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

# (source, extra flags).  The builders restate the reference expressions and
# must not contract a*b+c into fma.
UNITS = [
    ("kernels_f32.hip", DEVICE),
    ("kernels_f64.hip", DEVICE),
    ("kernels_f32_batch.hip", DEVICE),
    ("kernels_f64_batch.hip", DEVICE),
    ("kernels_f32_v1.hip", DEVICE),
    ("kernels_f32_v2.hip", DEVICE),
    ("kernels_f32_v3.hip", DEVICE),
    ("kernels_f64_v1.hip", DEVICE),
    ("kernels_f64_v2.hip", DEVICE),
    ("kernels_f64_v3.hip", DEVICE),
    ("observe.hip", DEVICE),
    ("device_build.hip", DEVICE),
    ("lattice.hip", DEVICE),
    ("noalloc.hip", DEVICE),
    ("tdse_api.cpp", ["-x", "hip"] + DEVICE),
    ("builders.cpp", ["-ffp-contract=off"]),
]


def _compile(src: str, extra: list[str]) -> str:
    obj = os.path.join(OBJ_DIR, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *COMMON, *extra, "-c", os.path.join(CSRC, src), "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ_DIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(len(UNITS), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(lambda u: _compile(*u), UNITS))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", *objs, "-o", tmp]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
