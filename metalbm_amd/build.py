"""Builds libmetalbm_amd.so in-tree with hipcc for gfx950.

    python -m metalbm_amd.build            # compile every source, then link
    python -m metalbm_amd.build --changed  # recompile only out-of-date objects

FP64-compute kernels, the initialisation kernels and the C ABI are built with
``-ffp-contract=off`` (reference operation order for FP64 parity); the FP32
kernels with FMA contraction.  Every HIP file also gets
``-Xclang -target-feature -Xclang -packed-fp32-ops`` so the compiler never
emits packed FP32 arithmetic.
"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmetalbm_amd.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

ARCH = "gfx950"
COMMON = ["-O3", f"--offload-arch={ARCH}", "-std=c++17", "-fPIC",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wno-unused-command-line-argument", f"-I{INCLUDE}"]
SOURCES = {
    "step_f32_d3q27.hip": ["-ffp-contract=fast"],
    "step_f32_d3q19.hip": ["-ffp-contract=fast"],
    "step_f32_d2q9.hip": ["-ffp-contract=fast"],
    "step_f64_d3q27.hip": ["-ffp-contract=off"],
    "step_f64_d3q19.hip": ["-ffp-contract=off"],
    "step_f64_d2q9.hip": ["-ffp-contract=off"],
    "aux_kernels.hip": ["-ffp-contract=off"],
    "mlp_train.hip": ["-ffp-contract=fast"],
    "capi.hip": ["-ffp-contract=off"],
    "hdf5_io.cpp": ["-ffp-contract=off"],
    "spectral.hip": ["-ffp-contract=off"],
    "probe.hip": ["-ffp-contract=off"],
}


def _hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.sep not in c or os.path.exists(c)):
            return c
    raise RuntimeError("hipcc not found")


def _headers():
    return [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hpp")] + \
        [os.path.join(INCLUDE, f) for f in os.listdir(INCLUDE) if f.endswith(".h")]


def _stale(obj: str, src: str) -> bool:
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(p) > t for p in [src] + _headers())


def _compile(src: str, flags, changed_only: bool) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    path = os.path.join(CSRC, src)
    if changed_only and not _stale(obj, path):
        return obj
    cmd = [_hipcc()] + COMMON + flags + ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(changed_only: bool = False, verbose: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    workers = max(1, min(len(SOURCES), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))))
    with ThreadPoolExecutor(max_workers=workers) as ex:
        objs = list(ex.map(lambda kv: _compile(kv[0], kv[1], changed_only),
                           SOURCES.items()))
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-L/opt/rocm/lib", "-lhipfft", "-ldl"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--changed", action="store_true",
                    help="recompile only objects older than their sources")
    args = ap.parse_args(argv)
    build(changed_only=args.changed)
    return 0


if __name__ == "__main__":
    sys.exit(main())
