"""Build libqsm_gpu.so (HIP, gfx950) in-tree with hipcc.

Usage: python qsm.rs_amd/build.py [--jobs N]
The library lands next to this file; nothing is installed elsewhere.
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
LIB = os.path.join(HERE, "libqsm_gpu.so")

# qsm_pass.hip is compiled once per specialised axis length, and with length
# 0 in four parts (QSM_GENERIC_PART) for the runtime-plan kernels; keep in
# step with QSM_FIXED_LENGTHS in qsm_misc.hip.
FIXED_LENGTHS = [48, 64, 72, 80, 96, 112, 120, 128, 144, 150, 160, 168, 176, 180, 192, 200, 208, 216, 220,
                 224, 240, 256, 288, 320, 384, 416, 448, 480, 512]
UNITS = ([("qsm_pass.hip", f"qsm_pass_0_{k}", ["-DQSM_PASS_N=0", f"-DQSM_GENERIC_PART={k}"]) for k in range(4)]
         + [("qsm_pass.hip", f"qsm_pass_{n}", [f"-DQSM_PASS_N={n}"]) for n in FIXED_LENGTHS]
         + [("qsm_misc.hip", "qsm_misc", []), ("qsm_runtime.hip", "qsm_runtime", []),
            ("qsm_nifti.hip", "qsm_nifti", [])])

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC",
    # no packed-FP32 instructions anywhere (v_pk_*_f32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.isabs(cand) and os.path.exists(cand) or not os.path.isabs(cand)):
            return cand
    return "hipcc"


def _compile(unit) -> str:
    src, name, defs = unit
    obj = os.path.join(BUILD, name + ".o")
    cmd = [hipcc(), *HIPCC_FLAGS, *defs, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed on {src}")
    return obj


def build(jobs: int = 8, verbose: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(_compile, UNITS))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, "-lz"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link of libqsm_gpu.so failed")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    a = ap.parse_args()
    build(a.jobs)
