"""Compile the native libraries in-tree.

    python -m rfba_amd.build          # everything
    python -m rfba_amd.build srrf     # one target

Targets (all outputs next to their sources, git-ignored):
  rfba_amd/librfba_srrf.so   HIP kernels + C ABI of include/rfba_srrf.h (gfx950)
  rfba_amd/librfba_3b.so     HIP kernels + C ABI of include/rfba_3b.h (gfx950)
  rfba_amd/librfba_prep.so   TIFF loading + normalisation, C ABI of include/rfba_prep.h (gfx950)
  rfba_amd/librfba_synth.so  host-only synthetic data generator
  oracle/liboracle.so        the CPU oracle (test infrastructure)
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "rfba_amd")
CSRC = os.path.join(PKG, "csrc")
ORACLE = os.path.join(ROOT, "oracle")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# Device code: gfx950 only; packed FP32 disabled; no FP contraction for the
# SRRF kernels (they reproduce the reference's separately rounded products
# and sums).
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared", "--offload-arch=gfx950",
    "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
GXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-shared", "-pthread", "-ffp-contract=off"]


def _run(cmd):
    r = subprocess.run(cmd, cwd=ROOT, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def _hip(out, srcs, contract="off"):
    flags = [f if f != "-ffp-contract=off" else "-ffp-contract=" + contract for f in HIP_FLAGS]
    return [HIPCC, *flags, "-x", "hip", *srcs, "-o", out]


TARGETS = {
    "srrf": lambda: _hip(os.path.join(PKG, "librfba_srrf.so"),
                         [os.path.join(CSRC, "srrf_kernel.hip"), os.path.join(CSRC, "srrf_api.cpp"),
                          os.path.join(CSRC, "srrf_cxx_linkage.cpp")]),
    # 3B: FP32 pixel terms checked against an FP64 oracle with a tolerance;
    # contraction to FMA is allowed there.
    "b3": lambda: _hip(os.path.join(PKG, "librfba_3b.so"),
                       [os.path.join(CSRC, "b3_kernel.hip"), os.path.join(CSRC, "b3_gibbs.hip"),
                        os.path.join(CSRC, "b3_api.cpp")], contract="fast"),
    # input loading and normalisation: FP32 order of libcvd's FIR, no contraction
    "prep": lambda: _hip(os.path.join(PKG, "librfba_prep.so"), [os.path.join(CSRC, "prep.hip")]),
    "synth": lambda: ["g++", *GXX_FLAGS, "-O3", os.path.join(CSRC, "synth.cpp"),
                      "-o", os.path.join(PKG, "librfba_synth.so")],
    "oracle": lambda: ["g++", *GXX_FLAGS,
                       *[os.path.join(ORACLE, f) for f in ("srrf_oracle.cpp", "b3_oracle.cpp", "prep_oracle.cpp")
                         if os.path.exists(os.path.join(ORACLE, f))],
                       "-o", os.path.join(ORACLE, "liboracle.so")],
}


def _available(name):
    cmd = TARGETS[name]()
    srcs = [a for a in cmd if a.endswith((".hip", ".cpp"))]
    return all(os.path.exists(s) for s in srcs)


def build(targets=None, verbose=True):
    names = [n for n in (targets or list(TARGETS)) if _available(n)]
    if any(n in ("srrf", "b3", "prep") for n in names) and not shutil.which(HIPCC) and not os.path.exists(HIPCC):
        raise RuntimeError("hipcc not found at %s" % HIPCC)
    with ThreadPoolExecutor(max_workers=4) as ex:
        futs = {n: ex.submit(_run, TARGETS[n]()) for n in names}
        for n, f in futs.items():
            f.result()
            if verbose:
                print("built", n, flush=True)
    return names


if __name__ == "__main__":
    build(sys.argv[1:] or None)
