"""Build libcvrpaco.so (HIP kernels + C ABI, gfx950) and the cvrp_aco CLI in-tree.

Invoked by ``__graft_entry__.build()``; also runnable as
``python cvrp_aco_ocl_amd/build.py`` (not through the package, whose import
loads the library this script builds).  Plain hipcc invocations, no CMake.
"""
from __future__ import annotations

import os
import shlex
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(PKG, "libcvrpaco.so")
PROBE = os.path.join(PKG, "libcvrpprobe.so")   # roof microbenchmarks: LDS, HBM copy, BF16 MFMA
CLI = os.path.join(PKG, "cvrp_aco")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CVRP_OFFLOAD_ARCH", "gfx950")

# Float results must equal the CPU restatement's: no FMA contraction, IEEE
# division/sqrt, denormals kept.  Packed FP32 instructions are disabled.
DEVICE_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-gpu-flush-denormals-to-zero",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
    f"-I{INCLUDE}", f"-I{CSRC}",
]

LIB_SOURCES = ["aco_kernels.hip", "aco_local_search.hip", "aco_prior.hip", "aco_nnlist.hip",
               "aco_setup.hip", "aco_capi.cpp", "aco_host.cpp", "aco_decompose.cpp"]
CLI_SOURCES = ["g_aco.cpp", "main.cpp"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(shlex.quote(c) for c in cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build failed: {cmd[0]} exited {r.returncode}")
    if verbose and r.stderr.strip():
        sys.stderr.write(r.stderr)


def build(verbose: bool = True) -> None:
    srcs = [os.path.join(CSRC, s) for s in LIB_SOURCES]
    tmp = LIB + ".tmp"
    _run([HIPCC, *DEVICE_FLAGS, "-shared", "-o", tmp, *srcs], verbose)
    os.replace(tmp, LIB)
    _run([HIPCC, *DEVICE_FLAGS, "-shared", "-o", PROBE + ".tmp",
          *[os.path.join(CSRC, s) for s in ("probe_lds.hip", "probe_roofs.hip")]], verbose)
    os.replace(PROBE + ".tmp", PROBE)
    cli_srcs = [os.path.join(CSRC, s) for s in CLI_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    if len(cli_srcs) == len(CLI_SOURCES):
        _run(["g++", "-O2", "-std=c++17", "-Wall", f"-I{INCLUDE}", f"-I{CSRC}", "-o", CLI + ".tmp",
              *cli_srcs, f"-L{PKG}", "-lcvrpaco", "-Wl,-rpath,$ORIGIN"], verbose)
        os.replace(CLI + ".tmp", CLI)


if __name__ == "__main__":
    build(verbose=True)
