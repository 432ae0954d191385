"""Build libbcsim_hip.so (the C-ABI library) for gfx950 with hipcc, in-tree.

    python -m openbcsim_amd.build

Compile flags:
  -ffp-contract=off   the radial coordinate and spline position must round like
                      the reference's float code (they select a bin); the FFT
                      code re-enables contraction locally with a pragma
  -munsafe-fp-atomics hardware float atomics (buffers are coarse-grained hipMalloc)
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      no packed FP32 instructions
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libbcsim_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", f"--offload-arch={ARCH}",
          "-munsafe-fp-atomics", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-I", os.path.join(ROOT, "include"), "-fvisibility=hidden"]

SOURCES = ["sim_kernels.hip", "project_tile.hip", "sim_host.cpp", "batch_host.cpp", "serve_host.cpp"]


def _run(cmd):
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in res.stdout.splitlines() if "packed-fp32-ops" not in l)
    if res.returncode != 0:
        raise RuntimeError(f"command failed ({res.returncode}): {' '.join(cmd)}\n{out}")
    return out


def build(verbose: bool = False) -> str:
    """Compile every HIP source and link the shared library; returns its path."""
    obj_dir = os.path.join(PKG, "_build")
    os.makedirs(obj_dir, exist_ok=True)
    objs = [os.path.join(obj_dir, os.path.splitext(src)[0] + ".o") for src in SOURCES]
    cmds = [[HIPCC, *COMMON, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj] for src, obj in zip(SOURCES, objs)]
    # one compiler process per source (each is single-threaded)
    with ThreadPoolExecutor(max_workers=min(len(cmds), max(1, os.cpu_count() or 1))) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-pthread", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


CEILINGS_LIB = os.path.join(PKG, "libbcsim_ceilings.so")


def build_ceilings() -> str:
    """Compile the benchmark's ceiling kernels (csrc/ceilings.hip) into
    libbcsim_ceilings.so: measurement infrastructure, not the simulation path."""
    tmp = CEILINGS_LIB + ".tmp"
    _run([HIPCC, *COMMON, "-shared", "-x", "hip", os.path.join(CSRC, "ceilings.hip"), "-o", tmp])
    os.replace(tmp, CEILINGS_LIB)
    return CEILINGS_LIB


def build_oracle() -> None:
    """Build the test-only checkers under oracle/ (the FP64 oracle, and the compiled
    reference CPU path when the reference tree is present)."""
    oracle = os.path.join(ROOT, "oracle")
    _run(["make", "-s", "-C", oracle, "-B", "oracle"])
    if os.path.isdir("/root/reference/src/core"):
        _run(["make", "-s", "-C", oracle, "-B", "ref"])


C_CALLER_FLAGS = ["-std=c99", "-Wall", "-Wextra", "-Werror", "-pedantic"]


def build_cxx_check() -> str:
    """Compile tests/cxx/facade_check.cpp (a C++ IAlgorithm user of the facade
    header) and tests/cxx/c_caller.c (a C99 user of the C-ABI, warnings as errors)
    against the built library, for the GPU tests to run."""
    out = os.path.join(PKG, "_build", "facade_check")
    _run(["g++", "-std=c++14", "-O2", "-I", os.path.join(ROOT, "include"),
          os.path.join(ROOT, "tests", "cxx", "facade_check.cpp"), "-o", out, "-L", PKG, "-lbcsim_hip",
          "-Wl,-rpath,$ORIGIN/.."])
    _run(["gcc", *C_CALLER_FLAGS, "-O2", "-I", os.path.join(ROOT, "include"),
          os.path.join(ROOT, "tests", "cxx", "c_caller.c"), "-o", os.path.join(PKG, "_build", "c_caller"),
          "-L", PKG, "-lbcsim_hip", "-lm", "-Wl,-rpath,$ORIGIN/.."])
    return out


if __name__ == "__main__":
    path = build(verbose="-v" in sys.argv)
    build_cxx_check()
    build_ceilings()
    print(path)
