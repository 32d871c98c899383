"""Builds libsm2d_hip.so in-tree (hipcc for gfx950 + g++ for the host table builders).

Run as `python -m sm_2d_amd.build`.  The library lands next to this file so it travels with
the source tree; nothing is installed.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsm2d_hip.so")
RUNNER = os.path.join(HERE, "sm2d_run")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SM2D_ARCH", "gfx950")

# -ffp-contract=off on both compilers: the shared elementary functions and the per-bin physics
# must round identically on the CPU oracle and on the GPU.  Packed FP32 instructions are
# disabled for device code.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result"]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + r.stdout + r.stderr)


# Diagnostic build with the transport kernel's stage clocks (sm2d_stage_clocks): same sources and
# work, into its own directory (load it with SM2D_LIB=<path>).
CLOCKS_LIB = os.path.join(HERE, "_clocks", "libsm2d_hip.so")


def build(verbose: bool = False, stage_clocks: bool = False) -> str:
    out_dir = os.path.dirname(CLOCKS_LIB) if stage_clocks else HERE
    bdir = os.path.join(BUILD, "clocks") if stage_clocks else BUILD
    lib = CLOCKS_LIB if stage_clocks else LIB
    os.makedirs(bdir, exist_ok=True)
    os.makedirs(out_dir, exist_ok=True)
    host_o = os.path.join(bdir, "sm2d_host.o")
    io_o = os.path.join(bdir, "sm2d_io.o")
    dev_o = os.path.join(bdir, "sm2d_runtime.o")
    extra = ["-DSM2D_STAGE_CLOCKS"] if stage_clocks else []
    steps = [
        ["g++", *HOST_FLAGS, "-c", os.path.join(CSRC, "sm2d_host.cpp"), "-o", host_o],
        ["g++", *HOST_FLAGS, "-c", os.path.join(CSRC, "sm2d_io.cpp"), "-o", io_o],
        [HIPCC, *HIP_FLAGS, *extra, "-c", os.path.join(CSRC, "sm2d_runtime.hip"), "-o", dev_o],
        [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", host_o, io_o, dev_o, "-o", lib + ".tmp"],
    ]
    for cmd in steps:
        if verbose:
            print(" ".join(cmd), flush=True)
        _run(cmd)
    os.replace(lib + ".tmp", lib)
    if stage_clocks:
        return lib
    # the INI driver (run_simulation's counterpart), linked against the library beside it
    _run(["g++", *HOST_FLAGS, os.path.join(CSRC, "sm2d_run.cpp"), "-o", RUNNER + ".tmp",
          "-L" + HERE, "-lsm2d_hip", "-Wl,-rpath,$ORIGIN"])
    os.replace(RUNNER + ".tmp", RUNNER)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv, stage_clocks="--stage-clocks" in sys.argv))
