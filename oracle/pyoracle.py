"""Python handle on the CPU oracle (oracle/_build/liboracle_sm2d.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "_build", "liboracle_sm2d.so")
LIB_OMP = os.path.join(HERE, "_build", "liboracle_sm2d_omp.so")
# the same transport step evaluated in binary64 (the FP32 results' error is stated against it)
LIB_FP64 = os.path.join(HERE, "_build", "liboracle_sm2d_fp64.so")
# -O3 -march=x86-64-v3 build of the same source (the CPU baseline; results equal the -O2 build's)
LIB_FAST = os.path.join(HERE, "_build", "liboracle_sm2d_fast.so")

if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from sm_2d_amd import abi  # noqa: E402
from sm_2d_amd.engine import Engine, Library  # noqa: E402


def build(quiet: bool = True) -> None:
    """Compiles the oracle (and oracle/_ref when the reference checkout is present)."""
    out = subprocess.run(["make", "-C", HERE, "-j4"], capture_output=quiet, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + (out.stdout or "") + (out.stderr or ""))


def load(omp: bool = False, fp64: bool = False, fast: bool = False) -> Library:
    path = LIB_FP64 if fp64 else LIB_OMP if omp else LIB_FAST if fast else LIB
    if not os.path.exists(path):
        build()
    return Library(path, "orc_", {
        "eval": [C.c_void_p, C.c_int, abi.fp, abi.fp, abi.fp, C.c_int],
        "set_threads": [C.c_void_p, C.c_int],
        "surrogate_set": [C.c_void_p, abi.fp, C.c_int, C.c_int, abi.i32p, C.c_int],
    })


class OracleEngine(Engine):
    def eval(self, which: int, x, y=None):
        import numpy as np
        x = np.ascontiguousarray(x, np.float32)
        yy = None if y is None else np.ascontiguousarray(y, np.float32)
        out = np.zeros_like(x)
        self.L.call("eval", self.h, which, x.ctypes.data_as(abi.fp),
                    None if yy is None else yy.ctypes.data_as(abi.fp),
                    out.ctypes.data_as(abi.fp), len(x))
        return out

    def surrogate_set(self, table, n_mat: int, n_nodes: int, rowmat, on: bool = True):
        """Uses the library's learned-physics table (the layout of sm2d_surrogate_export)."""
        import numpy as np
        t = np.ascontiguousarray(table, np.float32)
        r = np.ascontiguousarray(rowmat, np.int32)
        self._surr_keep = (t, r)
        self.L.call("surrogate_set", self.h, t.ctypes.data_as(abi.fp), n_mat, n_nodes,
                    r.ctypes.data_as(abi.i32p), 1 if on else 0)
