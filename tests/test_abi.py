"""The C ABI surface: every function declared in include/*.h is exported by libsm2d_hip.so, and
the ctypes mirrors (sm_2d_amd/abi.py, sm_2d_amd/io.py) have the C layout of each struct.
No GPU work is done here."""
import glob
import ctypes as C
import os
import re
import subprocess

import pytest

from sm_2d_amd import abi, io

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = sorted(glob.glob(os.path.join(ROOT, "include", "*.h")))


def declared_functions():
    text = "\n".join(open(h).read() for h in HEADERS)
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(sm2d_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("sm2d_ctx_create", "sm2d_lut_upload", "sm2d_run_spot", "sm2d_k1k2_mask_compact",
                 "sm2d_run_iterations", "sm2d_k5_audit", "sm2d_swap", "sm2d_inject"):
        assert must in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(abi.LIB_PATH):
        pytest.fail(f"{abi.LIB_PATH} missing: run `python -m sm_2d_amd.build`")
    lib = C.CDLL(abi.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


STRUCTS = {"sm2d_grid": abi.Grid, "sm2d_transport": abi.Transport, "sm2d_spot": abi.Spot,
           "sm2d_audit": abi.Audit, "sm2d_k5_report": abi.K5Report, "sm2d_timing": abi.Timing,
           "sm2d_run_config": io.RunConfig, "sm2d_memory_plan": io.MemoryPlan}


def test_ctypes_structs_match_c_layout(tmp_path):
    src = tmp_path / "sz.c"
    lines = ['#include <stdio.h>', '#include <stddef.h>'] + [f'#include "{h}"' for h in HEADERS]
    lines.append("int main(void){")
    for name, cls in STRUCTS.items():
        lines.append(f'printf("{name} %zu\\n", sizeof({name}));')
        for f, _ in cls._fields_:
            lines.append(f'printf("{name}.{f} %zu\\n", offsetof({name}, {f}));')
    lines.append("return 0;}")
    src.write_text("\n".join(lines))
    exe = tmp_path / "sz"
    subprocess.run(["gcc", str(src), "-o", str(exe)], check=True)
    out = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                 check=True).stdout.splitlines())
    for name, cls in STRUCTS.items():
        assert int(out[name]) == C.sizeof(cls), name
        for f, _ in cls._fields_:
            assert int(out[f"{name}.{f}"]) == getattr(cls, f).offset, (name, f)
