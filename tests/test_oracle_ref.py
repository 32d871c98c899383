"""Host table builders against the reference's own code (oracle/_ref/libsm2d_ref.so, compiled from
the reference's grids.cpp, r_lut.cpp and nist_loader.cpp by oracle/Makefile), and against the
NIST nodes.  Both the oracle's restatement and the product library's builders must be
bit-identical to it."""
import ctypes as C
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref", "libsm2d_ref.so")

pytestmark = pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref not built (no reference checkout)")

fp = C.POINTER(C.c_float)


@pytest.fixture(scope="module")
def ref():
    lib = C.CDLL(REF)
    for n in ("ref_energy_grid_log", "ref_angular_grid"):
        getattr(lib, n).restype = C.c_int
    return lib


def _f(a):
    return a.ctypes.data_as(fp)


GROUPS = [
    ((0.1, 2.0, 0.1), (2.0, 20.0, 0.2), (20.0, 100.0, 0.25), (100.0, 250.0, 0.25)),
    ((0.5, 3.0, 0.05), (3.0, 230.0, 0.5)),
    ((0.1, 250.0, 0.7),),
]


def _ref_piecewise(ref, groups):
    g = np.asarray(groups, np.float32).reshape(-1)
    n = C.c_int(0)
    ref.ref_energy_grid_piecewise(_f(g), len(groups), None, None, C.byref(n))
    e, r = np.zeros(n.value + 1, np.float32), np.zeros(n.value, np.float32)
    ref.ref_energy_grid_piecewise(_f(g), len(groups), _f(e), _f(r), C.byref(n))
    return e, r


def _ref_rlut(ref, groups=None, log=None):
    if groups is not None:
        g = np.asarray(groups, np.float32).reshape(-1)
        ng = len(groups)
        n = len(_ref_piecewise(ref, groups)[1])
    else:
        g = np.asarray(log, np.float32)
        ng = 0
        n = int(log[2])
    out = [np.zeros(n, np.float32) for _ in range(5)]
    ref.ref_build_rlut(_f(g), ng, *[_f(o) for o in out])
    return out


def _libs(orc):
    libs = [orc]
    from sm_2d_amd import abi
    if os.path.exists(abi.LIB_PATH):
        from sm_2d_amd.engine import Library
        libs.append(Library(abi.LIB_PATH, "sm2d_"))
    return libs


@pytest.mark.parametrize("groups", GROUPS)
def test_piecewise_grid_and_rlut_bit_exact(ref, orc, groups):
    e_ref, r_ref = _ref_piecewise(ref, groups)
    lut_ref = _ref_rlut(ref, groups=groups)
    for lib in _libs(orc):
        e, r = lib.energy_grid_piecewise(groups)
        assert np.array_equal(e, e_ref) and np.array_equal(r, r_ref), lib.prefix
        lut = lib.build_rlut(r)
        for a, b in zip(lut, lut_ref):
            assert np.array_equal(a, b), lib.prefix


@pytest.mark.parametrize("args", [(0.1, 250.0, 256), (0.1, 300.0, 256), (1.0, 230.0, 97)])
def test_log_grid_and_rlut_bit_exact(ref, orc, args):
    e_ref, r_ref = np.zeros(args[2] + 1, np.float32), np.zeros(args[2], np.float32)
    ref.ref_energy_grid_log(C.c_float(args[0]), C.c_float(args[1]), args[2], _f(e_ref), _f(r_ref))
    lut_ref = _ref_rlut(ref, log=args)
    for lib in _libs(orc):
        e, r = lib.energy_grid_log(*args)
        assert np.array_equal(e, e_ref) and np.array_equal(r, r_ref)
        for a, b in zip(lib.build_rlut(r), lut_ref):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("args", [(-0.35, 0.35, 36), (-0.35, 0.35, 360), (-np.pi / 2, np.pi / 2, 512)])
def test_angular_grid_bit_exact(ref, orc, args):
    e_ref = np.zeros(args[2] + 1, np.float32)
    ref.ref_angular_grid(C.c_float(args[0]), C.c_float(args[1]), args[2], _f(e_ref))
    for lib in _libs(orc):
        assert np.array_equal(lib.angular_grid(*args), e_ref)


def test_fe_precompute_uses_reference_host_lookups(ref, orc):
    """The FE precompute is restated (k1k6_pipeline.cu:1061-1176); its ingredients — the host
    log-LUT energy loss and the host Highland formula — are compared with the reference's own
    functions, and the recurrences are checked for the defining identities."""
    x = np.linspace(0.6, 230.0, 400).astype(np.float32)
    y = np.full_like(x, 0.5)
    dep = np.zeros_like(x)
    ref.ref_log_lut_eval(3, _f(x), _f(y), _f(dep), len(x))
    hl = np.zeros_like(x)
    ref.ref_log_lut_eval(5, _f(x), _f(y), _f(hl), len(x))
    for lib in _libs(orc):
        A, B, st = lib.fe_precompute(150.0, 3.0, 400, 0.5)
        assert A[0] == 0.0 and B[0] == 0.0 and st[0] == np.float32(3.0)
        dA = np.diff(A.astype(np.float64))
        live = dA > 0
        assert live.sum() > 250          # 150 MeV: ~158 mm of range over 0.5 mm rows
        assert np.all(np.diff(st[: live.sum()]) >= 0)  # lateral sigma grows with depth
    # A increments are T*dz = highland(E_mid, dz)^2 with E_mid from the reference's LUT:
    # recompute the first rows with the reference's own functions
    E = np.float32(150.0)
    A_ref = [np.float32(0.0)]
    for _ in range(5):
        xe = np.array([E], np.float32)
        d = np.zeros(1, np.float32)
        ref.ref_log_lut_eval(3, _f(xe), _f(np.array([0.5], np.float32)), _f(d), 1)
        dE = min(d[0], np.float32(E - np.float32(0.1)))
        Em = np.float32(E - np.float32(0.5) * dE)
        h = np.zeros(1, np.float32)
        ref.ref_log_lut_eval(5, _f(np.array([Em], np.float32)), _f(np.array([0.5], np.float32)), _f(h), 1)
        T = np.float32(h[0] * h[0] / np.float32(0.5))
        A_ref.append(np.float32(A_ref[-1] + T * np.float32(0.5)))
        E = np.float32(E - dE)
    A, _, _ = orc.fe_precompute(150.0, 3.0, 400, 0.5)
    assert np.array_equal(A[:6], np.array(A_ref, np.float32))


def test_encodings_match_reference(ref):
    def LAYOUT_REF_ENC(t, e, x, z):  # Layout<4,2,8,4>::enc of sm2d_device.h
        return t + 4 * (e + 2 * (x + 8 * z))
    ref.ref_local_bins.restype = C.c_int
    assert ref.ref_local_bins() == 256
    for l in range(256):
        t, e, x, z = (C.c_int() for _ in range(4))
        ref.ref_decode_local_idx_4d(l, C.byref(t), C.byref(e), C.byref(x), C.byref(z))
        assert LAYOUT_REF_ENC(t.value, e.value, x.value, z.value) == l
        assert ref.ref_encode_local_idx_4d(t.value, e.value, x.value, z.value) == l
    ref.ref_x_offset.restype = C.c_float
    ref.ref_z_offset.restype = C.c_float
    for xs in range(8):
        assert ref.ref_x_offset(xs, C.c_float(0.5)) == np.float32(0.5) * np.float32((2 * xs + 1 - 8) / 16)
    for zs in range(4):
        assert ref.ref_z_offset(zs, C.c_float(0.5)) == np.float32(0.5) * np.float32((2 * zs + 1 - 4) / 8)


def test_two_percent_step_rule_matches_reference(ref, orc):
    """The 2 % range step rule (step_rule = 1) against the reference's host
    compute_max_step_physics (step_control.hpp:18-66) on the same R-LUT (the log grid 0.1-300 MeV
    of 256 bins that ref_log_lut_eval holds).  Wherever the oracle's range lookup (the device rule
    of device_lut.cuh:66-85) returns the reference host lookup's value, the step is equal bit for
    bit, across every refinement branch (E < 5, 10, 20, 50 MeV and above, the 0.05 mm floor)."""
    from oracle.pyoracle import OracleEngine
    from sm_2d_amd import problems
    prob = problems.reference_grid_case(Nx=10, Nz=10)
    prob.energy = ("log", 0.1, 300.0, 256)
    eng = problems.setup_engine(orc, prob, OracleEngine)
    E = np.concatenate([np.geomspace(0.15, 250.0, 4000),
                        [4.99, 5.0, 9.99, 10.0, 19.99, 20.0, 49.99, 50.0]]).astype(np.float32)
    R_o, S_o = eng.eval(0, E), eng.eval(4, E)
    R_r, S_r = np.zeros_like(E), np.zeros_like(E)
    ref.ref_log_lut_eval(0, _f(E), None, _f(R_r), len(E))
    ref.ref_log_lut_eval(4, _f(E), None, _f(S_r), len(E))
    same = R_o == R_r
    assert same.mean() > 0.8
    assert np.array_equal(S_o[same], S_r[same])
    assert np.allclose(S_o, S_r, rtol=2e-6)
    assert S_o.min() == np.float32(0.05)  # the floor is reached near the end of range
