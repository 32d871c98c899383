"""ctypes mirror of include/sm2d_hip.h.

The POD structs and the function prototypes of the C ABI, shared by the host-side runner
(sm_2d_amd.runner) and by the tests.  Field order and types follow the header exactly.
"""
from __future__ import annotations

import ctypes as C
import os

KB = 32
EMPTY = 0xFFFFFFFF

OK = 0
ERR_ARG = -1
ERR_HIP = -2
ERR_CAPACITY = -3
ERR_AUDIT = -4
ERR_UNSUPPORTED = -5

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsm2d_hip.so")


class Grid(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint32),
        ("Nx", C.c_int), ("Nz", C.c_int),
        ("dx", C.c_float), ("dz", C.c_float),
        ("x_min", C.c_float), ("z_min", C.c_float),
        ("N_theta", C.c_int),
        ("theta_min", C.c_float), ("theta_max", C.c_float),
        ("N_E", C.c_int),
        ("n_theta_local", C.c_int), ("n_e_local", C.c_int),
        ("n_x_sub", C.c_int), ("n_z_sub", C.c_int),
    ]

    def __init__(self, **kw):
        super().__init__(struct_size=C.sizeof(Grid), **kw)

    @property
    def local_bins(self) -> int:
        return self.n_theta_local * self.n_e_local * self.n_x_sub * self.n_z_sub


class Transport(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint32),
        ("E_fine_on", C.c_float), ("E_fine_off", C.c_float),
        ("weight_active_min", C.c_float),
        ("step_coarse", C.c_float),
        ("max_iterations", C.c_int),
        ("fail_fast_on_audit", C.c_int),
        ("step_rule", C.c_int),
        ("enable_straggling", C.c_int),
        ("enable_nuclear", C.c_int),
        ("sigma_x_initial", C.c_float),
    ]

    def __init__(self, **kw):
        d = dict(E_fine_on=10.0, E_fine_off=11.0, weight_active_min=1e-12, step_coarse=5.0,
                 max_iterations=100, fail_fast_on_audit=0, step_rule=0, enable_straggling=1,
                 enable_nuclear=1, sigma_x_initial=0.0)
        d.update(kw)
        super().__init__(struct_size=C.sizeof(Transport), **d)


class Spot(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint32),
        ("x0", C.c_float), ("z0", C.c_float), ("theta0", C.c_float),
        ("E0", C.c_float), ("W_total", C.c_float),
        ("sigma_x", C.c_float), ("sigma_theta", C.c_float), ("sigma_E", C.c_float),
        ("n_samples", C.c_int),
        ("random_seed", C.c_uint32),
    ]

    def __init__(self, **kw):
        d = dict(x0=0.0, z0=0.0, theta0=0.0, E0=150.0, W_total=1.0, sigma_x=0.0,
                 sigma_theta=0.0, sigma_E=0.0, n_samples=1000, random_seed=42)
        d.update(kw)
        super().__init__(struct_size=C.sizeof(Spot), **d)


class Audit(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint32),
        ("iterations", C.c_int), ("terminated", C.c_int),
        ("W_fail_iters", C.c_int), ("E_fail_iters", C.c_int),
        ("max_W_error", C.c_double), ("max_E_error", C.c_double),
        ("W_source_injected", C.c_double), ("W_source_out_of_grid", C.c_double),
        ("W_source_slot_drop", C.c_double),
        ("E_source_injected", C.c_double), ("E_source_out_of_grid", C.c_double),
        ("E_source_slot_drop", C.c_double), ("E_representation_loss", C.c_double),
        ("E_dep", C.c_double), ("W_cutoff", C.c_double), ("E_cutoff", C.c_double),
        ("W_nuclear", C.c_double), ("E_nuclear", C.c_double),
        ("W_boundary", C.c_double), ("E_boundary", C.c_double),
        ("W_transport_drop", C.c_double), ("E_transport_drop", C.c_double),
        ("E_audit_residual", C.c_double),
        ("W_leftover", C.c_double), ("E_leftover", C.c_double),
        ("max_cell_W_closure", C.c_double), ("max_cell_E_closure", C.c_double),
        ("cell_bin_updates", C.c_uint64),
        ("processed_cells", C.c_uint64), ("processed_components", C.c_uint64),
        ("cross_cell_emissions", C.c_uint64),
        ("surrogate_active", C.c_int), ("surrogate_fallback", C.c_int),
        ("fine_cell_bin_updates", C.c_uint64), ("surrogate_substeps", C.c_uint64),
    ]

    def __init__(self):
        super().__init__(struct_size=C.sizeof(Audit))

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_ if f != "struct_size"}


class K5Report(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint32),
        ("W_pass", C.c_int), ("E_pass", C.c_int),
        ("W_error", C.c_double), ("E_error", C.c_double),
        ("W_in", C.c_double), ("W_out", C.c_double), ("W_cutoff", C.c_double),
        ("W_nuclear", C.c_double), ("W_boundary", C.c_double),
        ("W_source_out_of_grid", C.c_double), ("W_source_slot_drop", C.c_double),
        ("W_transport_drop", C.c_double),
        ("E_in", C.c_double), ("E_out", C.c_double), ("E_dep", C.c_double),
        ("E_cutoff", C.c_double), ("E_nuclear", C.c_double), ("E_boundary", C.c_double),
        ("E_source_out_of_grid", C.c_double), ("E_source_slot_drop", C.c_double),
        ("E_transport_drop", C.c_double),
    ]

    def __init__(self):
        super().__init__(struct_size=C.sizeof(K5Report))

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_ if f != "struct_size"}


class Timing(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint32),
        ("enabled", C.c_int),
        ("iterations", C.c_uint64), ("transport_launches", C.c_uint64),
        ("gather_launches", C.c_uint64),
        ("mask_ms", C.c_double), ("transport_ms", C.c_double), ("gather_ms", C.c_double),
        ("audit_ms", C.c_double),
        ("cell_bin_updates", C.c_uint64),
        ("coarse_transport_ms", C.c_double), ("fine_transport_ms", C.c_double),
        ("coarse_updates", C.c_uint64), ("fine_updates", C.c_uint64),
        ("coarse_gather_ms", C.c_double), ("fine_gather_ms", C.c_double),
        ("coarse_gather_values", C.c_uint64), ("fine_gather_values", C.c_uint64),
    ]

    def __init__(self):
        super().__init__(struct_size=C.sizeof(Timing))

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_ if f != "struct_size"}


P = C.c_void_p
fp = C.POINTER(C.c_float)
dp = C.POINTER(C.c_double)
i32p = C.POINTER(C.c_int32)
u32p = C.POINTER(C.c_uint32)
u8p = C.POINTER(C.c_uint8)
ip = C.POINTER(C.c_int)

# Prototypes shared by libsm2d_hip.so (prefix "sm2d_") and the oracle (prefix "orc_").
class MlpTrainCfg(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint32),
        ("n_samples", C.c_int), ("batch", C.c_int), ("steps", C.c_int),
        ("lr", C.c_float), ("w_att", C.c_float), ("w_mom", C.c_float), ("seed", C.c_uint32),
    ]

    def __init__(self, **kw):
        d = dict(n_samples=1 << 22, batch=4096, steps=3000, lr=2e-3, w_att=1.0, w_mom=1.0, seed=7)
        d.update(kw)
        super().__init__(struct_size=C.sizeof(MlpTrainCfg), **d)


COMMON_PROTOS = {
    "energy_grid_log": [C.c_float, C.c_float, C.c_int, fp, fp],
    "energy_grid_piecewise": [fp, C.c_int, fp, fp, ip],
    "angular_grid": [C.c_float, C.c_float, C.c_int, fp],
    "build_rlut": [fp, C.c_int, fp, fp, fp, fp, fp],
    "fe_precompute": [C.c_float, C.c_float, C.c_int, C.c_float, fp, fp, fp],
    "post_convolve": [dp, C.c_int, C.c_int, C.c_float, fp, C.c_float],
    "ctx_destroy": [P],
    "lut_upload": [P, fp, fp, fp, fp, fp, fp, fp],
    "fe_upload": [P, fp, fp],
    "material_upload": [P, fp, fp],
    "reset": [P],
    "inject": [P, C.POINTER(Spot), C.POINTER(Audit)],
    "psi_load": [P, C.c_int, i32p, i32p, u32p, i32p, fp],
    "psi_dump": [P, u32p, C.c_int, i32p, i32p, u32p, i32p, fp, ip],
    "k1k2_mask_compact": [P, C.POINTER(Transport), u8p, u8p, i32p, ip, i32p, ip],
    "run_iterations": [P, C.POINTER(Transport), C.c_int, C.c_int, dp, C.POINTER(K5Report),
                       C.POINTER(Audit)],
    "k5_audit": [P, dp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                 C.c_double, C.c_int, C.POINTER(K5Report)],
    "swap": [P],
    "run_spot": [P, C.POINTER(Spot), C.POINTER(Transport), dp, C.POINTER(Audit)],
}


def bind(lib: C.CDLL, prefix: str, protos: dict) -> None:
    for name, args in protos.items():
        f = getattr(lib, prefix + name)
        f.argtypes = args
        f.restype = C.c_int
