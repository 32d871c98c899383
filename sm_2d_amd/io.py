"""Run configuration and file formats of the library (include/sm2d_io.h), from Python.

The INI loader, validation and runner derivation of the reference's host driver
(src/include/core/config_loader.hpp, incident_particle_config.hpp:272-391,
src/gpu/gpu_transport_runner.cpp:28-213), the dose writers of run_simulation.cpp:31-116 and
the raw phase-space CSV of src/cuda/k1k6_pipeline.cu:195-291.  All of these are host
functions of libsm2d_hip.so; none needs a GPU.  The device memory plan
(sm2d_memory_estimate / sm2d_memory_preflight, include/sm2d_hip.h) is here too.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi

MAX_ENERGY_GROUPS = 64
PATH_MAX = 512


class RunConfig(C.Structure):
    """sm2d_run_config: IncidentParticleConfig flattened (incident_particle_config.hpp:201-244)."""
    _fields_ = [
        ("struct_size", C.c_uint32),
        ("particle_type", C.c_int),
        ("particle_mass_amu", C.c_float), ("particle_charge_e", C.c_float),
        ("beam_profile", C.c_int),
        ("W_total", C.c_float),
        ("E_mean", C.c_float), ("E_sigma", C.c_float), ("E_min", C.c_float), ("E_max", C.c_float),
        ("x0", C.c_float), ("z0", C.c_float), ("sigma_x", C.c_float), ("sigma_z", C.c_float),
        ("width_x", C.c_float), ("width_z", C.c_float), ("radius", C.c_float),
        ("theta0", C.c_float), ("phi0", C.c_float), ("sigma_theta", C.c_float),
        ("sigma_phi", C.c_float),
        ("n_samples", C.c_int), ("random_seed", C.c_uint32), ("use_stratified", C.c_int),
        ("Nx", C.c_int), ("Nz", C.c_int), ("dx", C.c_float), ("dz", C.c_float),
        ("max_steps", C.c_int),
        ("N_theta", C.c_int), ("N_theta_local", C.c_int), ("N_E_local", C.c_int),
        ("n_energy_groups", C.c_int),
        ("energy_groups", (C.c_float * 3) * MAX_ENERGY_GROUPS),
        ("E_fine_on", C.c_float), ("E_fine_off", C.c_float), ("E_trigger", C.c_float),
        ("weight_active_min", C.c_float), ("E_coarse_max", C.c_float), ("step_coarse", C.c_float),
        ("n_steps_per_cell", C.c_int), ("fine_batch_max_cells", C.c_int),
        ("fine_halo_cells", C.c_int),
        ("preflight_vram_margin", C.c_float),
        ("validation_mode", C.c_int), ("fail_fast_on_audit", C.c_int), ("debug_dumps", C.c_int),
        ("max_iterations", C.c_int), ("log_level", C.c_int),
        ("output_dir", C.c_char * PATH_MAX), ("dose_2d_file", C.c_char * PATH_MAX),
        ("pdd_file", C.c_char * PATH_MAX), ("let_file", C.c_char * PATH_MAX),
        ("format", C.c_char * 32),
        ("normalize_dose", C.c_int), ("save_2d", C.c_int), ("save_pdd", C.c_int),
        ("save_lat_profiles", C.c_int),
    ]

    def __init__(self):
        super().__init__(struct_size=C.sizeof(RunConfig))

    def groups(self) -> list[tuple[float, float, float]]:
        return [tuple(self.energy_groups[k]) for k in range(self.n_energy_groups)]

    def as_dict(self) -> dict:
        out = {}
        for f, _ in self._fields_:
            if f == "struct_size":
                continue
            v = getattr(self, f)
            if f == "energy_groups":
                v = self.groups()
            elif isinstance(v, bytes):
                v = v.decode()
            out[f] = v
        return out


class MemoryPlan(C.Structure):
    """sm2d_memory_plan (include/sm2d_hip.h)."""
    _fields_ = [
        ("struct_size", C.c_uint32),
        ("listed_cells", C.c_uint32), ("occupied_cells", C.c_uint32),
        ("payload_slots", C.c_uint32),
        ("psi_bytes", C.c_uint64), ("bucket_bytes", C.c_uint64), ("cell_bytes", C.c_uint64),
        ("other_bytes", C.c_uint64), ("total_bytes", C.c_uint64), ("dense_bytes", C.c_uint64),
        ("free_bytes", C.c_uint64), ("usable_bytes", C.c_uint64),
        ("ok", C.c_int),
    ]

    def __init__(self):
        super().__init__(struct_size=C.sizeof(MemoryPlan))

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_ if f != "struct_size"}


PROTOS = {
    "config_defaults": [C.POINTER(RunConfig)],
    "config_load": [C.c_char_p, C.POINTER(RunConfig)],
    "config_validate": [C.POINTER(RunConfig)],
    "config_runtime": [C.POINTER(RunConfig), C.POINTER(abi.Grid), C.POINTER(abi.Transport),
                       C.POINTER(abi.Spot)],
    "write_dose_2d": [C.c_char_p, abi.dp, C.c_int, C.c_int, C.c_float, C.c_float, C.c_int],
    "write_pdd": [C.c_char_p, abi.dp, C.c_int, C.c_int, C.c_float, C.c_int],
    "write_psi_csv": [C.c_char_p, C.c_int, C.c_int, C.POINTER(abi.Grid), abi.fp, abi.fp, C.c_int,
                      abi.i32p, abi.i32p, abi.u32p, abi.i32p, abi.fp],
    "read_psi_csv": [C.c_char_p, C.c_int, abi.i32p, abi.i32p, abi.u32p, abi.i32p, abi.fp, abi.ip],
    "write_cells_csv": [C.c_char_p, C.POINTER(abi.Grid), abi.fp, abi.fp, C.c_int,
                        abi.i32p, abi.i32p, abi.u32p, abi.i32p, abi.fp],
    "memory_estimate": [C.POINTER(abi.Grid), C.c_int, C.POINTER(MemoryPlan)],
    "memory_preflight": [C.c_int, C.POINTER(abi.Grid), C.c_int, C.c_float, C.POINTER(MemoryPlan)],
    "ctx_memory": [C.c_void_p, C.POINTER(C.c_uint64)],
}


def _b(path) -> bytes:
    return str(path).encode()


def _arr(a, dt):
    return np.ascontiguousarray(a, dt)


def config_load(lib, path) -> RunConfig:
    """load_incident_particle_config; raises on an unreadable file."""
    c = RunConfig()
    lib.call("config_load", _b(path), C.byref(c))
    return c


def config_validate(lib, cfg: RunConfig) -> None:
    """IncidentParticleConfig::validate; raises LibError with the reference's message."""
    lib.call("config_validate", C.byref(cfg))


def config_runtime(lib, cfg: RunConfig):
    """(grid, transport, spot) as GPUTransportRunner::run derives them."""
    g, tp, sp = abi.Grid(), abi.Transport(), abi.Spot()
    lib.call("config_runtime", C.byref(cfg), C.byref(g), C.byref(tp), C.byref(sp))
    return g, tp, sp


def write_dose_2d(lib, path, edep, dx: float, dz: float, normalize: bool = True) -> None:
    e = _arr(edep, np.float64)
    Nz, Nx = e.shape
    lib.call("write_dose_2d", _b(path), e.ctypes.data_as(abi.dp), Nx, Nz, dx, dz, int(normalize))


def write_pdd(lib, path, edep, dz: float, normalize: bool = True) -> None:
    e = _arr(edep, np.float64)
    Nz, Nx = e.shape
    lib.call("write_pdd", _b(path), e.ctypes.data_as(abi.dp), Nx, Nz, dz, int(normalize))


def write_psi_csv(lib, path, rows: dict, grid: abi.Grid, theta_edges, E_edges, iteration: int = 1,
                  decimals: int = 8) -> None:
    """rows: dict of arrays cell, slot, bid, lidx, w (as Engine.psi_dump returns them)."""
    cell, slot, lidx = (_arr(rows[k], np.int32) for k in ("cell", "slot", "lidx"))
    bid, w = _arr(rows["bid"], np.uint32), _arr(rows["w"], np.float32)
    te, ee = _arr(theta_edges, np.float32), _arr(E_edges, np.float32)
    lib.call("write_psi_csv", _b(path), iteration, decimals, C.byref(grid),
             te.ctypes.data_as(abi.fp), ee.ctypes.data_as(abi.fp), len(cell),
             cell.ctypes.data_as(abi.i32p), slot.ctypes.data_as(abi.i32p),
             bid.ctypes.data_as(abi.u32p), lidx.ctypes.data_as(abi.i32p), w.ctypes.data_as(abi.fp))


def write_cells_csv(lib, path, rows: dict, grid: abi.Grid, theta_edges, E_edges) -> None:
    """Per-cell summary (debug_cells_iter_NN_<stage>.csv format) of psi_dump rows."""
    cell, slot, lidx = (_arr(rows[k], np.int32) for k in ("cell", "slot", "lidx"))
    bid, w = _arr(rows["bid"], np.uint32), _arr(rows["w"], np.float32)
    te, ee = _arr(theta_edges, np.float32), _arr(E_edges, np.float32)
    lib.call("write_cells_csv", _b(path), C.byref(grid), te.ctypes.data_as(abi.fp),
             ee.ctypes.data_as(abi.fp), len(cell), cell.ctypes.data_as(abi.i32p),
             slot.ctypes.data_as(abi.i32p), bid.ctypes.data_as(abi.u32p),
             lidx.ctypes.data_as(abi.i32p), w.ctypes.data_as(abi.fp))


def read_psi_csv(lib, path) -> dict:
    n = C.c_int(0)
    lib.call("read_psi_csv", _b(path), 0, None, None, None, None, None, C.byref(n))
    out = {"cell": np.zeros(n.value, np.int32), "slot": np.zeros(n.value, np.int32),
           "bid": np.zeros(n.value, np.uint32), "lidx": np.zeros(n.value, np.int32),
           "w": np.zeros(n.value, np.float32)}
    m = C.c_int(0)
    lib.call("read_psi_csv", _b(path), n.value, out["cell"].ctypes.data_as(abi.i32p),
             out["slot"].ctypes.data_as(abi.i32p), out["bid"].ctypes.data_as(abi.u32p),
             out["lidx"].ctypes.data_as(abi.i32p), out["w"].ctypes.data_as(abi.fp), C.byref(m))
    return out


def memory_estimate(lib, grid: abi.Grid, n_spots: int = 1) -> MemoryPlan:
    p = MemoryPlan()
    lib.call("memory_estimate", C.byref(grid), n_spots, C.byref(p))
    return p


def memory_preflight(lib, grid: abi.Grid, n_spots: int = 1, margin: float = 0.85,
                     device: int = 0) -> MemoryPlan:
    p = MemoryPlan()
    lib.call("memory_preflight", device, C.byref(grid), n_spots, margin, C.byref(p))
    return p
