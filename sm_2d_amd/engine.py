"""Host-side driver of a transport-step library through its C ABI.

`Engine` wraps one context of a library that exports the entry points of
include/sm2d_hip.h (the HIP library, prefix ``sm2d_``).  It mirrors the reference's
boundary `sm_2d::run_k1k6_pipeline_transport` (src/cuda/gpu_transport_wrapper.cu:207-641):
grids and the R-LUT are built on the host exactly as the reference runner does
(src/gpu/gpu_transport_runner.cpp:63-98,145), uploaded once, and spots are run against the
persistent context.  The same class drives the CPU oracle in the tests (with another
prefix), so parity tests call identical methods on both.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import abi

DEFAULT_ENERGY_GROUPS = ((0.1, 2.0, 0.1), (2.0, 20.0, 0.2), (20.0, 100.0, 0.25),
                         (100.0, 250.0, 0.25))


def _f(a):
    return a.ctypes.data_as(abi.fp)


def _d(a):
    return a.ctypes.data_as(abi.dp)


class LibError(RuntimeError):
    pass


@dataclass
class Tables:
    """Phase-space grids and R-LUT of one problem (all float32 host arrays)."""
    theta_edges: np.ndarray
    E_edges: np.ndarray
    E_rep: np.ndarray
    R: np.ndarray
    S: np.ndarray
    log_E: np.ndarray
    log_R: np.ndarray
    log_S: np.ndarray


class Library:
    """A loaded C-ABI library (HIP product or oracle) with bound prototypes."""

    def __init__(self, path: str, prefix: str, extra_protos: dict | None = None):
        self.path = path
        self.prefix = prefix
        self.lib = C.CDLL(path)
        protos = dict(abi.COMMON_PROTOS)
        protos["ctx_create"] = ([C.c_int, C.POINTER(abi.Grid), C.POINTER(C.c_void_p)]
                                if prefix == "sm2d_" else
                                [C.POINTER(abi.Grid), C.POINTER(C.c_void_p)])
        protos["math_eval"] = [C.c_int, abi.fp, abi.fp, C.c_int]
        if extra_protos:
            protos.update(extra_protos)
        abi.bind(self.lib, prefix, protos)
        # attribute access caches the function object (item access would make a new one)
        self._last_error = getattr(self.lib, prefix + "last_error")
        self._last_error.restype = C.c_char_p
        self._last_error.argtypes = []

    def call(self, name: str, *args) -> int:
        rc = getattr(self.lib, self.prefix + name)(*args)
        if rc != abi.OK:
            raw = self._last_error()
            msg = raw.decode(errors="replace") if raw else ""
            raise LibError(f"{self.prefix}{name} failed ({rc}): {msg}")
        return rc

    # ---- host table builders ---------------------------------------------------------------
    def energy_grid_log(self, E_min: float, E_max: float, N_E: int):
        edges = np.zeros(N_E + 1, np.float32)
        rep = np.zeros(N_E, np.float32)
        self.call("energy_grid_log", E_min, E_max, N_E, _f(edges), _f(rep))
        return edges, rep

    def energy_grid_piecewise(self, groups=DEFAULT_ENERGY_GROUPS):
        g = np.asarray(groups, np.float32).reshape(-1)
        n = C.c_int(0)
        self.call("energy_grid_piecewise", _f(g), len(groups), None, None, C.byref(n))
        edges = np.zeros(n.value + 1, np.float32)
        rep = np.zeros(n.value, np.float32)
        self.call("energy_grid_piecewise", _f(g), len(groups), _f(edges), _f(rep), C.byref(n))
        return edges, rep

    def angular_grid(self, theta_min: float, theta_max: float, N_theta: int):
        e = np.zeros(N_theta + 1, np.float32)
        self.call("angular_grid", theta_min, theta_max, N_theta, _f(e))
        return e

    def build_rlut(self, rep: np.ndarray):
        n = len(rep)
        out = [np.zeros(n, np.float32) for _ in range(5)]
        rep = np.ascontiguousarray(rep, np.float32)
        self.call("build_rlut", _f(rep), n, *[_f(o) for o in out])
        return out

    def fe_precompute(self, E0: float, sigma_x0: float, Nz: int, dz: float):
        A, B, st = (np.zeros(Nz, np.float32) for _ in range(3))
        self.call("fe_precompute", E0, sigma_x0, Nz, dz, _f(A), _f(B), _f(st))
        return A, B, st

    def post_convolve(self, edep: np.ndarray, dx: float, sigma_total: np.ndarray,
                      sigma_x0: float) -> np.ndarray:
        e = np.ascontiguousarray(edep, np.float64).copy()
        Nz, Nx = e.shape
        st = np.ascontiguousarray(sigma_total, np.float32)
        self.call("post_convolve", _d(e), Nx, Nz, dx, _f(st), sigma_x0)
        return e

    def math_eval(self, which: int, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros_like(x)
        self.call("math_eval", which, _f(x), _f(y), len(x))
        return y

    def make_tables(self, N_theta: int, theta_min: float, theta_max: float,
                    energy: str | tuple = "piecewise", groups=DEFAULT_ENERGY_GROUPS) -> Tables:
        """Grids + R-LUT as the reference runner builds them (gpu_transport_runner.cpp:89-98,145).

        energy="piecewise" uses `groups`; energy=("log", E_min, E_max, N_E) a log grid."""
        th = self.angular_grid(theta_min, theta_max, N_theta)
        if energy == "piecewise":
            E_edges, rep = self.energy_grid_piecewise(groups)
        else:
            _, emin, emax, ne = energy
            E_edges, rep = self.energy_grid_log(emin, emax, ne)
        R, S, lE, lR, lS = self.build_rlut(rep)
        return Tables(th, E_edges, rep, R, S, lE, lR, lS)


class Engine:
    """One context (grid + tables) of a Library."""

    def __init__(self, lib: Library, grid: abi.Grid, tables: Tables, device: int = 0,
                 n_spots: int = 1):
        self.L = lib
        self.grid = grid
        self.tables = tables
        self.N = grid.Nx * grid.Nz
        self.LB = grid.local_bins
        h = C.c_void_p()
        self.n_spots = n_spots
        if n_spots > 1:
            lib.call("ctx_create_batch", device, C.byref(grid), n_spots, C.byref(h))
        elif lib.prefix == "sm2d_":
            lib.call("ctx_create", device, C.byref(grid), C.byref(h))
        else:
            lib.call("ctx_create", C.byref(grid), C.byref(h))
        self.h = h
        t = tables
        self._keep = [np.ascontiguousarray(a, np.float32) for a in
                      (t.theta_edges, t.E_edges, t.R, t.S, t.log_E, t.log_R, t.log_S)]
        lib.call("lut_upload", self.h, *[_f(a) for a in self._keep])
        lib.call("fe_upload", self.h, None, None)

    def close(self):
        if self.h:
            self.L.call("ctx_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- configuration -------------------------------------------------------------------
    def upload_fe(self, A, B):
        if A is None:
            self.L.call("fe_upload", self.h, None, None)
            return
        self._fe = (np.ascontiguousarray(A, np.float32), np.ascontiguousarray(B, np.float32))
        self.L.call("fe_upload", self.h, _f(self._fe[0]), _f(self._fe[1]))

    def upload_material(self, rho=None, x0_ratio=None):
        if rho is None:
            self.L.call("material_upload", self.h, None, None)
            return
        self._mat = (np.ascontiguousarray(rho, np.float32),
                     np.ascontiguousarray(x0_ratio, np.float32))
        self.L.call("material_upload", self.h, _f(self._mat[0]), _f(self._mat[1]))

    # ---- stages ------------------------------------------------------------------------------
    def reset(self):
        self.L.call("reset", self.h)

    def inject(self, spot: abi.Spot) -> abi.Audit:
        a = abi.Audit()
        self.L.call("inject", self.h, C.byref(spot), C.byref(a))
        return a

    def psi_load(self, rows: dict):
        n = len(rows["cell"])
        cols = dict(cell=np.int32, slot=np.int32, bid=np.uint32, lidx=np.int32, w=np.float32)
        arr = {k: np.ascontiguousarray(rows[k], t) for k, t in cols.items()}
        self.L.call("psi_load", self.h, n, arr["cell"].ctypes.data_as(abi.i32p),
                    arr["slot"].ctypes.data_as(abi.i32p), arr["bid"].ctypes.data_as(abi.u32p),
                    arr["lidx"].ctypes.data_as(abi.i32p), _f(arr["w"]))

    def psi_dump(self, capacity: int = 1 << 20):
        bids = np.zeros(self.N * abi.KB, np.uint32)
        out = dict(cell=np.zeros(capacity, np.int32), slot=np.zeros(capacity, np.int32),
                   bid=np.zeros(capacity, np.uint32), lidx=np.zeros(capacity, np.int32),
                   w=np.zeros(capacity, np.float32))
        n = C.c_int(0)
        self.L.call("psi_dump", self.h, bids.ctypes.data_as(abi.u32p), capacity,
                    out["cell"].ctypes.data_as(abi.i32p), out["slot"].ctypes.data_as(abi.i32p),
                    out["bid"].ctypes.data_as(abi.u32p), out["lidx"].ctypes.data_as(abi.i32p),
                    _f(out["w"]), C.byref(n))
        if n.value > capacity:
            return self.psi_dump(n.value)
        rows = {k: v[: n.value].copy() for k, v in out.items()}
        return bids.reshape(self.N, abi.KB), rows

    def mask_compact(self, tp: abi.Transport, mask_prev=None):
        mask = np.zeros(self.N, np.uint8)
        act = np.zeros(self.N, np.int32)
        crs = np.zeros(self.N, np.int32)
        na, nc = C.c_int(0), C.c_int(0)
        mp = None
        if mask_prev is not None:
            self._mp = np.ascontiguousarray(mask_prev, np.uint8)
            mp = self._mp.ctypes.data_as(abi.u8p)
        self.L.call("k1k2_mask_compact", self.h, C.byref(tp), mp, mask.ctypes.data_as(abi.u8p),
                    act.ctypes.data_as(abi.i32p), C.byref(na), crs.ctypes.data_as(abi.i32p),
                    C.byref(nc))
        return mask, act[: na.value].copy(), crs[: nc.value].copy()

    def run_iterations(self, tp: abi.Transport, n_iter: int, include_source_terms: bool = False):
        t = np.zeros((7, self.N), np.float64)
        k5 = abi.K5Report()
        a = abi.Audit()
        self.L.call("run_iterations", self.h, C.byref(tp), n_iter, int(include_source_terms),
                    _d(t), C.byref(k5), C.byref(a))
        return t, k5, a

    def k5_audit(self, tallies_delta, W_src_out=0.0, W_src_slot=0.0, E_src_out=0.0,
                 E_src_slot=0.0, W_drop=0.0, E_drop=0.0, include_source_terms=False):
        td = np.ascontiguousarray(tallies_delta, np.float64)
        r = abi.K5Report()
        self.L.call("k5_audit", self.h, _d(td), W_src_out, W_src_slot, E_src_out, E_src_slot,
                    W_drop, E_drop, int(include_source_terms), C.byref(r))
        return r

    def swap(self):
        self.L.call("swap", self.h)

    def set_profiling(self, on: bool = True):
        self.L.call("set_profiling", self.h, int(on))

    def timing(self) -> abi.Timing:
        t = abi.Timing()
        self.L.call("get_timing", self.h, C.byref(t))
        return t

    def run_spot(self, spot: abi.Spot, tp: abi.Transport):
        edep = np.zeros((self.grid.Nz, self.grid.Nx), np.float64)
        a = abi.Audit()
        self.L.call("run_spot", self.h, C.byref(spot), C.byref(tp), _d(edep), C.byref(a))
        return edep, a

    def set_transport_waves(self, waves: int):
        """Fixes the transport launch shape (1, 2 or 4 waves per listed cell; 0 = tuned)."""
        self.L.call("set_transport_waves", self.h, waves)

    def set_gather_rows(self, kib: int):
        """Fixes the K4 gather launch shape (16, 8 or 4 KiB of LDS rows per destination; 0 = tuned)."""
        self.L.call("set_gather_rows", self.h, kib)

    def set_graphs(self, on: bool):
        """Captured-and-replayed chunks of iterations (True, the default) or eager launches."""
        self.L.call("set_graphs", self.h, 1 if on else 0)

    def tuner_report(self) -> dict:
        import json
        buf = C.create_string_buffer(8192)
        self.L.call("tuner_report", self.h, buf, len(buf))
        return json.loads(buf.value.decode())

    def surrogate_attach(self, mlp: "Mlp", n_nodes: int = 64):
        """Builds the learned-physics table of this context and turns it on."""
        self.L.call("surrogate_attach", self.h, mlp.h, n_nodes)

    def surrogate_check(self, tol: float = 0.25):
        """Re-validates the attached table against the plain physics; returns (nodes checked,
        nodes rejected, largest deviation, RMS deviation)."""
        st = np.zeros(4, np.float32)
        self.L.call("surrogate_check", self.h, C.c_float(tol), _f(st))
        return tuple(float(v) for v in st)

    def surrogate_stats(self):
        """Statistics of the last validation (attach or check)."""
        st = np.zeros(4, np.float32)
        self.L.call("surrogate_stats", self.h, _f(st))
        return tuple(float(v) for v in st)

    def surrogate_enable(self, on):
        """0 / False: plain physics; 1 / True: the attached table; 2: the attached network
        evaluated inside the fine transport."""
        self.L.call("surrogate_enable", self.h, int(on))

    def surrogate_export(self):
        """(table [n_mat][N_E][n_nodes][3], n_mat, n_nodes, rowmat [Nz])."""
        nm, nn = C.c_int(), C.c_int()
        self.L.call("surrogate_export", self.h, None, 0, C.byref(nm), C.byref(nn), None)
        tab = np.zeros(nm.value * self.grid.N_E * nn.value * 3, np.float32)
        rows = np.zeros(self.grid.Nz, np.int32)
        self.L.call("surrogate_export", self.h, _f(tab), len(tab), C.byref(nm), C.byref(nn),
                    rows.ctypes.data_as(abi.i32p))
        return tab, nm.value, nn.value, rows

    def surrogate_set_table(self, table, n_mat: int, n_nodes: int, rowmat):
        """Installs a table in the layout of surrogate_export and turns the learned physics on."""
        tab = np.ascontiguousarray(table, np.float32)
        rows = np.ascontiguousarray(rowmat, np.int32)
        if tab.size != n_mat * self.grid.N_E * n_nodes * 3 or rows.size != self.grid.Nz:
            raise ValueError("surrogate table / rowmat size does not match the grid")
        self.L.call("surrogate_set_table", self.h, _f(tab), n_mat, n_nodes, rows.ctypes.data_as(abi.i32p))

    def run_batch(self, spots, tp: abi.Transport, per_spot: bool = False):
        """Transports `spots` together; returns (summed EdepC [Nz, Nx] FP64, per-spot EdepC
        [n, Nz, Nx] FP32 or None, batch audit)."""
        n = len(spots)
        arr = (abi.Spot * n)(*spots)
        edep = np.zeros((self.grid.Nz, self.grid.Nx), np.float64)
        per = np.zeros((n, self.grid.Nz, self.grid.Nx), np.float32) if per_spot else None
        a = abi.Audit()
        self.L.call("run_batch", self.h, arr, n, C.byref(tp), _d(edep),
                    _f(per) if per is not None else None, C.byref(a))
        return edep, per, a

    def batch_spot_edep(self, s: int) -> np.ndarray:
        """Spot s's FP64 EdepC [Nz, Nx] of the last run_batch (equal to run_spot's)."""
        e = np.zeros((self.grid.Nz, self.grid.Nx), np.float64)
        self.L.call("batch_spot_edep", self.h, s, _d(e))
        return e


class Mlp:
    """The learned-physics network of the fine transport (sm2d_mlp_* of the C ABI)."""

    def __init__(self, lib: Library, seed: int = 1, device: int = 0):
        self.L = lib
        h = C.c_void_p()
        lib.call("mlp_create", device, seed, C.byref(h))
        self.h = h
        self.n = lib.lib.sm2d_mlp_param_count()

    def close(self):
        if self.h:
            self.L.call("mlp_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        p = np.zeros(self.n, np.float32)
        d = np.zeros(4, np.float32)
        self.L.call("mlp_get_params", self.h, _f(p), _f(d))
        return p, d

    def set_params(self, p=None, dict4=None):
        pp = None if p is None else np.ascontiguousarray(p, np.float32)
        dd = None if dict4 is None else np.ascontiguousarray(dict4, np.float32)
        self.L.call("mlp_set_params", self.h, None if pp is None else _f(pp), None if dd is None else _f(dd))

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        z = np.zeros((len(x), 33), np.float32)
        self.L.call("mlp_forward", self.h, _f(x), len(x), _f(z))
        return z

    def infer(self, x):
        """Decoded (ln sigma_E, ln sigma_theta, a) of normalised inputs by the in-kernel inference."""
        x = np.ascontiguousarray(x, np.float32)
        out = np.zeros((len(x), 3), np.float32)
        self.L.call("mlp_infer", self.h, _f(x), len(x), _f(out))
        return out

    def grad(self, x, labels, w_att: float = 1.0, w_mom: float = 1.0):
        x = np.ascontiguousarray(x, np.float32)
        lb = np.ascontiguousarray(labels, np.float32)
        g = np.zeros(self.n, np.float32)
        loss = np.zeros(1, np.float32)
        self.L.call("mlp_grad", self.h, _f(x), _f(lb), len(x), C.c_float(w_att), C.c_float(w_mom),
                    _f(loss), _f(g))
        return float(loss[0]), g

    def train(self, eng: Engine, n_hist: int = 20, **cfg):
        c = abi.MlpTrainCfg(**cfg)
        hist = np.zeros(n_hist, np.float32)
        self.L.call("mlp_train", self.h, eng.h, C.byref(c), _f(hist), n_hist)
        return hist

    def evaluate(self, eng: Engine, n: int = 65536, seed: int = 99):
        e = np.zeros(6, np.float32)
        self.L.call("mlp_evaluate", self.h, eng.h, n, seed, _f(e))
        return e
