"""Readers for the reference's one-step dumps under tests/golden/one_step_case_{A,B,C,D}, and
the dose-error measure used against the binary64 oracle.

The dumps were written by the reference's debug path (src/cuda/k1k6_pipeline.cu:195-291 for
the raw ψ rows, :2127-2157 for the channel totals) for the INI cases stored beside them.
"""
from __future__ import annotations

import configparser
import csv
import os

import numpy as np

from sm_2d_amd import abi

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def case_dir(case: str) -> str:
    return os.path.join(GOLDEN, f"one_step_case_{case}")


def read_ini(case: str) -> configparser.ConfigParser:
    cp = configparser.ConfigParser(inline_comment_prefixes=("#", ";"))
    cp.read(os.path.join(case_dir(case), f"case_{case.lower()}.ini"))
    return cp


def read_psi(case: str, which: str) -> dict:
    """which = 'initial' (iter 00) or 'after' (iter 01 after K4)."""
    name = ("debug_ps_raw_iter_00_initial.csv" if which == "initial"
            else "debug_ps_raw_iter_01_after_K4.csv")
    rows = {k: [] for k in ("cell", "slot", "bid", "lidx", "w")}
    with open(os.path.join(case_dir(case), name)) as f:
        for r in csv.DictReader(f):
            rows["cell"].append(int(r["cell"]))
            rows["slot"].append(int(r["slot"]))
            rows["bid"].append(int(r["bid"]))
            rows["lidx"].append(int(r["lidx"]))
            rows["w"].append(float(r["weight"]))
    return {k: np.asarray(v) for k, v in rows.items()}


def read_totals(case: str) -> dict:
    out = {}
    with open(os.path.join(case_dir(case), "debug_channels_totals_iter01.csv")) as f:
        for r in csv.DictReader(f):
            if r["iter"] == "1":
                out[r["channel"]] = float(r["value"])
    return out


def problem(case: str):
    """(grid, transport, spot, theta range, N_theta) of a case, as the reference runner builds
    them (gpu_transport_runner.cpp:63-98,172-173) with the layout 4x2x8x4."""
    cp = read_ini(case)
    Nx = cp.getint("grid", "Nx")
    Nz = cp.getint("grid", "Nz")
    dx = cp.getfloat("grid", "dx_mm")
    dz = cp.getfloat("grid", "dz_mm")
    NT = cp.getint("transport", "N_theta", fallback=36)
    E_on = cp.getfloat("transport", "E_fine_on_MeV")
    E_off = cp.getfloat("transport", "E_fine_off_MeV")
    spot = abi.Spot(
        x0=cp.getfloat("spatial", "x0_mm"), z0=cp.getfloat("spatial", "z0_mm"),
        theta0=cp.getfloat("angular", "theta0_rad"), E0=cp.getfloat("energy", "mean_MeV"),
        W_total=cp.getfloat("beam", "weight"), sigma_x=cp.getfloat("spatial", "sigma_x_mm"),
        sigma_theta=cp.getfloat("angular", "sigma_theta_rad"),
        sigma_E=cp.getfloat("energy", "sigma_MeV"),
        n_samples=cp.getint("sampling", "n_samples"),
        random_seed=cp.getint("sampling", "random_seed"))
    tp = abi.Transport(E_fine_on=E_on, E_fine_off=E_off,
                       max_iterations=cp.getint("transport", "max_iterations"),
                       sigma_x_initial=spot.sigma_x)
    grid = dict(Nx=Nx, Nz=Nz, dx=dx, dz=dz, x_min=-(Nx * dx) / 2.0, z_min=0.0, N_theta=NT,
                theta_min=-0.35, theta_max=0.35, n_theta_local=4, n_e_local=2, n_x_sub=8,
                n_z_sub=4)
    return grid, tp, spot


def rows_as_set(rows: dict, wdigits: int) -> dict:
    """{(cell, bid, lidx): weight} with slots canonicalised away."""
    out = {}
    for c, b, l, w in zip(rows["cell"], rows["bid"], rows["lidx"], rows["w"]):
        key = (int(c), int(b), int(l))
        out[key] = out.get(key, 0.0) + float(w)
    return out


def dose_errors(e, e64):
    """Largest relative difference of dose grid `e` from `e64` on voxels above 1 % and 10 % of
    Dmax (keys 0.01, 0.1), and of the PDD on depths above 1 % of its maximum (key "pdd")."""
    D = e64.max()
    out = {}
    for f in (0.01, 0.1):
        m = e64 > f * D
        out[f] = float((np.abs(e - e64)[m] / e64[m]).max())
    p, p64 = e.sum(axis=1), e64.sum(axis=1)
    m = p64 > 0.01 * p64.max()
    out["pdd"] = float((np.abs(p - p64)[m] / p64[m]).max())
    return out
