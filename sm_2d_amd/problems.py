"""Named transport problems: the benchmark configurations and the small parity cases.

Each problem is the tuple the reference runner would build from its INI file
(src/gpu/gpu_transport_runner.cpp:63-98,172-173): spatial grid, phase-space grids, R-LUT,
transport settings and the beam spot(s).  Inputs are synthetic (seeded), as the workloads are
defined by their parameters only.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import abi
from .engine import DEFAULT_ENERGY_GROUPS, Library, Tables


@dataclass
class Problem:
    name: str
    grid: dict
    energy: object            # "piecewise" or ("log", E_min, E_max, N_E)
    transport: dict
    spots: list = field(default_factory=list)   # list of abi.Spot kwargs
    materials: object = None  # None (water) or (rho[Nz], x0_ratio[Nz])
    note: str = ""

    def make_grid(self, N_E: int) -> abi.Grid:
        return abi.Grid(N_E=N_E, **self.grid)

    def make_transport(self, **over) -> abi.Transport:
        d = dict(self.transport)
        d.update(over)
        return abi.Transport(**d)

    def make_spot(self, i: int = 0) -> abi.Spot:
        return abi.Spot(**self.spots[i])

    def tables(self, lib: Library) -> Tables:
        g = self.grid
        return lib.make_tables(g["N_theta"], g["theta_min"], g["theta_max"], self.energy,
                               DEFAULT_ENERGY_GROUPS)


def _grid(Nx, Nz, dx, dz, N_theta, tmin, tmax, layout):
    tl, el, xs, zs = layout
    return dict(Nx=Nx, Nz=Nz, dx=dx, dz=dz, x_min=-(Nx * dx) / 2.0, z_min=0.0, N_theta=N_theta,
                theta_min=tmin, theta_max=tmax, n_theta_local=tl, n_e_local=el, n_x_sub=xs,
                n_z_sub=zs)


LAYOUT_REF = (4, 2, 8, 4)   # src/include/core/local_bins.hpp:19-23
LAYOUT_CFG = (8, 4, 4, 1)   # 128 local bins: 8 theta x 4 E x 4 x_sub


def config1(layout=LAYOUT_CFG, step_rule=0, n_samples=1000) -> Problem:
    """70 MeV, sigma_x 3 mm, 1 % energy spread, water 40 x 60 mm at 0.5 mm (80 x 120),
    256 log E bins 0.1-250 MeV, 512 theta bins over [-pi/2, pi/2]."""
    return Problem(
        "config1_70MeV_80x120", _grid(80, 120, 0.5, 0.5, 512, -math.pi / 2, math.pi / 2, layout),
        ("log", 0.1, 250.0, 256),
        dict(E_fine_on=10.0, E_fine_off=11.0, max_iterations=400, step_rule=step_rule),
        [dict(E0=70.0, sigma_x=3.0, sigma_E=0.7, sigma_theta=0.0, n_samples=n_samples, random_seed=42)])


def config2(layout=LAYOUT_CFG, n_samples=1000) -> Problem:
    """150 MeV, sigma_x 3.8 mm, 1 % energy spread, 2 mrad, water 100 x 200 mm at 0.5 mm
    (200 x 400), 256 log E bins, 512 theta bins over [-pi/2, pi/2], 128 local bins."""
    return Problem(
        "config2_150MeV_200x400", _grid(200, 400, 0.5, 0.5, 512, -math.pi / 2, math.pi / 2, layout),
        ("log", 0.1, 250.0, 256),
        dict(E_fine_on=10.0, E_fine_off=11.0, max_iterations=1000),
        [dict(E0=150.0, sigma_x=3.8, sigma_E=1.5, sigma_theta=0.002, n_samples=n_samples,
              random_seed=42)])


def reference_grid_case(Nx=80, Nz=120, E0=70.0, sigma_x=3.0, n_samples=1000,
                        layout=LAYOUT_REF, N_theta=36, max_iterations=400) -> Problem:
    """The reference runner's own defaults: piecewise 1029-bin E grid, +-0.35 rad theta."""
    return Problem(
        f"refgrid_{E0:g}MeV_{Nx}x{Nz}", _grid(Nx, Nz, 0.5, 0.5, N_theta, -0.35, 0.35, layout),
        "piecewise", dict(E_fine_on=10.0, E_fine_off=11.0, max_iterations=max_iterations),
        [dict(E0=E0, sigma_x=sigma_x, sigma_E=0.01 * E0, sigma_theta=0.0, n_samples=n_samples,
              random_seed=42)])


def slab_materials(Nz: int, dz: float, seed: int = 7):
    """3-6 seeded slabs of water / bone / lung along depth, thicknesses uniform in 5-40 mm."""
    rng = np.random.default_rng(seed)
    mats = {"water": (1.0, 1.0), "bone": (1.85, 0.55), "lung": (0.26, 1.0 / 0.26)}
    rho = np.ones(Nz, np.float32)
    x0r = np.ones(Nz, np.float32)
    n = int(rng.integers(3, 7))
    z = 0.0
    depth = Nz * dz
    for _ in range(n):
        t = float(rng.uniform(5.0, 40.0))
        name = ["water", "bone", "lung"][int(rng.integers(0, 3))]
        z0 = z + float(rng.uniform(0.0, 10.0))
        i0, i1 = int(z0 / dz), min(Nz, int((z0 + t) / dz))
        rho[i0:i1], x0r[i0:i1] = mats[name]
        z = z0 + t
        if z >= depth:
            break
    return rho, x0r


def config3(layout=LAYOUT_CFG, n_samples=1000, slab_seed: int = 4) -> Problem:
    """230 MeV, sigma_x 5 mm, 1 % energy spread, 100 x 300 mm at 0.25 mm (400 x 1200) through
    3-6 seeded water / bone / lung slabs, 256 log E bins, 512 theta bins over [-pi/2, pi/2]."""
    p = Problem(
        "config3_230MeV_slabs_400x1200",
        _grid(400, 1200, 0.25, 0.25, 512, -math.pi / 2, math.pi / 2, layout),
        ("log", 0.1, 250.0, 256),
        dict(E_fine_on=10.0, E_fine_off=11.0, max_iterations=3000),
        [dict(E0=230.0, sigma_x=5.0, sigma_E=2.3, sigma_theta=0.0, n_samples=n_samples,
              random_seed=42)])
    p.materials = slab_materials(1200, 0.25, slab_seed)
    return p


def config4(n_spots: int = 256, layout=LAYOUT_CFG, n_samples=1000, seed: int = 2024,
            Nz: int = 600) -> Problem:
    """Batch of pencil beams on water 100 x 300 mm at 0.5 mm (200 x 600): E0 ~ U(70, 230) MeV,
    x0 ~ U(-40, 40) mm, sigma_x ~ U(3, 6) mm, 1 % energy spread, seeded."""
    rng = np.random.default_rng(seed)
    E0 = rng.uniform(70.0, 230.0, n_spots)
    x0 = rng.uniform(-40.0, 40.0, n_spots)
    sx = rng.uniform(3.0, 6.0, n_spots)
    spots = [dict(E0=float(E0[i]), x0=float(x0[i]), sigma_x=float(sx[i]), sigma_E=0.01 * float(E0[i]),
                  sigma_theta=0.0, n_samples=n_samples, random_seed=int(1000 + i))
             for i in range(n_spots)]
    return Problem(
        f"config4_{n_spots}spots_200x{Nz}",
        _grid(200, Nz, 0.5, 0.5, 512, -math.pi / 2, math.pi / 2, layout),
        ("log", 0.1, 250.0, 256),
        dict(E_fine_on=10.0, E_fine_off=11.0, max_iterations=2000), spots)


def config4_fine_grid(n_spots: int = 256, n_samples=1000, seed: int = 2024) -> Problem:
    """The batch of config4 on the reference runner's default piecewise 1029-bin energy grid
    (src/include/core/incident_particle_config.hpp:164-171).  Coarse steps lose more than half
    a bin there, so components slow down into the fine transport (where the learned physics
    acts).  A few fine components reach a fixed point of the reference algorithm (a hash-seeded
    draw with dE = 0 that keeps the component in its sub-bin, repeated with the same seed every
    iteration), so the run is capped at Nz + 50 iterations and reports the leftover weight."""
    p = config4(n_spots=n_spots, n_samples=n_samples, seed=seed)
    p.name = f"config4_fine_grid_{n_spots}spots_200x600"
    p.energy = "piecewise"
    p.transport = dict(p.transport, max_iterations=600 + 50)
    p.note = (f"{n_spots} pencil beams, E0 U(70,230) MeV, x0 U(-40,40) mm, sigma_x U(3,6) mm, water "
              "200x600 @ 0.5 mm, piecewise 1029-bin E grid (reference runner default) x 512 theta, "
              "128 local bins, 650 iterations")
    return p


def config5(n_spots: int = 256, energy="piecewise", n_samples=1000, seed: int = 2024,
            slab_seed: int = 4) -> Problem:
    """The config4 spot distribution on the config3 slab phantom (water / bone / lung slabs,
    100 x 300 mm at 0.25 mm = 400 x 1200 cells): the learned-physics workload.  With the
    256-bin log grid of the benchmark configurations no component reaches the fine transport
    (coarse steps re-centre every component in its energy bin); the piecewise 1029-bin grid
    (default) is where the fine transport and the learned physics act (capped at Nz + 50
    iterations, see config4_fine_grid)."""
    p = config4(n_spots=n_spots, n_samples=n_samples, seed=seed)
    p3 = config3(slab_seed=slab_seed)
    p.grid = p3.grid
    p.materials = p3.materials
    if energy == "piecewise":
        p.energy = "piecewise"
        p.transport = dict(p.transport, max_iterations=1200 + 50)
        tag, grid_txt = "fine_grid", "piecewise 1029-bin E grid (reference runner default), 1250 iterations"
    else:
        p.energy = ("log", 0.1, 250.0, 256)
        p.transport = dict(p.transport, max_iterations=3000)
        tag, grid_txt = "log256", "256 log E bins 0.1-250 MeV"
    p.name = f"config5_{tag}_{n_spots}spots_400x1200"
    p.note = (f"{n_spots} pencil beams, E0 U(70,230) MeV, x0 U(-40,40) mm, sigma_x U(3,6) mm, seeded "
              f"water/bone/lung slabs 400x1200 @ 0.25 mm, {grid_txt} x 512 theta, 128 local bins")
    return p


def setup_engine(lib: Library, prob: Problem, engine_cls=None, spot_index: int = 0,
                 n_spots: int = 1):
    """Builds tables, creates a context (of `n_spots` spot slots) and uploads FE moments for
    spot `spot_index` (the stage entry points use them; full-spot calls compute their own)."""
    from .engine import Engine
    cls = engine_cls or Engine
    tab = prob.tables(lib)
    grid = prob.make_grid(len(tab.E_rep))
    eng = cls(lib, grid, tab, n_spots=n_spots) if n_spots > 1 else cls(lib, grid, tab)
    sp = prob.make_spot(spot_index)
    A, B, st = lib.fe_precompute(sp.E0, sp.sigma_x, grid.Nz, grid.dz)
    eng.upload_fe(A, B)
    eng.sigma_total = st
    if prob.materials is not None:
        eng.upload_material(*prob.materials)
    return eng
