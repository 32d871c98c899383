"""GPU tests of the drivers above the transport step: the INI driver executable
(sm_2d_amd/sm2d_run, run_simulation's counterpart), the batch sweep driver (sm_2d_amd.batch,
batch_run.py's counterpart) and the device memory plan.

The expected files are written by the same writers from the CPU oracle's EdepC (the writers'
formats are pinned against the reference's own outputs in tests/test_io.py), so a driver's
files must equal them byte for byte.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from sm_2d_amd import abi, io, problems
from sm_2d_amd.engine import Engine

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUNNER = os.path.join(ROOT, "sm_2d_amd", "sm2d_run")

SPOT_INI = """# 70 MeV Gaussian spot in water, 40 x 60 cells at 1 mm
[beam]
profile = gaussian
weight = 1.0
[energy]
mean_MeV = 70.0
sigma_MeV = 0.7
[spatial]
x0_mm = 0.0
sigma_x_mm = 2.0
[angular]
sigma_theta_rad = 0.002
[sampling]
n_samples = 400
random_seed = 7
[grid]
Nx = 40
Nz = 60
dx_mm = 1.0
dz_mm = 1.0
max_steps = 600
[transport]
N_theta_local = {tl}
N_E_local = {el}
[output]
output_dir = out
"""


def _expected(lib, orc, ini, out_dir):
    """dose_2d.txt and pdd.txt of an INI from the oracle's EdepC."""
    from oracle.pyoracle import OracleEngine
    cfg = io.config_load(lib, ini)
    g, tp, sp = io.config_runtime(lib, cfg)
    tab = orc.make_tables(g.N_theta, g.theta_min, g.theta_max, groups=cfg.groups())
    e = OracleEngine(orc, g, tab)
    edep, _ = e.run_spot(sp, tp)
    _, _, st = orc.fe_precompute(sp.E0, sp.sigma_x, g.Nz, g.dz)
    edep = orc.post_convolve(edep, g.dx, st, sp.sigma_x)
    os.makedirs(out_dir, exist_ok=True)
    io.write_dose_2d(lib, os.path.join(out_dir, "dose_2d.txt"), edep, g.dx, g.dz, True)
    io.write_pdd(lib, os.path.join(out_dir, "pdd.txt"), edep, g.dz, True)
    return edep


def _same_files(a, b):
    for f in ("dose_2d.txt", "pdd.txt"):
        assert open(os.path.join(a, f)).read() == open(os.path.join(b, f)).read(), f


@pytest.mark.parametrize("layout", [(4, 2), (8, 4)], ids=["ref", "cfg"])
def test_driver_writes_the_oracles_files(hip, orc, tmp_path, layout):
    ini = tmp_path / "spot.ini"
    ini.write_text(SPOT_INI.format(tl=layout[0], el=layout[1]))
    print("driver", layout, flush=True)
    r = subprocess.run([RUNNER, str(ini), "--dump-psi"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "=== Simulation Complete ===" in r.stdout
    _expected(hip, orc, ini, tmp_path / "expected")
    _same_files(tmp_path / "out", tmp_path / "expected")
    # the injected phase space in the reference's initial-dump format, as the oracle injects it
    from oracle.pyoracle import OracleEngine
    cfg = io.config_load(hip, ini)
    g, tp, sp = io.config_runtime(hip, cfg)
    tab = orc.make_tables(g.N_theta, g.theta_min, g.theta_max)
    o = OracleEngine(orc, g, tab)
    A, B, _ = orc.fe_precompute(sp.E0, sp.sigma_x, g.Nz, g.dz)
    o.upload_fe(A, B)
    o.reset()
    o.inject(sp)
    _, rows = o.psi_dump()
    io.write_psi_csv(hip, tmp_path / "init.csv", rows, g, tab.theta_edges, tab.E_edges, 0, 6)
    assert (open(tmp_path / "out" / "debug_ps_raw_iter_00_initial.csv").read() ==
            open(tmp_path / "init.csv").read())
    # after-K4 dumps at the reference's iterations, equal to the oracle's phase space there
    done, checked = 0, []
    for it in list(range(1, 11)) + [50, 100]:
        o.run_iterations(tp, it - done, include_source_terms=(done == 0))
        done = it
        raw = tmp_path / "out" / f"debug_ps_raw_iter_{it:02d}_after_K4.csv"
        cells = tmp_path / "out" / f"debug_cells_iter_{it:02d}_after_K4.csv"
        _, rows = o.psi_dump()
        if len(rows["w"]) == 0:  # the transport ended before this iteration: nothing dumped
            assert not raw.exists() and not cells.exists()
            break
        io.write_psi_csv(hip, tmp_path / "k4.csv", rows, g, tab.theta_edges, tab.E_edges, it, 8)
        io.write_cells_csv(hip, tmp_path / "cells.csv", rows, g, tab.theta_edges, tab.E_edges)
        assert open(raw).read() == open(tmp_path / "k4.csv").read(), it
        assert open(cells).read() == open(tmp_path / "cells.csv").read(), it
        checked.append(it)
    assert checked[:10] == list(range(1, 11)) and 50 in checked, checked


def test_driver_reports_errors(tmp_path):
    bad = tmp_path / "bad.ini"
    bad.write_text("[transport]\nE_fine_on = 20\nE_fine_off = 10\n")
    r = subprocess.run([RUNNER, str(bad)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert "Error: validate: IncidentParticleConfig: transport.E_fine_off must be >= transport.E_fine_on" in r.stderr
    r = subprocess.run([RUNNER, str(tmp_path / "missing.ini")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "Failed to load configuration file" in r.stderr


def test_batch_sweep_equals_single_runs(hip, orc, tmp_path):
    """Four runs (2 energies x 2 positions) in one batched launch; every run directory holds
    the files the oracle's single-spot EdepC gives for its sim_run.ini."""
    from sm_2d_amd import batch
    tpl = tmp_path / "template.ini"
    tpl.write_text(SPOT_INI.format(tl=4, el=2))
    yml = tmp_path / "b.yaml"
    yml.write_text(f"""template: "{tpl}"
output_base: "{tmp_path / 'runs'}"
sweep:
  energy:
    section: "[energy]"
    mean_MeV: [60, 70]
  position:
    section: "[spatial]"
    x0_mm: [-3, 3]
execution:
  parallel: 1
""")
    runner = batch.BatchRunner(str(yml), lib=hip)
    res = runner.run_all()
    assert res["completed"] == 4 and res["launches"] == 1 and not runner.failed
    for run in runner.runs:
        d = run["dir"]
        _expected(hip, orc, os.path.join(d, "sim_run.ini"), os.path.join(d, "expected"))
        _same_files(d, os.path.join(d, "expected"))


@pytest.mark.parametrize("n_spots,layout", [(1, problems.LAYOUT_REF), (1, problems.LAYOUT_CFG),
                                            (8, problems.LAYOUT_CFG)])
def test_context_memory_equals_plan(hip, n_spots, layout):
    prob = problems.reference_grid_case(Nx=64, Nz=96, layout=layout)
    tab = prob.tables(hip)
    grid = prob.make_grid(len(tab.E_rep))
    plan = io.memory_estimate(hip, grid, n_spots)
    eng = Engine(hip, grid, tab, n_spots=n_spots)
    import ctypes as C
    b = C.c_uint64(0)
    hip.call("ctx_memory", eng.h, C.byref(b))
    assert b.value == plan.total_bytes
    pre = io.memory_preflight(hip, grid, n_spots, 0.85)
    assert pre.total_bytes == plan.total_bytes and pre.free_bytes > 0
    assert pre.ok == (pre.total_bytes <= pre.usable_bytes) == 1
    eng.close()
