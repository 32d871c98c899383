"""Run configuration, output files and memory planning of the library (include/sm2d_io.h and the
planning entry points of include/sm2d_hip.h).  CPU only.

Pins:
  INI loader + validation   the reference's own loader (config_loader.hpp, header-only) compiled
                            into oracle/_ref/libsm2d_ref.so, on the reference's INI files
                            (tests/golden/ini, copied from the checkout root and validation/)
                            and on edge-case files written here
  dose writers              the reference's outputs results/pdd.txt and the first three rows of
                            results/dose_2d.txt (tests/golden/output); depth dose and Bragg
                            peak against deterministic_beam.cpp compiled into _ref
  raw ψ CSV                 the reference's one-step dumps (tests/golden/one_step_case_*):
                            the initial dump byte for byte from the oracle's injection, the
                            after-K4 dump column by column from its own rows
  per-cell ψ summary CSV    the reference's debug_cells_iter_NN_after_K4.csv of its 150 MeV run
                            (iterations 1..10) and of the one-step cases, rewritten from the raw
                            dumps of the same iterations (tests/golden/multi_step)
  dense memory estimate     the reference's preflight line in output_message.txt:14-22
                            (100 x 320 cells, 36 x 1029 bins: required 5.96 GiB)
"""
import ctypes as C
import glob
import os

import numpy as np
import pytest

import fixtures
from sm_2d_amd import abi, io, runtime

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
REF = os.path.join(ROOT, "oracle", "_ref", "libsm2d_ref.so")
INIS = sorted(glob.glob(os.path.join(GOLD, "ini", "*.ini")) +
              glob.glob(os.path.join(GOLD, "one_step_case_*", "*.ini")))


@pytest.fixture(scope="module")
def lib():
    return runtime.load_library()


@pytest.fixture(scope="module")
def ref():
    if not os.path.exists(REF):
        pytest.skip("oracle/_ref not built (no reference checkout)")
    r = C.CDLL(REF)
    r.ref_config_load.argtypes = [C.c_char_p, C.POINTER(io.RunConfig), C.c_char_p, C.c_int]
    r.ref_config_load.restype = C.c_int
    r.ref_depth_dose.argtypes = [abi.dp, C.c_int, C.c_int, abi.dp, C.POINTER(C.c_int)]
    r.ref_depth_dose.restype = C.c_int
    return r


def _ref_load(ref, path):
    c = io.RunConfig()
    msg = C.create_string_buffer(512)
    rc = ref.ref_config_load(str(path).encode(), C.byref(c), msg, 512)
    return rc, c, msg.value.decode()


def _ours(lib, path):
    from sm_2d_amd.engine import LibError
    c = io.config_load(lib, path)
    try:
        io.config_validate(lib, c)
        return 0, c, ""
    except LibError as e:
        return 1, c, str(e).split(": ", 1)[1]


def _same(a: io.RunConfig, b: io.RunConfig):
    da, db = a.as_dict(), b.as_dict()
    for k in da:
        va, vb = da[k], db[k]
        if isinstance(va, float):
            assert np.float32(va).tobytes() == np.float32(vb).tobytes(), (k, va, vb)
        else:
            assert va == vb, (k, va, vb)


@pytest.mark.parametrize("path", INIS, ids=[os.path.relpath(p, GOLD) for p in INIS])
def test_config_loader_matches_reference_on_its_ini_files(lib, ref, path):
    rc_r, cr, mr = _ref_load(ref, path)
    rc_o, co, mo = _ours(lib, path)
    assert rc_r != 2
    assert (rc_o, mo) == (rc_r, mr)
    _same(co, cr)


EDGE_CASES = {
    "aliases_and_comments": """
        # comment
        ; other comment
        [ transport ]
        E_fine_on_MeV = 12.5
        E_fine_on = 13.0        # the unit-less key wins, the trailing text is ignored by stof
        E_fine_off_MeV = 14
        step_coarse_mm = 2.5
        E_coarse_max = 250
        debug_dumps = off
        enable_debug_dumps = YES
        energy_groups = 0.1:2:0.1, 2:250:0.5
        [output]
        output_dir = "quoted dir"
        dose_2d_file = 'd.txt'
        normalize_dose = maybe
        save_pdd = 0
    """,
    "legacy_trigger_only": """
        [transport]
        E_trigger_MeV = 15
    """,
    "legacy_trigger_with_fine_on": """
        [transport]
        E_trigger = 15
        E_fine_on = 9
        E_fine_off = 9.5
    """,
    "bad_numbers_keep_defaults": """
        [grid]
        Nx = abc
        Nz = 57.9
        dx_mm = 1e3x
        dz_mm = -
        max_steps = 99999999999
        [sampling]
        random_seed = 4294967295
        n_samples = 12
        use_stratified = True
    """,
    "malformed_energy_groups": """
        [transport]
        energy_groups = 0.1:2, 2:250:0.5
    """,
    "energy_groups_extra_fields": """
        [transport]
        energy_groups = 0.1:2:0.1:9, ,2:250:1
    """,
    "invalid_fine_off": """
        [transport]
        E_fine_on = 20
        E_fine_off = 15
    """,
    "invalid_overlapping_groups": """
        [transport]
        energy_groups = 0.1:20:0.1, 10:250:0.5
    """,
    "invalid_margin": """
        [transport]
        preflight_vram_margin = 1.5
    """,
    "particle_and_profile": """
        [particle]
        type = carbon_ion
        [beam]
        profile = flat
        weight = 0
    """,
    "no_sections": """
        Nx = 5
        key without equals
        = value without key
        []
    """,
}


@pytest.mark.parametrize("name", sorted(EDGE_CASES))
def test_config_loader_matches_reference_on_edge_cases(lib, ref, tmp_path, name):
    p = tmp_path / f"{name}.ini"
    p.write_text("\n".join(l.strip() for l in EDGE_CASES[name].splitlines()) + "\n")
    rc_r, cr, mr = _ref_load(ref, p)
    rc_o, co, mo = _ours(lib, p)
    assert (rc_o, mo) == (rc_r, mr)
    _same(co, cr)


def test_config_load_missing_file(lib, tmp_path):
    from sm_2d_amd.engine import LibError
    with pytest.raises(LibError, match="Failed to load configuration file"):
        io.config_load(lib, tmp_path / "none.ini")


def test_runtime_derivation(lib, tmp_path):
    """gpu_transport_runner.cpp:29-98,172-173 and gpu_transport_wrapper.cu:273."""
    cfg = io.config_load(lib, os.path.join(GOLD, "one_step_case_C", "case_c.ini"))
    g, tp, sp = io.config_runtime(lib, cfg)
    grid, tpf, spf = fixtures.problem("C")
    for k, v in grid.items():
        assert np.float32(getattr(g, k)) == np.float32(v), k
    assert g.N_E == 1029
    assert (tp.E_fine_on, tp.E_fine_off, tp.max_iterations) == (tpf.E_fine_on, tpf.E_fine_off,
                                                               tpf.max_iterations)
    assert tp.sigma_x_initial == sp.sigma_x == spf.sigma_x
    for k in ("x0", "z0", "theta0", "E0", "W_total", "sigma_theta", "sigma_E", "n_samples",
              "random_seed"):
        assert getattr(sp, k) == getattr(spf, k), k
    # legacy lift: E_trigger customised while the fine thresholds are at their defaults
    p = tmp_path / "t.ini"
    p.write_text("[grid]\nmax_steps = 77\n")
    c = io.config_load(lib, p)
    c.E_trigger = 20.0
    _, tp2, _ = io.config_runtime(lib, c)
    assert (tp2.E_fine_on, tp2.E_fine_off, tp2.max_iterations) == (20.0, 20.0, 77)
    c.validation_mode = 1
    assert io.config_runtime(lib, c)[1].fail_fast_on_audit == 1
    c.N_theta_local, c.N_E_local = 8, 4
    g3, _, _ = io.config_runtime(lib, c)
    assert (g3.n_x_sub, g3.n_z_sub) == (4, 1)
    c.N_theta_local = 5
    from sm_2d_amd.engine import LibError
    with pytest.raises(LibError, match="layout"):
        io.config_runtime(lib, c)


# ---------------------------------------------------------------------------------------------
# dose writers
# ---------------------------------------------------------------------------------------------

def _table(path):
    head, rows = [], []
    for line in open(path).read().split("\n"):
        if line.startswith("#"):
            head.append(line)
        elif line:
            rows.append(line.split("\t"))
    return head, rows


def test_pdd_writer_reproduces_reference_file(lib, tmp_path):
    """results/pdd.txt (dz = 1 mm): header, depth and dose columns byte for byte from its own
    4-decimal doses; the normalised column and the peak dose's 6 significant digits come from
    the unrounded doses, so they agree to the file's last digit."""
    head, rows = _table(os.path.join(GOLD, "output", "pdd.txt"))
    dose = np.array([float(r[1]) for r in rows])
    out = tmp_path / "pdd.txt"
    io.write_pdd(lib, out, dose[:, None], dz=1.0, normalize=True)
    h2, r2 = _table(out)
    assert h2[:2] == head[:2]
    assert h2[2].startswith("# Bragg Peak at: 161 mm, ") and head[2].startswith("# Bragg Peak at: 161 mm, ")
    assert float(h2[2].split(", ")[1].split()[0]) == pytest.approx(float(head[2].split(", ")[1].split()[0]), abs=6e-5)
    assert [r[:2] for r in r2] == [r[:2] for r in rows]
    assert np.abs(np.array([float(r[2]) for r in r2]) - np.array([float(r[2]) for r in rows])).max() <= 1.0001e-4
    assert open(out).read().endswith("\n") and not open(out).read().endswith("\n\n")


def test_dose_2d_writer_reproduces_reference_rows(lib, tmp_path):
    """First three depth rows of results/dose_2d.txt (100 x 320 at 1 mm): x, z and dose columns
    and the blank line after every depth row byte for byte; dose_norm within the rounding of
    the file's 4-decimal doses, with the file's maximum dose (recovered from its own
    dose / dose_norm pairs) placed in a fourth row."""
    src = os.path.join(GOLD, "output", "dose_2d_rows0-2.txt")
    text = open(src).read()
    head, rows = _table(src)
    dose = np.array([float(r[2]) for r in rows]).reshape(3, 100)
    norm = np.array([float(r[3]) for r in rows]).reshape(3, 100)
    sel = norm > 0.01
    dmax = float((dose[sel] ** 2).sum() / (dose[sel] * norm[sel]).sum())  # least squares
    edep = np.zeros((4, 100))
    edep[:3] = dose
    edep[3, 0] = dmax
    out = tmp_path / "d.txt"
    io.write_dose_2d(lib, out, edep, dx=1.0, dz=1.0, normalize=True)
    t2 = open(out).read()
    blocks_ref = text.split("\n\n")[:3]
    blocks_out = t2.split("\n\n")[:3]
    for a, b in zip(blocks_ref, blocks_out):
        la, lb = a.split("\n"), b.split("\n")
        assert len(la) == len(lb)
        for x, y in zip(la, lb):
            if x.startswith("#"):
                assert x == y
                continue
            fx, fy = x.split("\t"), y.split("\t")
            assert fx[:3] == fy[:3]
            # 4-decimal doses (+-5e-5 over a maximum of ~0.4) and the rounding of both files
            assert abs(float(fx[3]) - float(fy[3])) <= 2.0001e-4


def test_writers_replace_nan_and_skip_normalisation(lib, tmp_path):
    e = np.array([[np.nan, 1.0], [np.inf, 3.0]])
    io.write_dose_2d(lib, tmp_path / "d.txt", e, dx=0.5, dz=0.5, normalize=False)
    assert open(tmp_path / "d.txt").read() == (
        "# 2D Dose Distribution\n# x(mm) z(mm) dose(Gy)\n"
        "-0.2500\t0.0000\t0.0000\n0.2500\t0.0000\t1.0000\n\n"
        "-0.2500\t0.5000\t0.0000\n0.2500\t0.5000\t3.0000\n\n")
    z = np.zeros((3, 2))
    io.write_pdd(lib, tmp_path / "p.txt", z, dz=2.0, normalize=True)
    assert open(tmp_path / "p.txt").read() == (
        "# Depth-Dose Distribution (PDD)\n# Depth(mm) Dose(Gy) Dose_norm\n"
        "# Bragg Peak at: 0 mm, 0 Gy\n0.0000\t0.0000\n2.0000\t0.0000\n4.0000\t0.0000\n")


def test_pdd_peak_and_depth_dose_match_reference(lib, ref, tmp_path):
    """get_depth_dose / find_bragg_peak_z of the reference (first maximum on ties)."""
    rng = np.random.default_rng(4)
    for trial in range(20):
        Nz, Nx = int(rng.integers(1, 40)), int(rng.integers(1, 30))
        e = np.round(rng.random((Nz, Nx)) * 3, 1 if trial % 2 else 6)
        if trial % 3 == 0:
            e[Nz // 2] = e[0]  # a tie
        dd = np.zeros(Nz)
        pk = C.c_int(0)
        ref.ref_depth_dose(np.ascontiguousarray(e).ctypes.data_as(abi.dp), Nx, Nz,
                           dd.ctypes.data_as(abi.dp), C.byref(pk))
        io.write_pdd(lib, tmp_path / "p.txt", e, dz=0.25, normalize=True)
        lines = open(tmp_path / "p.txt").read().split("\n")
        assert lines[2] == "# Bragg Peak at: %s mm, %s Gy" % ("%g" % np.float32(pk.value * np.float32(0.25)),
                                                               "%g" % dd[pk.value])
        assert [l.split("\t")[1] for l in lines[3:3 + Nz]] == ["%.4f" % v for v in dd]


# ---------------------------------------------------------------------------------------------
# raw phase-space CSV
# ---------------------------------------------------------------------------------------------

def _case_tables(lib, case):
    grid, tp, sp = fixtures.problem(case)
    g = abi.Grid(**grid)
    tab = lib.make_tables(g.N_theta, g.theta_min, g.theta_max)
    g.N_E = len(tab.E_rep)
    return g, tab


@pytest.mark.parametrize("case", ["A", "B", "C", "D"])
def test_initial_psi_dump_byte_for_byte(lib, orc, case, tmp_path):
    """The oracle's injection written in the initial-dump format equals the reference's
    debug_ps_raw_iter_00_initial.csv of the case byte for byte."""
    from oracle.pyoracle import OracleEngine
    grid, tp, sp = fixtures.problem(case)
    g = abi.Grid(**grid)
    tab = orc.make_tables(g.N_theta, g.theta_min, g.theta_max)
    g.N_E = len(tab.E_rep)
    e = OracleEngine(orc, g, tab)
    e.inject(sp)
    _, rows = e.psi_dump()
    out = tmp_path / "init.csv"
    io.write_psi_csv(lib, out, rows, g, tab.theta_edges, tab.E_edges, iteration=0, decimals=6)
    gold = os.path.join(fixtures.case_dir(case), "debug_ps_raw_iter_00_initial.csv")
    assert open(out).read() == open(gold).read()


@pytest.mark.parametrize("case", ["A", "B", "C", "D"])
def test_after_k4_dump_format(lib, case, tmp_path):
    """The reader returns the rows of debug_ps_raw_iter_01_after_K4.csv, and writing them back
    reproduces every decoded column byte for byte and every weight to float32."""
    g, tab = _case_tables(lib, case)
    gold = os.path.join(fixtures.case_dir(case), "debug_ps_raw_iter_01_after_K4.csv")
    rows = io.read_psi_csv(lib, gold)
    f = fixtures.read_psi(case, "after")
    for k in ("cell", "slot", "bid", "lidx"):
        assert np.array_equal(rows[k], f[k].astype(rows[k].dtype)), k
    assert np.array_equal(rows["w"], f["w"].astype(np.float32))
    out = tmp_path / "k4.csv"
    io.write_psi_csv(lib, out, rows, g, tab.theta_edges, tab.E_edges, iteration=1, decimals=8)
    a = open(gold).read().splitlines()
    b = open(out).read().splitlines()
    # weights in (1e-12, 5e-9) print as 0.00000000, read back as 0 and are skipped on writing
    a = a[:1] + [l for l in a[1:] if np.float32(l.split(",")[7]) > 1e-12]
    assert a[0] == b[0] and len(a) == len(b)
    for x, y in zip(a[1:], b[1:]):
        fx, fy = x.split(","), y.split(",")
        assert fx[:7] == fy[:7] and fx[8:] == fy[8:]
        assert np.float32(fx[7]) == np.float32(fy[7])


def _printed_rows(lib, path):
    """Rows of a raw dump; a weight printed as 0.00000000 was above the dump's 1e-12 threshold,
    so it is read back as a tiny positive weight that counts as a bin and adds nothing visible."""
    rows = io.read_psi_csv(lib, path)
    rows["w"] = np.where(rows["w"] > 0, rows["w"], np.float32(2e-12)).astype(np.float32)
    return rows


def _same_cells(mine_path, gold_path):
    a = open(gold_path).read().splitlines()
    b = open(mine_path).read().splitlines()
    assert a[0] == b[0] and len(a) == len(b), (len(a), len(b))
    for x, y in zip(a[1:], b[1:]):
        fx, fy = x.split(","), y.split(",")
        assert fx[:5] == fy[:5] and fx[6] == fy[6], (x, y)      # cell, ix, iz, x_mm, z_mm, bins
        assert fx[9:] == fy[9:], (x, y)                        # min/max E and theta: bin centres
        # total weight and the weighted means come from weights printed with 8 decimals
        assert abs(float(fx[5]) - float(fy[5])) <= 1.5e-6, (x, y)
        assert abs(float(fx[7]) - float(fy[7])) <= 2e-6 * abs(float(fx[7])) + 1.5e-6, (x, y)
        assert abs(float(fx[8]) - float(fy[8])) <= 1.5e-6, (x, y)


@pytest.mark.parametrize("n", range(1, 11))
def test_cells_dump_matches_reference_multi_step(lib, n, tmp_path):
    """debug_cells_iter_NN_after_K4.csv of the reference's 150 MeV run, rewritten from the raw
    dump of the same iteration: cells, positions and bin counts exactly, energy and angle ranges
    exactly, sums and means within the raw dump's print precision."""
    from sm_2d_amd import problems
    g = abi.Grid(**problems._grid(100, 200, 1.0, 1.0, 36, -0.35, 0.35, problems.LAYOUT_REF))
    tab = lib.make_tables(g.N_theta, g.theta_min, g.theta_max)
    g.N_E = len(tab.E_rep)
    rows = _printed_rows(lib, os.path.join(GOLD, "multi_step", f"iter_{n:02d}.csv"))
    io.write_cells_csv(lib, tmp_path / "c.csv", rows, g, tab.theta_edges, tab.E_edges)
    _same_cells(tmp_path / "c.csv", os.path.join(GOLD, "multi_step", f"cells_{n:02d}.csv"))


@pytest.mark.parametrize("case", ["A", "B", "C", "D"])
def test_cells_dump_matches_reference_cases(lib, case, tmp_path):
    g, tab = _case_tables(lib, case)
    rows = _printed_rows(lib, os.path.join(fixtures.case_dir(case), "debug_ps_raw_iter_01_after_K4.csv"))
    io.write_cells_csv(lib, tmp_path / "c.csv", rows, g, tab.theta_edges, tab.E_edges)
    _same_cells(tmp_path / "c.csv", os.path.join(GOLD, "multi_step", f"cells_case_{case}.csv"))


def test_psi_csv_skips_tiny_weights_and_sorts(lib, tmp_path):
    g, tab = _case_tables(lib, "A")
    rows = {"cell": np.array([41, 40, 40]), "slot": np.array([0, 1, 0]),
            "bid": np.array([5, 7, 7], np.uint32), "lidx": np.array([3, 2, 1]),
            "w": np.array([0.5, 1e-13, 0.25], np.float32)}
    io.write_psi_csv(lib, tmp_path / "p.csv", rows, g, tab.theta_edges, tab.E_edges)
    back = io.read_psi_csv(lib, tmp_path / "p.csv")
    assert back["cell"].tolist() == [40, 41] and back["w"].tolist() == [0.25, 0.5]


# ---------------------------------------------------------------------------------------------
# memory planning
# ---------------------------------------------------------------------------------------------

def test_dense_estimate_matches_reference_preflight_line(lib):
    """output_message.txt:14-22: a 100 x 320 grid with 36 x 1029 bins at LOCAL_BINS 256 needs
    "required 5.96 GiB" (psi 1.96 GiB, buckets 3.93 GiB, pipeline 3.48 MiB, lut 24.12 KiB,
    overhead 64 MiB)."""
    g = abi.Grid(Nx=100, Nz=320, dx=1.0, dz=1.0, x_min=-50.0, z_min=0.0, N_theta=36,
                 theta_min=-0.35, theta_max=0.35, N_E=1029, n_theta_local=4, n_e_local=2,
                 n_x_sub=8, n_z_sub=4)
    p = io.memory_estimate(lib, g, 1)
    assert "%.2f GiB" % (p.dense_bytes / 2 ** 30) == "5.96 GiB"
    parts = 1.96 * 2 ** 30 + 3.93 * 2 ** 30 + 3.48 * 2 ** 20 + 24.12 * 2 ** 10 + 64 * 2 ** 20
    assert abs(p.dense_bytes - parts) < 0.01 * 2 ** 30
    # the sparse store of the same grid is a small fraction of it
    assert p.total_bytes < 0.25 * p.dense_bytes
    assert p.total_bytes == p.psi_bytes + p.bucket_bytes + p.cell_bytes + p.other_bytes
    assert p.ok == 1


def test_memory_estimate_scales_and_rejects_bad_grids(lib):
    from sm_2d_amd.engine import LibError
    g = abi.Grid(Nx=200, Nz=400, dx=0.5, dz=0.5, x_min=-50.0, z_min=0.0, N_theta=512,
                 theta_min=-1.5707964, theta_max=1.5707964, N_E=256, n_theta_local=8,
                 n_e_local=4, n_x_sub=4, n_z_sub=1)
    one = io.memory_estimate(lib, g, 1)
    many = io.memory_estimate(lib, g, 256)
    assert one.listed_cells == 48 * 200 + 1024 and many.listed_cells == 256 * 24 * 200 + 1024
    # a batch plans half the listed rows and payload slots per spot of a single-spot context
    assert 50 * one.total_bytes < many.total_bytes < 256 * one.total_bytes
    g.n_x_sub = 3
    with pytest.raises(LibError, match="layout"):
        io.memory_estimate(lib, g, 1)
