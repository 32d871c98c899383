"""Known-answer and closed-form checks of the oracle's stages and physics.

K1 truth table   tests/kernels/test_k1_activemask.cpp:7-144 (reference KATs)
K5 closure KAT   tests/gpu/test_k5_drop_channels_gpu.cu:120-211 (reference KAT)
Physics          closed forms written from the published formulas (Highland/PDG, ICRU-63
                 nuclear slope, Bohr/Landau straggling limits, CSDA range inversion) and the
                 NIST PSTAR nodes (src/data/nist/pstar_water.txt)
"""
import math

import numpy as np
import pytest

from sm_2d_amd import abi, problems


def _k1_engine(orc, Nx=4, Nz=4):
    from oracle.pyoracle import OracleEngine
    # log grid 0.1-300 MeV / 256 bins, N_E_local = 4 as in the reference's K1 tests; 96 angular
    # bins so that the tests' angular block 10 lies on the grid
    tab = orc.make_tables(96, -0.35, 0.35, ("log", 0.1, 300.0, 256))
    grid = abi.Grid(Nx=Nx, Nz=Nz, dx=1.0, dz=1.0, x_min=-2.0, z_min=0.0, N_theta=96,
                    theta_min=-0.35, theta_max=0.35, N_E=256, n_theta_local=8, n_e_local=4,
                    n_x_sub=4, n_z_sub=1)
    return OracleEngine(orc, grid, tab), tab


def _thr_energy(tab, b_E, n_e_local=4):
    """An energy whose bin lies in block b_E (so compute_b_E_threshold returns b_E)."""
    return float(tab.E_rep[b_E * n_e_local])


def _row(cell, b_theta, b_E, lidx, w):
    return dict(cell=[cell], slot=[0], bid=[(b_theta & 0xFFF) | ((b_E & 0xFFF) << 12)], lidx=[lidx], w=[w])


def _merge(*rows):
    out = {k: [] for k in ("cell", "slot", "bid", "lidx", "w")}
    for r in rows:
        for k in out:
            out[k] += r[k]
    return out


CASES = [  # (b_E of the block, on, off, w, prev, expected) — test_k1_activemask.cpp
    ("LowEnergyTriggersFineTransport", 5, 10, 10, 1.0, 0, 1),
    ("ThresholdEnergyTriggersFineTransport", 10, 10, 10, 1.0, 0, 1),
    ("HighEnergyDoesNotTrigger", 15, 10, 10, 1.0, 0, 0),
    ("WeightThreshold", 5, 10, 10, 1e-12, 0, 0),
    ("HysteresisKeepsActiveUntilFineOff", 11, 10, 12, 1.0, 1, 1),
    ("HysteresisDeactivatesAboveFineOff", 13, 10, 12, 1.0, 1, 0),
]


@pytest.mark.parametrize("name,bE,on,off,w,prev,exp", CASES, ids=[c[0] for c in CASES])
def test_k1_truth_table(orc, name, bE, on, off, w, prev, exp):
    e, tab = _k1_engine(orc)
    e.psi_load(_row(0, 10, bE, 0, w))
    tp = abi.Transport(E_fine_on=_thr_energy(tab, on), E_fine_off=_thr_energy(tab, off),
                       weight_active_min=1e-10)
    mp = np.zeros(16, np.uint8)
    mp[0] = prev
    mask, act, crs = e.mask_compact(tp, mp)
    assert mask[0] == exp
    if exp:
        assert list(act) == [0] and list(crs) == []
    else:
        assert list(act) == [] and (list(crs) == [0] if w > 0 else True)


def test_k1_multiple_cells(orc):
    e, tab = _k1_engine(orc)
    e.psi_load(_merge(_row(0, 10, 5, 0, 1.0), _row(1, 10, 15, 0, 1.0)))
    tp = abi.Transport(E_fine_on=_thr_energy(tab, 10), E_fine_off=_thr_energy(tab, 10),
                       weight_active_min=1e-10)
    mask, act, crs = e.mask_compact(tp)
    assert mask[0] == 1 and mask[1] == 0
    assert list(act) == [0] and list(crs) == [1]


def test_compute_b_E_trigger_ranges(orc):
    """ComputeBETriggerFromEnergy: EnergyGrid(0.1, 300, 256), N_E_local 4 -> block of 10 MeV in
    (30, 45), block of 50 MeV in (45, 55)."""
    e, rep = orc.energy_grid_log(0.1, 300.0, 256)
    b10 = (np.searchsorted(e, 10.0, side="right") - 1) // 4
    b50 = (np.searchsorted(e, 50.0, side="right") - 1) // 4
    assert 30 < b10 < 45 and 45 < b50 < 55 and b50 > b10


def test_psi_load_rejects_bins_off_the_grid(orc):
    """A block may reach past the grid (a partial last block); its local bins there have no
    energy centre, so loading one is an error while the last bins on the grid load."""
    from sm_2d_amd.engine import LibError
    from oracle.pyoracle import OracleEngine
    for n_e in (256, 255):  # 64 whole blocks of 4 / a last block with 3 bins on the grid
        tab = orc.make_tables(96, -0.35, 0.35, ("log", 0.1, 300.0, n_e))
        grid = abi.Grid(Nx=2, Nz=2, dx=1.0, dz=1.0, x_min=-1.0, z_min=0.0, N_theta=96,
                        theta_min=-0.35, theta_max=0.35, N_E=n_e, n_theta_local=8, n_e_local=4,
                        n_x_sub=4, n_z_sub=1)
        e = OracleEngine(orc, grid, tab)
        last_e = (n_e - 1) % 4  # local energy index of the last bin on the grid
        e.psi_load(_row(0, 11, 63, 7 + 8 * last_e, 1.0))
        with pytest.raises(LibError, match="outside the grid"):
            e.psi_load(_row(0, 11, 64, 0, 1.0))
        with pytest.raises(LibError, match="outside the grid"):
            e.psi_load(_row(0, 12, 0, 0, 1.0))
        if last_e < 3:
            with pytest.raises(LibError, match="outside the grid"):
                e.psi_load(_row(0, 11, 63, 8 * (last_e + 1), 1.0))


def test_k5_drop_channel_kat(orc):
    """1x1 grid, psi_in weight 1.0 at E = 100 MeV (one 1-bin energy grid), empty psi_out; source
    losses 0.2/0.3 (W) and 20/30 (E), transport drop 1.0 / 100: closes exactly; without the
    drop it fails both checks (test_k5_drop_channels_gpu.cu:120-211)."""
    from oracle.pyoracle import OracleEngine
    from sm_2d_amd.engine import Tables
    th = np.array([-0.35, 0.35], np.float32)
    ee = np.array([99.5, 100.5], np.float32)
    one = np.ones(1, np.float32)
    tab = Tables(th, ee, np.array([100.0], np.float32), one, one, one, one, one)
    grid = abi.Grid(Nx=1, Nz=1, dx=1.0, dz=1.0, x_min=0.0, z_min=0.0, N_theta=1, theta_min=-0.35,
                    theta_max=0.35, N_E=1, n_theta_local=8, n_e_local=4, n_x_sub=4, n_z_sub=1)
    e = OracleEngine(orc, grid, tab)
    e.psi_load(_row(0, 0, 0, 0, 1.0))
    zeros = np.zeros((7, 1))
    r = e.k5_audit(zeros, 0.2, 0.3, 20.0, 30.0, 1.0, 100.0, include_source_terms=True)
    assert r.W_pass == 1 and r.E_pass == 1
    assert r.W_in == pytest.approx(1.5) and r.E_in == pytest.approx(150.0)
    assert r.W_transport_drop == pytest.approx(1.0) and r.E_transport_drop == pytest.approx(100.0)
    r = e.k5_audit(zeros, 0.2, 0.3, 20.0, 30.0, 0.0, 0.0, include_source_terms=True)
    assert r.W_pass == 0 and r.E_pass == 0
    assert r.W_error > 1e-6 and r.E_error > 1e-5


# ---------------------------------------------------------------------------------------------
# Physics closed forms
# ---------------------------------------------------------------------------------------------
MP = 938.272


def _beta_p(E):
    g = (E + MP) / MP
    beta = math.sqrt(1 - 1 / g ** 2)
    p = math.sqrt((E + MP) ** 2 - MP ** 2)
    return beta, p


@pytest.fixture(scope="module")
def phys(orc):
    from oracle.pyoracle import OracleEngine
    tab = orc.make_tables(36, -0.35, 0.35)
    grid = abi.Grid(Nx=4, Nz=4, dx=0.5, dz=0.5, x_min=-1.0, z_min=0.0, N_theta=36,
                    theta_min=-0.35, theta_max=0.35, N_E=len(tab.E_rep), n_theta_local=4,
                    n_e_local=2, n_x_sub=8, n_z_sub=4)
    return OracleEngine(orc, grid, tab), tab


def test_highland_closed_form(phys):
    e, _ = phys
    E = np.array([1.0, 10.0, 70.0, 150.0, 230.0], np.float32)
    ds = np.array([0.05, 0.5, 1.0, 5.0, 20.0], np.float32)
    for d in ds:
        got = e.eval(6, E, np.full_like(E, d))
        for Ei, g in zip(E, got):
            beta, p = _beta_p(float(Ei))
            t = float(d) / 360.8
            exp = 13.6 / (beta * p) * math.sqrt(t) * max(1 + 0.038 * math.log(t), 0.25)
            assert g == pytest.approx(exp, rel=1e-4)  # FP32 1 - 1/gamma^2 cancels at low E
    assert e.eval(6, np.array([100.0], np.float32), np.array([1e-4], np.float32))[0] == 0.0


def test_nuclear_cross_section_closed_form(phys):
    e, _ = phys
    E = np.array([1.0, 4.99, 5.0, 12.5, 19.99, 20.0, 100.0, 250.0, 1e4], np.float32)
    got = e.eval(8, E)
    for Ei, g in zip(E.astype(float), got):
        if Ei < 5:
            exp = 0.0
        elif Ei < 20:
            exp = 0.0016 * (Ei - 5) / 15
        else:
            exp = 0.0012 * max(1 - 0.15 * math.log(Ei / 100), 0.4)
        assert g == pytest.approx(exp, rel=2e-6, abs=1e-12)


def test_straggling_regimes(phys):
    """Bohr limit (kappa > 10): 0.156 sqrt(ds)/beta; Landau limit (kappa < 0.01): 4 xi / 2.355
    with xi = 0.1535 * 0.555 / beta^2 * ds_cm; in between the 1/(1+kappa) blend."""
    e, _ = phys
    E = np.array([0.2, 1.0, 10.0, 150.0], np.float32)
    for d in (0.01, 0.5, 5.0):
        got = e.eval(7, E, np.full_like(E, d))
        for Ei, g in zip(E.astype(float), got):
            beta, _ = _beta_p(Ei)
            gam = (Ei + MP) / MP
            xi = 0.307 / 2 * 0.555 / beta ** 2 * d / 10
            mr = 0.511 / MP
            tmax = 2 * 0.511 * beta ** 2 * gam ** 2 / (1 + 2 * gam * mr + mr * mr)
            k = max(xi / tmax, 1e-6)
            bohr = 0.156 * math.sqrt(d) / math.sqrt(max(beta ** 2, 0.01))
            land = 4 * xi / 2.355
            exp = bohr if k > 10 else (land if k < 0.01 else (land / (1 + k) + (1 - 1 / (1 + k)) * bohr))
            assert g == pytest.approx(exp, rel=1e-5)


def test_range_table_and_inversion(phys):
    """R(E) from the log-log NIST interpolation reproduces the PSTAR CSDA ranges; R(70) = 40.8 mm
    and R(150) = 157.7 mm (test_r_lut.cpp tolerances are 15 %; here 1 %); E(R(E) - s) is the
    CSDA energy after a step and E(R(E)) returns E to 0.5 %."""
    e, tab = phys
    R = e.eval(0, np.array([70.0, 150.0], np.float32))
    assert R[0] == pytest.approx(40.80, rel=0.01) and R[1] == pytest.approx(157.7, rel=0.01)
    E = np.linspace(2.0, 240.0, 200).astype(np.float32)
    back = e.eval(2, e.eval(0, E))
    assert np.allclose(back, E, rtol=5e-3)
    # stepping 1 mm at 150 MeV removes ~ S(150) * 0.1 = 0.544 MeV (stopping power 5.443 MeV cm2/g)
    E1 = e.eval(5, np.array([150.0], np.float32), np.array([1.0], np.float32))[0]
    assert 150.0 - E1 == pytest.approx(0.5443, rel=0.03)


def test_device_step_rule(phys):
    e, _ = phys
    E = np.array([0.3, 1.0, 3.0, 10.0, 50.0, 150.0, 230.0], np.float32)
    got = e.eval(3, E)
    R = e.eval(0, E)
    for Ei, Ri, g in zip(E.astype(float), R.astype(float), got):
        fl = 0.1 if Ei < 2 else (0.25 if Ei < 20 else (0.5 if Ei < 100 else 1.0))
        d = max(0.01 * Ri, fl)
        if Ei < 0.5:
            d = min(d, 0.05)
        elif Ei < 5:
            d = min(d, 0.2)
        elif Ei < 20:
            d = min(d, 0.5)
        assert g == pytest.approx(d, rel=1e-6)


def _min_E_block_and_fine(orc, prob, n_iter):
    from oracle.pyoracle import OracleEngine
    e = problems.setup_engine(orc, prob, OracleEngine)
    tp = prob.make_transport()
    e.inject(prob.make_spot())
    _, r0 = e.psi_dump()
    fine = 0
    for _ in range(n_iter // 10):
        _, act, _ = e.mask_compact(tp)
        fine += len(act)
        e.run_iterations(tp, 10)
    _, r = e.psi_dump()
    return int(((r0["bid"] >> 12) & 0xFFF).min()), int(((r["bid"] >> 12) & 0xFFF).min()), fine


def test_coarse_steps_lock_to_wide_energy_bins(orc):
    """K2 takes a component's energy at its bin centre and re-bins E - dE
    (k2_coarsetransport.cu:167-170, 235, 488).  On the 256-bin log grid (bins 3.1 % wide) a
    0.5 mm step loses less than half a bin, so a 70 MeV component stays in its injected energy
    block and never reaches the fine phase (E <= 10 MeV); on the reference runner's piecewise
    1029-bin grid the same beam slows down and the fine list fills."""
    b0, b1, fine = _min_E_block_and_fine(orc, problems.config1(n_samples=100), 80)
    assert b1 == b0 and fine == 0
    b0, b1, fine = _min_E_block_and_fine(
        orc, problems.reference_grid_case(Nx=40, Nz=100, E0=70.0, sigma_x=2.0, n_samples=100), 100)
    assert b1 < b0 and fine > 0
