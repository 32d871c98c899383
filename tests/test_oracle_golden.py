"""The oracle against the reference's committed one-step dumps (tests/golden/one_step_case_*).

What the dumps pin (all four cases, INI inputs stored beside them):
  - source injection: every (cell, slot, block id, local bin) row and its weight;
  - one K1..K6 iteration: the weight in every (cell, E bin, x_sub, z_sub) after K4, and the
    channel totals (EdepC, nuclear weight and energy, boundary, source accounting, K5 residual).
What they do not pin: the split of each scattered component over neighbouring theta bins.  The
dumps come from a build whose angular sampling differs from the current reference source (the
pencil components of cases A and B show exactly zero deflection, which the current hash-seeded
draw cannot produce for them), so theta-bin weights are "parity unpinned" by these fixtures.
"""
import numpy as np
import pytest

from sm_2d_amd import abi

import fixtures as fx


def _engine(orc, case):
    from oracle.pyoracle import OracleEngine
    g, tp, spot = fx.problem(case)
    tab = orc.make_tables(g["N_theta"], g["theta_min"], g["theta_max"])
    grid = abi.Grid(N_E=len(tab.E_rep), **g)
    e = OracleEngine(orc, grid, tab)
    A, B, _ = orc.fe_precompute(spot.E0, spot.sigma_x, grid.Nz, grid.dz)
    e.upload_fe(A, B)
    e.reset()
    return e, tp, spot


@pytest.mark.parametrize("case", ["A", "B", "C", "D"])
def test_injection_matches_dump(orc, case):
    e, tp, spot = _engine(orc, case)
    a = e.inject(spot)
    _, rows = e.psi_dump()
    mine = fx.rows_as_set(rows, 6)
    ref = fx.rows_as_set(fx.read_psi(case, "initial"), 6)
    assert set(mine) == set(ref)
    for k, w in ref.items():  # printed with 6 decimals
        assert abs(mine[k] - w) <= 5e-7 + 1e-6 * w
    tot = fx.read_totals(case)
    assert a.W_source_injected == pytest.approx(tot["source_injected_weight"], rel=3e-5)
    assert a.E_representation_loss == pytest.approx(tot["source_representation_loss_energy"], rel=1e-5)


def _marginal(rows):
    out = {}
    for c, b, l, w in zip(rows["cell"], rows["bid"], rows["lidx"], rows["w"]):
        if w <= 1e-12:
            continue
        el = (l // 4) % 2
        xs = (l // 8) % 8
        zs = l // 64
        k = (int(c), int(b >> 12) * 2 + el, xs, zs)
        out[k] = out.get(k, 0.0) + float(w)
    return out


@pytest.mark.parametrize("case", ["A", "B", "C", "D"])
def test_one_iteration_matches_dump(orc, case):
    e, tp, spot = _engine(orc, case)
    e.inject(spot)
    t, k5, _ = e.run_iterations(tp, 1, include_source_terms=True)
    _, rows = e.psi_dump()
    mine = _marginal(rows)
    ref = _marginal(fx.read_psi(case, "after"))
    assert set(mine) == set(ref)
    for k, w in ref.items():  # 8 printed decimals per theta row
        assert abs(mine[k] - w) <= 1.6e-8 + 1e-6 * w, (k, mine[k], w)  # <= 3 rows x 5e-9 print error
    tot = fx.read_totals(case)
    assert t[0].sum() == pytest.approx(tot["EdepC"], rel=5e-6)
    if case != "D":  # D's dumped nuclear totals equal C's although its step is shorter
        assert t[3].sum() == pytest.approx(tot["AbsorbedWeight_nuclear"], rel=5e-6)
        assert t[4].sum() == pytest.approx(tot["AbsorbedEnergy_nuclear"], rel=5e-6)
    assert t[5].sum() == tot["BoundaryLoss_weight"] == 0.0
    assert t[1].sum() == tot["AbsorbedWeight_cutoff"] == 0.0
    res = k5.E_in - (k5.E_out + k5.E_dep + k5.E_cutoff + k5.E_nuclear + k5.E_boundary +
                     k5.E_transport_drop + k5.E_source_out_of_grid + k5.E_source_slot_drop)
    assert res == pytest.approx(tot["transport_audit_residual_energy"], abs=4e-6)
    assert k5.W_pass == 1


def test_dumps_show_theta_split_is_unpinned(orc):
    """Documents the one known disagreement: the current source scatters case B's single
    component into a neighbouring theta bin with a small weight; the dump has none."""
    e, tp, spot = _engine(orc, "B")
    e.inject(spot)
    e.run_iterations(tp, 1, include_source_terms=True)
    _, rows = e.psi_dump()
    assert len(rows["w"]) == 2 and len(fx.read_psi("B", "after")["w"]) == 1
