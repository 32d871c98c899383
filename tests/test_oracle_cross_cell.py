"""Lateral emission into another listed cell's +z bucket (device_emit_lateral_spread,
device_bucket.cuh:857-987), brute force on a tiny grid: one 150 MeV component in the middle of a
row of cells 0.02 mm wide, so that the lateral spread of its +z exit (width >= 0.01 mm) reaches
two cells to either side.

Run A lists only the source cell: the weight spread toward its neighbours has no bucket and is a
drop.  Run B also lists the neighbours (components of 1e-13, pruned by the transport, keep them
on the list): the same weight now arrives, after K4, in the neighbours' +z cells, at the x sub-bin
facing the source (the last for targets on the left, the first on the right) and z sub-bin 0.
The source's own emission is the same in both runs, the dropped weight of A equals the
neighbours' weight in B, and nothing else moves."""
import numpy as np

from oracle.pyoracle import OracleEngine
from sm_2d_amd import problems

NX, NZ = 9, 4
SRC = 4  # source cell (ix 4, iz 0)


def _run(orc, listed_neighbours: bool):
    prob = problems.Problem("cross_cell", problems._grid(NX, NZ, 0.02, 0.02, 36, -0.35, 0.35, problems.LAYOUT_REF),
                            "piecewise", dict(E_fine_on=10.0, E_fine_off=11.0, max_iterations=1), [dict(E0=150.0)])
    e = problems.setup_engine(orc, prob, OracleEngine)
    e.reset()
    # block: theta bins 16-19 (b_theta 4), E bins 628-629 (b_E 314); local bin theta 2, E 1,
    # x_sub 4, z_sub 0
    bid = 4 | (314 << 12)
    lidx = 2 + 4 * (1 + 2 * (4 + 8 * 0))
    cells = [SRC] + ([c for c in range(NX) if c != SRC] if listed_neighbours else [])
    rows = dict(cell=np.array(cells), slot=np.zeros(len(cells), int), bid=np.full(len(cells), bid),
                lidx=np.full(len(cells), lidx),
                w=np.array([1.0e-3] + [1e-13] * (len(cells) - 1), np.float32))
    e.psi_load(rows)
    _, _, a = e.run_iterations(prob.make_transport(sigma_x_initial=3.0), 1)
    _, out = e.psi_dump()
    per_cell = np.zeros(NX * NZ)
    np.add.at(per_cell, out["cell"], out["w"].astype(np.float64))
    return out, per_cell.reshape(NZ, NX), a


def test_lateral_emission_reaches_listed_neighbours(orc):
    oa, wa, aa = _run(orc, False)
    ob, wb, ab = _run(orc, True)
    assert aa.cross_cell_emissions == 0 and ab.cross_cell_emissions > 0
    # the source's own +z cell gets the same weight; row 0 and rows 2.. hold nothing
    assert wa[1, SRC] == wb[1, SRC] > 0
    assert wa[0].sum() == wb[0].sum() == 0.0 and wa[2:].sum() == wb[2:].sum() == 0.0
    side = wb[1].sum() - wb[1, SRC]
    assert side > 1e-4  # a large share (~30 %) of the component spreads sideways
    assert wa[1].sum() == wa[1, SRC]
    # what A drops, B delivers to the neighbours (the same FP32 weights, split over theta bins in
    # B, so equal to FP32 rounding)
    assert abs(aa.W_transport_drop - side) <= 1e-6 * side
    assert ab.W_transport_drop < 1e-11  # only the pruned components of 1e-13
    # arrival sub-bins: x_sub 7 on the left of the source, 0 on the right, z_sub 0
    for c, l, w in zip(ob["cell"], ob["lidx"], ob["w"]):
        ix, iz = c % NX, c // NX
        if iz == 1 and ix != SRC and w > 0:
            xs, zs = (l // 8) % 8, l // 64
            assert zs == 0
            assert xs == (7 if ix < SRC else 0), (ix, xs)
