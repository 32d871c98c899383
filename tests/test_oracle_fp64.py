"""The oracle's binary64 build (oracle/_build/liboracle_sm2d_fp64.so): the same transport step
with every binary32 quantity in double.  The FP32 results (oracle, and hence the bit-identical
HIP library) state their error against it.

What pins the FP64 build: it is the FP32 source compiled with one type switch, so the two agree
on every discrete outcome (iterations, cell-bin updates) on these cases. Its per-cell weight
closure is at double-precision level (a term dropped or double-counted in the weight flow
would leave a closure of the size of that term). And its dose lies within the FP32 rounding
error of the pinned FP32 oracle.
"""
import numpy as np
import pytest

from sm_2d_amd import problems

from fixtures import dose_errors


@pytest.fixture(scope="module")
def orc64():
    from oracle.pyoracle import load
    return load(fp64=True)


def _run(lib, prob):
    from oracle.pyoracle import OracleEngine
    e = problems.setup_engine(lib, prob, OracleEngine)
    return e.run_spot(prob.make_spot(), prob.make_transport())


def test_fp32_error_against_fp64_small_spot(orc, orc64):
    """70 MeV on 80 x 120 cells: the FP32 dose is within 1.7e-5 relative of the FP64 dose
    (measured 1.65e-5 on voxels above 1 % of Dmax, 1.58e-5 on the PDD)."""
    prob = problems.config1(n_samples=1000)
    e32, a32 = _run(orc, prob)
    e64, a64 = _run(orc64, prob)
    assert a32.iterations == a64.iterations and a32.terminated == a64.terminated == 1
    err = dose_errors(e32, e64)
    print("FP32 vs FP64:", err, flush=True)
    assert err[0.01] < 5e-5 and err["pdd"] < 5e-5
    assert abs(a32.E_dep - a64.E_dep) < 5e-5 * a64.E_dep
    assert a64.max_cell_W_closure < 1e-12   # double-precision closure
    assert a32.max_cell_W_closure < 1e-6    # the FP32 bar
    assert a64.W_fail_iters == 0 and a32.W_fail_iters == 0


def test_fp64_build_differs_only_by_rounding(orc, orc64):
    """Three iterations from the same injected phase space: the same bins are occupied, and the
    weights agree to FP32 rounding.  Small tail weights of the lateral spread carry the
    rounding of erf differences (up to ~2e-3 of themselves), so the bound is taken against the
    largest weight; the total weight agrees to 1e-6."""
    from oracle.pyoracle import OracleEngine
    prob = problems.reference_grid_case(Nx=40, Nz=60, E0=60.0, sigma_x=2.0, n_samples=300)
    outs = []
    for lib in (orc, orc64):
        e = problems.setup_engine(lib, prob, OracleEngine)
        e.inject(prob.make_spot())
        e.run_iterations(prob.make_transport(), 3)
        _, rows = e.psi_dump()
        outs.append(rows)
    key = [sorted(zip(r["cell"].tolist(), r["bid"].tolist(), r["lidx"].tolist())) for r in outs]
    assert key[0] == key[1]
    w = [np.array([x for _, x in sorted(zip(zip(r["cell"].tolist(), r["bid"].tolist(), r["lidx"].tolist()),
                                               r["w"].tolist()))]) for r in outs]
    assert (np.abs(w[0] - w[1]) / w[1].max()).max() < 1e-5
    assert abs(w[0].sum() - w[1].sum()) < 1e-6 * w[1].sum()
