"""The -O3 -march=x86-64-v3 build of the oracle (the CPU baseline of bench.py) computes what the
-O2 build computes, bit for bit: floating-point contraction is off in both, and no other
optimisation may reorder IEEE operations."""
import numpy as np

from oracle import pyoracle
from oracle.pyoracle import OracleEngine
from sm_2d_amd import problems


def test_fast_build_equals_reference_build():
    prob = problems.reference_grid_case(Nx=40, Nz=60, E0=60.0, sigma_x=2.0, n_samples=200,
                                        max_iterations=200)
    out = []
    for fast in (False, True):
        lib = pyoracle.load(fast=fast)
        eng = problems.setup_engine(lib, prob, OracleEngine)
        e, a = eng.run_spot(prob.make_spot(), prob.make_transport())
        out.append((e, a))
        eng.close()
    (e0, a0), (e1, a1) = out
    assert np.array_equal(e0, e1)
    assert a0.iterations == a1.iterations and a0.cell_bin_updates == a1.cell_bin_updates
    assert a0.fine_cell_bin_updates == a1.fine_cell_bin_updates > 0
