"""Heterogeneous media in the oracle: the density and radiation-length scalings that the slab
phantom relies on, checked against what the physics fixes (range scales as 1/rho; water
material tables reproduce the water run exactly)."""
import numpy as np
import pytest

from sm_2d_amd import problems


def _pencil(orc, rho_value, x0_ratio, Nz=120):
    from oracle.pyoracle import OracleEngine
    prob = problems.reference_grid_case(Nx=20, Nz=Nz, E0=70.0, sigma_x=0.0, n_samples=1,
                                        max_iterations=600)
    prob.spots[0].update(sigma_E=0.0)
    if rho_value is not None:
        prob.materials = (np.full(Nz, rho_value, np.float32), np.full(Nz, x0_ratio, np.float32))
    eng = problems.setup_engine(orc, prob, OracleEngine)
    edep, au = eng.run_spot(prob.make_spot(), prob.make_transport())
    eng.close()
    return edep, au


def test_water_material_equals_default(orc):
    e0, a0 = _pencil(orc, None, None)
    e1, a1 = _pencil(orc, 1.0, 1.0)
    assert np.array_equal(e0, e1)
    assert a0.iterations == a1.iterations


def test_density_scales_range(orc):
    """CSDA range in a uniform medium of density rho is R_water / rho: the Bragg peak of a
    70 MeV pencil (R_water = 40.8 mm) moves from ~40 mm to ~20 mm at rho = 2."""
    dz = 0.5
    pk = {}
    for rho in (1.0, 2.0):
        e, au = _pencil(orc, rho, 1.0 / rho)
        assert au.W_leftover < 0.25  # slow low-energy remnants may outlive the iteration cap
        pdd = e.sum(axis=1)
        pk[rho] = (int(np.argmax(pdd)) + 0.5) * dz
    assert 38.0 < pk[1.0] < 43.0
    assert pk[2.0] == pytest.approx(pk[1.0] / 2.0, abs=1.5 * dz)


def test_slab_materials_are_seeded():
    r1, x1 = problems.slab_materials(1200, 0.25, 4)
    r2, x2 = problems.slab_materials(1200, 0.25, 4)
    assert np.array_equal(r1, r2) and np.array_equal(x1, x2)
    assert set(np.unique(r1).tolist()) == {np.float32(0.26), np.float32(1.0), np.float32(1.85)}
