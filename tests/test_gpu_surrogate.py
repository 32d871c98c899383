"""GPU tests of the learned straggling / scattering / attenuation physics of the fine transport.

The network (4-128-128-128-33, GELU, BF16 operands on the matrix cores, FP32 accumulation) is
checked against the numpy restatement in oracle/mlp_ref.py (forward pass and gradients, which
that file pins by finite differences), trained on the device, and then used through its
per-material lookup table, which is validated node by node against the plain physics.  With the
table fixed, the transport is deterministic and must equal the CPU oracle bit for bit (the
oracle reads the same table), including out-of-range rows and rejected nodes.  The learned physics changes the sampled straggling and scattering
widths, so its dose is compared with the plain physics through integral quantities; the
voxel-level comparison is documented with the measured sensitivity of the plain path itself.
"""
import numpy as np
import pytest

from sm_2d_amd import problems
from sm_2d_amd.engine import Engine, Mlp

pytestmark = pytest.mark.gpu


def _problem():
    """60 MeV pencil beam in water, every energy in the fine transport (the learned physics
    replaces the fine-transport widths), reference piecewise energy grid."""
    prob = problems.reference_grid_case(Nx=40, Nz=80, E0=60.0, sigma_x=2.0, n_samples=600,
                                        max_iterations=400)
    prob.transport.update(E_fine_on=300.0, E_fine_off=300.0)
    return prob


@pytest.fixture(scope="module")
def trained(hip):
    prob = _problem()
    eng = problems.setup_engine(hip, prob, Engine)
    m = Mlp(hip, seed=11)
    hist = m.train(eng, n_hist=6, n_samples=1 << 21, batch=4096, steps=4000, lr=2e-3, w_mom=4.0,
                   w_att=4.0)
    yield prob, eng, m, hist
    m.close()
    eng.close()


def test_forward_matches_numpy(hip):
    from oracle import mlp_ref as R
    rng = np.random.default_rng(0)
    m = Mlp(hip, seed=3)
    p, _ = m.params()
    x = rng.uniform(-1, 1, (256, 4)).astype(np.float32)
    z = m.forward(x)
    zr, _ = R.forward(p, x, emulate=True)
    # same BF16 operands; only the FP32 accumulation order differs
    assert np.abs(z - zr).max() <= 1e-3 * np.abs(zr).max()
    m.close()


def test_gradients_match_numpy(hip):
    from oracle import mlp_ref as R
    rng = np.random.default_rng(1)
    m = Mlp(hip, seed=5)
    p, _ = m.params()
    dict4 = np.array([-7.0, 0.5, -10.0, -0.5], np.float32)
    m.set_params(dict4=dict4)
    x = rng.uniform(-1, 1, (256, 4)).astype(np.float32)
    lab = np.stack([rng.uniform(-7, 0.5, 256), rng.uniform(-10, -0.5, 256),
                    rng.uniform(-20, -4, 256)], axis=1).astype(np.float32)
    for w_att, w_mom in ((1.0, 1.0), (0.5, 4.0)):
        loss, g = m.grad(x, lab, w_att, w_mom)
        lr_, gr = R.loss_and_grad(p, x, lab, dict4, w_att, w_mom, emulate=True)
        assert loss == pytest.approx(lr_, rel=1e-4)
        for name, off, shape in R.SHAPES:
            n = int(np.prod(shape))
            a, b = g[off:off + n], gr[off:off + n]
            assert np.linalg.norm(a - b) <= 2e-3 * np.linalg.norm(b) + 1e-9, name
    m.close()


def test_training_learns_the_plain_physics(trained):
    _, eng, m, hist = trained
    assert np.all(np.isfinite(hist))
    assert hist[-1] < 0.05 * hist[0]
    assert hist[-1] <= hist[1]
    e = m.evaluate(eng, 65536, seed=123)
    # held-out RMS errors of ln sigma_E and ln sigma_theta (fresh samples)
    assert e[0] < 0.1 and e[1] < 0.1
    # ln(a + 1e-9) of the attenuation exponent a (largest errors where sigma_n vanishes, 5 MeV)
    assert e[2] < 0.35


def test_learned_transport_bit_exact_vs_oracle(hip, trained, orc):
    from oracle.pyoracle import OracleEngine
    prob, _, m, _ = trained
    g = problems.setup_engine(hip, prob, Engine)
    o = problems.setup_engine(orc, prob, OracleEngine)
    g.surrogate_attach(m, 64)
    tab, nm, nn, rows = g.surrogate_export()
    assert np.all(rows >= 0)
    o.surrogate_set(tab, nm, nn, rows)
    sp, tp = prob.make_spot(), prob.make_transport()
    eg, ag = g.run_spot(sp, tp)
    eo, ao = o.run_spot(sp, tp)
    assert ag.surrogate_active == 1 and ag.surrogate_fallback == 0
    assert ag.iterations == ao.iterations
    assert ag.cell_bin_updates == ao.cell_bin_updates
    assert np.array_equal(eg, eo)
    assert ag.max_cell_W_closure < 1e-6 and ag.W_fail_iters == 0


def test_batched_learned_transport_equals_single_spots(hip, trained):
    """The learned physics in a batch context: three all-fine spots transported together with
    the single-spot context's table give, spot by spot, the single-spot learned EdepC, and the
    FP64 sum of those in spot order."""
    prob, _, m, _ = trained
    single = problems.setup_engine(hip, prob, Engine)
    single.surrogate_attach(m, 64)
    tab, nm, nn, rows = single.surrogate_export()
    spots = []
    for i, (E0, x0, sx) in enumerate([(60.0, 0.0, 2.0), (45.0, -3.0, 1.5), (55.0, 4.0, 2.5)]):
        sp = prob.make_spot()
        sp.E0, sp.x0, sp.sigma_x, sp.random_seed = E0, x0, sx, 77 + i
        spots.append(sp)
    tp = prob.make_transport()
    ref = []
    for sp in spots:
        e, a = single.run_spot(sp, tp)
        assert a.surrogate_active == 1 and a.surrogate_fallback == 0
        ref.append(e)
    single.close()
    batch = problems.setup_engine(hip, prob, Engine, n_spots=3)
    batch.surrogate_set_table(tab, nm, nn, rows)
    esum, per, ab = batch.run_batch(spots, tp, per_spot=True)
    batch.close()
    assert ab.surrogate_active == 1 and ab.surrogate_fallback == 0
    for k in range(3):
        assert np.array_equal(per[k], ref[k].astype(np.float32)), k
    assert np.array_equal(esum, ref[0] + ref[1] + ref[2])
    assert ab.max_cell_W_closure < 1e-6 and ab.W_fail_iters == 0


def test_out_of_range_material_uses_plain_physics(hip, trained):
    prob, _, m, _ = trained
    g = problems.setup_engine(hip, prob, Engine)
    sp, tp = prob.make_spot(), prob.make_transport()
    rho = np.full(g.grid.Nz, 2.5, np.float32)  # above the training range (0.2 - 2.0)
    x0r = np.full(g.grid.Nz, 0.5, np.float32)
    g.upload_material(rho, x0r)
    ep, _ = g.run_spot(sp, tp)
    g.surrogate_attach(m, 64)
    _, _, _, rows = g.surrogate_export()
    assert np.all(rows == -1)
    es, a = g.run_spot(sp, tp)
    assert np.array_equal(es, ep)


def test_validation_rejects_a_perturbed_network(hip, trained, orc):
    """Weights perturbed after training (drift injection): the table check against the plain
    physics rejects the nodes they spoil, those nodes run the plain physics, and the result still
    equals the oracle reading the same (validated) table."""
    from oracle.pyoracle import OracleEngine
    prob, _, m, _ = trained
    g = problems.setup_engine(hip, prob, Engine)
    g.surrogate_attach(m, 64)
    n_ok, rej_ok, mx_ok, rms_ok = g.surrogate_stats()
    assert n_ok > 0 and rej_ok <= 0.05 * n_ok and rms_ok < 0.1, (n_ok, rej_ok, mx_ok, rms_ok)
    p, d = m.params()
    rng = np.random.default_rng(5)
    bad = Mlp(hip, seed=11)
    bad.set_params((p * (1.0 + 0.3 * rng.standard_normal(p.shape))).astype(np.float32), d)
    g.surrogate_attach(bad, 64)
    n_bad, rej_bad, mx_bad, _ = g.surrogate_stats()
    assert rej_bad > 10 * max(rej_ok, 1.0) and mx_bad > 0.25, (n_bad, rej_bad, mx_bad, rej_ok)
    # every node left in the table is within the tolerance
    n2, rej2, mx2, _ = g.surrogate_check(0.25)
    assert rej2 == 0 and mx2 <= 0.25
    tab, nm, nn, rows = g.surrogate_export()
    o = problems.setup_engine(orc, prob, OracleEngine)
    o.surrogate_set(tab, nm, nn, rows)
    sp, tp = prob.make_spot(), prob.make_transport()
    eg, ag = g.run_spot(sp, tp)
    eo, ao = o.run_spot(sp, tp)
    assert ag.iterations == ao.iterations
    assert np.array_equal(eg, eo)
    bad.close()


def test_learned_dose_against_plain(hip, trained):
    """Integral agreement with the plain physics.  Voxel doses of this transport are not a smooth
    function of the widths (each component takes one hashed draw into one bin), so they are
    compared in 4 x 4 blocks; the bounds sit above the measured change of the plain path itself
    under a 1 % change of the widths (block mean 5 %)."""
    prob, _, m, _ = trained
    g = problems.setup_engine(hip, prob, Engine)
    sp, tp = prob.make_spot(), prob.make_transport()
    ep, ap = g.run_spot(sp, tp)
    g.surrogate_attach(m, 64)
    es, as_ = g.run_spot(sp, tp)
    assert as_.surrogate_active == 1 and as_.surrogate_fallback == 0
    assert abs(es.sum() / ep.sum() - 1.0) < 0.01

    def blk(d, b=4):
        nz, nx = d.shape
        return d[: nz // b * b, : nx // b * b].reshape(nz // b, b, nx // b, b).sum((1, 3))

    bp, bs = blk(ep), blk(es)
    mask = bp > 0.1 * bp.max()
    rel = np.abs(bs - bp)[mask] / bp[mask]
    assert rel.mean() < 0.12
    assert as_.max_cell_W_closure < 1e-6
