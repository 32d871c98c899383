"""GPU parity: the HIP library against the CPU oracle, through the C ABI.

The bar is bit-exactness for everything the transport step computes (block ids, masks,
compacted lists, phase-space weights, per-cell FP64 tallies): both sides evaluate the same
binary32 expressions in the same order with the same elementary functions.  K5 report sums
are FP64 reductions whose order differs (the oracle sums serially, the GPU with fixed trees),
so they are compared to 1e-12 relative.
"""
import numpy as np
import pytest

from sm_2d_amd import abi, problems
from sm_2d_amd.engine import Engine

import fixtures as fx

pytestmark = pytest.mark.gpu


def _rows_key(rows):
    return sorted(zip(rows["cell"].tolist(), rows["slot"].tolist(), rows["bid"].tolist(),
                      rows["lidx"].tolist(), rows["w"].view(np.uint32).tolist()))


def _pair(hip, orc, prob):
    from oracle.pyoracle import OracleEngine
    g = problems.setup_engine(hip, prob, Engine)
    o = problems.setup_engine(orc, prob, OracleEngine)
    return g, o


def test_math_functions_bit_exact(hip, orc):
    rng = np.random.default_rng(1)
    ranges = {0: (-80, 80), 1: (1e-12, 1e6), 2: (-7, 7), 3: (-7, 7), 4: (-7, 7), 5: (-7, 7), 6: (-7, 7)}
    for which, (a, b) in ranges.items():
        if which == 1:
            x = np.exp(rng.uniform(np.log(a), np.log(b), 200_000)).astype(np.float32)
        else:
            x = rng.uniform(a, b, 200_000).astype(np.float32)
        y_gpu = hip.math_eval(which, x)
        y_cpu = orc.math_eval(which, x)
        assert np.array_equal(y_gpu.view(np.uint32), y_cpu.view(np.uint32)), which


@pytest.mark.parametrize("case", ["A", "B", "C", "D"])
def test_injection_and_one_iteration_bit_exact(hip, orc, case):
    g_, tp, spot = fx.problem(case)
    prob = problems.Problem(f"case{case}", g_, "piecewise",
                            dict(E_fine_on=tp.E_fine_on, E_fine_off=tp.E_fine_off, max_iterations=1),
                            [dict(x0=spot.x0, z0=spot.z0, theta0=spot.theta0, E0=spot.E0,
                                  W_total=spot.W_total, sigma_x=spot.sigma_x,
                                  sigma_theta=spot.sigma_theta, sigma_E=spot.sigma_E,
                                  n_samples=spot.n_samples, random_seed=spot.random_seed)])
    g, o = _pair(hip, orc, prob)
    for e in (g, o):
        e.reset()
    ag, ao = g.inject(spot), o.inject(spot)
    bg, rg = g.psi_dump()
    bo, ro = o.psi_dump()
    assert np.array_equal(bg, bo)
    assert _rows_key(rg) == _rows_key(ro)
    assert ag.W_source_injected == ao.W_source_injected
    assert abs(ag.E_representation_loss - ao.E_representation_loss) <= 1e-9
    tpx = prob.make_transport(sigma_x_initial=spot.sigma_x)
    tg, kg, aug = g.run_iterations(tpx, 1, include_source_terms=True)
    to, ko, auo = o.run_iterations(tpx, 1, include_source_terms=True)
    bg, rg = g.psi_dump()
    bo, ro = o.psi_dump()
    assert np.array_equal(bg, bo)
    assert _rows_key(rg) == _rows_key(ro)
    assert np.array_equal(tg, to)
    for k in ("W_in", "W_out", "E_in", "E_out", "E_dep", "W_nuclear", "E_nuclear"):
        a, b = getattr(kg, k), getattr(ko, k)
        assert abs(a - b) <= 1e-12 * max(1.0, abs(b)), (k, a, b)
    assert kg.W_pass == ko.W_pass == 1
    assert aug.cell_bin_updates == auo.cell_bin_updates


def _on_grid(gr, bid, LB):
    """Local bins of block `bid` whose global (theta, E) bins lie on the grid (a partial last
    block has local bins past N_theta or N_E, which psi_load rejects)."""
    TL, EL = gr.n_theta_local, gr.n_e_local
    l = np.arange(LB)
    t = (bid & 0xFFF) * TL + l % TL
    e = ((bid >> 12) & 0xFFF) * EL + (l // TL) % EL
    return l[(t < gr.N_theta) & (e < gr.N_E)]


def _blocks(gr):
    return (gr.N_E + gr.n_e_local - 1) // gr.n_e_local, (gr.N_theta + gr.n_theta_local - 1) // gr.n_theta_local


def _random_psi(rng, gr, N, LB, n_cells=200, max_slots=6):
    NE_blocks, NT_blocks = _blocks(gr)
    cells = rng.choice(N, size=n_cells, replace=False)
    rows = {k: [] for k in ("cell", "slot", "bid", "lidx", "w")}
    for c in sorted(cells.tolist()):
        ns = int(rng.integers(1, max_slots + 1))
        bids = set()
        while len(bids) < ns:
            bids.add(int(rng.integers(0, NT_blocks)) | (int(rng.integers(0, NE_blocks)) << 12))
        for s, b in enumerate(sorted(bids)):
            ok = _on_grid(gr, b, LB)
            nl = min(int(rng.integers(1, 12)), len(ok))
            for l in rng.choice(ok, size=nl, replace=False).tolist():
                w = float(10 ** rng.uniform(-14, -2))
                rows["cell"].append(c)
                rows["slot"].append(s)
                rows["bid"].append(b)
                rows["lidx"].append(l)
                rows["w"].append(w)
    return {k: np.asarray(v) for k, v in rows.items()}


@pytest.mark.parametrize("layout", [problems.LAYOUT_REF, problems.LAYOUT_CFG])
def test_mask_compact_random_bit_exact(hip, orc, layout):
    prob = problems.reference_grid_case(Nx=60, Nz=50, layout=layout)
    g, o = _pair(hip, orc, prob)
    LB = g.LB
    rng = np.random.default_rng(3)
    rows = _random_psi(rng, g.grid, g.N, LB)
    prev = (rng.random(g.N) < 0.3).astype(np.uint8)
    for wmin in (1e-12, 1e-6):
        tp = prob.make_transport(weight_active_min=wmin)
        for e in (g, o):
            e.psi_load(rows)
        mg, ag, cg = g.mask_compact(tp, prev)
        mo, ao, co = o.mask_compact(tp, prev)
        assert np.array_equal(mg, mo)
        assert np.array_equal(ag, ao)
        assert np.array_equal(cg, co)
        assert np.all(np.diff(ag) > 0) and np.all(np.diff(cg) > 0)


@pytest.mark.parametrize("layout", [problems.LAYOUT_REF, problems.LAYOUT_CFG])
def test_random_psi_iterations_bit_exact(hip, orc, layout):
    """Random phase space (several blocks per cell, all energies) through 3 iterations: exercises
    both transports, slot claims, bucket emission to all faces and the K4 gather."""
    prob = problems.reference_grid_case(Nx=40, Nz=40, layout=layout)
    g, o = _pair(hip, orc, prob)
    rng = np.random.default_rng(11)
    rows = _random_psi(rng, g.grid, g.N, g.LB, n_cells=120, max_slots=8)
    tp = prob.make_transport(sigma_x_initial=2.0)
    for e in (g, o):
        e.psi_load(rows)
    tg, kg, aug = g.run_iterations(tp, 3)
    to, ko, auo = o.run_iterations(tp, 3)
    bg, rg = g.psi_dump()
    bo, ro = o.psi_dump()
    assert np.array_equal(bg, bo)
    assert _rows_key(rg) == _rows_key(ro)
    assert np.array_equal(tg, to)
    assert aug.cell_bin_updates == auo.cell_bin_updates
    assert aug.cross_cell_emissions == auo.cross_cell_emissions


def _step_rule_2pct(prob):
    """The same problem with the fine transport's 2 % energy-loss step rule (step_rule = 1)."""
    prob.transport = dict(prob.transport, step_rule=1)
    return prob


@pytest.mark.parametrize("prob", [
    problems.reference_grid_case(Nx=60, Nz=100, E0=70.0, sigma_x=3.0, n_samples=500),
    problems.config1(n_samples=500),
    _step_rule_2pct(problems.reference_grid_case(Nx=60, Nz=100, E0=70.0, sigma_x=3.0, n_samples=500)),
], ids=["refgrid70_LB256", "config1_LB128", "refgrid70_LB256_step_rule_2pct"])
def test_full_spot_bit_exact(hip, orc, prob):
    g, o = _pair(hip, orc, prob)
    sp = prob.make_spot()
    tp = prob.make_transport()
    eg, ag = g.run_spot(sp, tp)
    eo, ao = o.run_spot(sp, tp)
    if prob.transport.get("step_rule", 0) == 1:
        assert ag.fine_cell_bin_updates > 0  # the rule acts in the fine transport
    assert ag.fine_cell_bin_updates == ao.fine_cell_bin_updates
    assert ag.terminated == ao.terminated
    assert ag.iterations == ao.iterations
    assert np.array_equal(eg, eo)
    assert ag.W_leftover == pytest.approx(ao.W_leftover, rel=1e-12, abs=1e-15)
    assert ag.cell_bin_updates == ao.cell_bin_updates
    assert ag.processed_components == ao.processed_components
    for k in ("E_dep", "W_nuclear", "E_nuclear", "W_boundary", "E_boundary", "W_cutoff", "E_cutoff"):
        assert getattr(ag, k) == pytest.approx(getattr(ao, k), rel=1e-12, abs=1e-15), k
    assert ag.max_cell_W_closure < 1e-6
    assert ag.W_fail_iters == 0


def test_batched_spots_match_single_spots(hip, orc):
    """Spots transported together in one batch context give, spot by spot, the same EdepC as
    the single-spot path (and hence the oracle), and the FP64 sum is taken in spot order."""
    from oracle.pyoracle import OracleEngine
    base = problems.reference_grid_case(Nx=40, Nz=60, E0=60.0, sigma_x=2.0, n_samples=300)
    spots = []
    for i, (E0, x0, sx) in enumerate([(60.0, 0.0, 2.0), (45.0, -3.0, 1.5), (70.0, 4.0, 3.0)]):
        sp = base.make_spot()
        sp.E0, sp.x0, sp.sigma_x, sp.random_seed = E0, x0, sx, 1234 + 7 * i
        spots.append(sp)
    tp = base.make_transport()
    single = problems.setup_engine(hip, base, Engine)
    oracle = problems.setup_engine(orc, base, OracleEngine)
    ref, auds = [], []
    for sp in spots:
        e, a = single.run_spot(sp, tp)
        eo, _ = oracle.run_spot(sp, tp)
        assert np.array_equal(e, eo)
        ref.append(e)
        auds.append(a)
    batch = problems.setup_engine(hip, base, Engine, n_spots=4)
    esum, per, ab = batch.run_batch(spots, tp, per_spot=True)
    expect = ref[0] + ref[1] + ref[2]
    assert np.array_equal(esum, expect)
    for k in range(3):
        assert np.array_equal(per[k], ref[k].astype(np.float32))
    assert ab.terminated == int(all(a.terminated for a in auds))
    assert ab.iterations == max(a.iterations for a in auds)
    assert ab.cell_bin_updates == sum(a.cell_bin_updates for a in auds)
    assert ab.W_fail_iters == 0
    # the context is reusable: a second, smaller batch on the same context
    esum2, _, _ = batch.run_batch(spots[1:2], tp)
    assert np.array_equal(esum2, ref[1])


def test_slab_phantom_bit_exact(hip, orc):
    """Heterogeneous depth slabs (water / bone / lung rows): density-scaled steps, energy loss,
    straggling, nuclear attenuation and radiation-length-scaled scattering on both sides."""
    Nz = 100
    prob = problems.reference_grid_case(Nx=40, Nz=Nz, E0=70.0, sigma_x=2.0, n_samples=300,
                                        max_iterations=500)
    rho = np.ones(Nz, np.float32)
    x0r = np.ones(Nz, np.float32)
    rho[10:25], x0r[10:25] = 1.85, 0.55      # bone
    rho[25:45], x0r[25:45] = 0.26, 1 / 0.26  # lung
    prob.materials = (rho, x0r)
    g, o = _pair(hip, orc, prob)
    sp, tp = prob.make_spot(), prob.make_transport()
    eg, ag = g.run_spot(sp, tp)
    eo, ao = o.run_spot(sp, tp)
    assert ag.iterations == ao.iterations
    assert np.array_equal(eg, eo)
    assert ag.cell_bin_updates == ao.cell_bin_updates
    assert ag.W_fail_iters == 0


@pytest.mark.parametrize("waves", [1, 2, 4])
@pytest.mark.parametrize("prob", [
    problems.reference_grid_case(Nx=40, Nz=80, E0=60.0, sigma_x=2.5, n_samples=400, max_iterations=300),
    problems.config1(n_samples=400),
], ids=["refgrid60_LB256", "config1_LB128"])
def test_transport_shapes_bit_exact(hip, orc, prob, waves):
    """Every launch shape of the transport (waves per listed cell) reproduces the oracle."""
    g, o = _pair(hip, orc, prob)
    g.set_transport_waves(waves)
    sp, tp = prob.make_spot(), prob.make_transport()
    eg, ag = g.run_spot(sp, tp)
    eo, ao = o.run_spot(sp, tp)
    assert ag.iterations == ao.iterations
    assert np.array_equal(eg, eo)
    assert ag.cell_bin_updates == ao.cell_bin_updates


@pytest.mark.parametrize("split", [[1, 3], [2, 1, 1]], ids=["1+3", "2+1+1"])
def test_split_iteration_runs_match_one_run(hip, orc, split):
    """Iterations queued over several calls continue exactly where the last call stopped (epochs,
    mask hysteresis, buffer parity), so the state equals one call over all of them and the
    oracle's."""
    prob = problems.reference_grid_case(Nx=40, Nz=40)
    g, o = _pair(hip, orc, prob)
    rng = np.random.default_rng(11)
    rows = _random_psi(rng, g.grid, g.N, g.LB, n_cells=120, max_slots=8)
    tp = prob.make_transport(sigma_x_initial=2.0)
    for e in (g, o):
        e.psi_load(rows)
    for n in split:
        tg, _, _ = g.run_iterations(tp, n)
    to, _, _ = o.run_iterations(tp, sum(split))
    bg, rg = g.psi_dump()
    bo, ro = o.psi_dump()
    assert np.array_equal(bg, bo)
    assert _rows_key(rg) == _rows_key(ro)
    assert np.array_equal(tg, to)


@pytest.mark.parametrize("change", [dict(E_fine_on=30.0, E_fine_off=33.0), dict(weight_active_min=1e-5)],
                         ids=["fine_on", "weight_min"])
def test_split_runs_with_changed_thresholds(hip, orc, change):
    """A call whose K1 thresholds differ from the previous call's decides its first iteration's
    mask with its own thresholds (the previous call's closing audit decided that iteration ahead
    with the old ones), from the same previous-iteration mask as the oracle."""
    prob = problems.reference_grid_case(Nx=40, Nz=40)
    g, o = _pair(hip, orc, prob)
    rng = np.random.default_rng(17)
    rows = _random_psi(rng, g.grid, g.N, g.LB, n_cells=150, max_slots=8)
    tp1 = prob.make_transport(sigma_x_initial=2.0)
    tp2 = prob.make_transport(sigma_x_initial=2.0, **change)
    for e in (g, o):
        e.psi_load(rows)
    for tp in (tp1, tp2, tp1):
        tg, _, ag = g.run_iterations(tp, 2)
        to, _, ao = o.run_iterations(tp, 2)
        assert ag.processed_cells == ao.processed_cells
        assert ag.cell_bin_updates == ao.cell_bin_updates
    bg, rg = g.psi_dump()
    bo, ro = o.psi_dump()
    assert np.array_equal(bg, bo)
    assert _rows_key(rg) == _rows_key(ro)
    assert np.array_equal(tg, to)


def test_chunk_graphs_replay_bit_exact(hip, orc):
    """Chunks of iterations captured as graphs and replayed (with eager host work and other
    launches between the replays) give the oracle's dose on every run, as do eager launches."""
    prob = problems.reference_grid_case(Nx=40, Nz=80, E0=60.0, sigma_x=2.5, n_samples=400,
                                        max_iterations=300)
    g, o = _pair(hip, orc, prob)
    sp, tp = prob.make_spot(), prob.make_transport()
    eo, ao = o.run_spot(sp, tp)
    g.set_graphs(True)
    for rep in range(3):
        eg, ag = g.run_spot(sp, tp)
        assert ag.iterations == ao.iterations, rep
        assert np.array_equal(eg, eo), rep
        assert ag.cell_bin_updates == ao.cell_bin_updates, rep
        g.psi_dump()  # eager copies between replays
    rep = g.tuner_report()["chunk_graphs"]
    assert rep["on"] == 1 and rep["captured"] > 0 and rep["replayed"] > 0
    g.set_graphs(False)
    eg, ag = g.run_spot(sp, tp)
    assert np.array_equal(eg, eo)
    assert ag.iterations == ao.iterations


@pytest.mark.parametrize("fine_on", [10.0, 300.0], ids=["coarse-then-fine", "fine-only"])
def test_phase_skipping_chunks_bit_exact(hip, orc, fine_on):
    """Graph chunks leave out a phase whose list was empty when the previous chunk ended.  With
    a 60 MeV beam (0.5 mm cells) the fine phase starts after the first chunks: its first list pauses the
    transport before that iteration runs and the host queues the rest again with both phases.
    With every energy in the fine phase the coarse kernels stay out while the coarse list is
    empty.  Both equal the oracle."""
    prob = problems.reference_grid_case(Nx=40, Nz=80, E0=60.0, sigma_x=2.0, n_samples=300,
                                        max_iterations=400)
    prob.transport.update(E_fine_on=fine_on, E_fine_off=fine_on + 1.0)
    g, o = _pair(hip, orc, prob)
    sp, tp = prob.make_spot(), prob.make_transport()
    eo, ao = o.run_spot(sp, tp)
    g.set_graphs(True)
    g.set_transport_waves(2)  # no tuner exploration: every chunk is a graph chunk
    eg, ag = g.run_spot(sp, tp)
    assert ag.iterations == ao.iterations and ag.terminated == ao.terminated
    assert np.array_equal(eg, eo)
    assert ag.cell_bin_updates == ao.cell_bin_updates
    rep = g.tuner_report()["chunk_graphs"]
    assert rep["phase_skipping_chunks"] > 0
    # the fine phase's arrival pauses once; with every energy fine, a coarse list can still come
    # back (cells whose weight fell below w_min are coarse-listed), and after one pause a run
    # keeps both phases
    assert rep["pauses"] == 1 if fine_on == 10.0 else rep["pauses"] <= 1


def test_phase_skipping_long_lists(hip):
    """Phase skipping with lists too long for the one-workgroup compaction: 128 spots on the
    1029-bin grid start coarse-only (the fine kernels stay out of the graph chunks) until the
    first fine list, which k_scan<2> turns into a pause.  The graph run equals the eager run
    (no chunk leaves a phase out there) spot by spot."""
    prob = problems.config4_fine_grid(n_spots=128, n_samples=300)
    spots = [prob.make_spot(i) for i in range(len(prob.spots))]
    tp = prob.make_transport()
    g = problems.setup_engine(hip, prob, Engine, n_spots=len(spots))
    g.set_transport_waves(1)  # no tuner exploration: every chunk is a graph chunk
    import time
    g.set_graphs(False)
    t0 = time.perf_counter()
    ee, pe, ae = g.run_batch(spots, tp, per_spot=True)
    t_eager = time.perf_counter() - t0
    g.set_graphs(True)
    g.run_batch(spots, tp, per_spot=True)  # captures the chunk graphs
    t0 = time.perf_counter()
    eg, pg, ag = g.run_batch(spots, tp, per_spot=True)
    t_graph = time.perf_counter() - t0
    rep = g.tuner_report()["chunk_graphs"]
    g.close()
    # the chunk queued after the pause is sized from the list counts of the iteration that paused
    # (a pause that cleared them left that chunk with 64-workgroup launches, about twice as slow)
    assert t_graph < 1.5 * t_eager, (t_graph, t_eager)
    assert ae.fine_cell_bin_updates > 0 and ae.iterations == ag.iterations
    assert rep["phase_skipping_chunks"] > 0 and rep["pauses"] == 2, rep  # one per graph run
    assert ag.cell_bin_updates == ae.cell_bin_updates
    assert np.array_equal(eg, ee)
    for k in range(len(spots)):
        assert np.array_equal(pg[k], pe[k]), k


def _saturating_psi(rng, g, LB):
    """Cells spaced 7 apart in x and 3 in z (no lateral emission into another listed cell), each
    with 32 occupied blocks over the upper half of the energy range and 8-39 weighted bins per
    block: the transport's own and bucket tables fill (32 slots) and its LDS rows run out, so the
    closest-block fallback and the serial apply with global payloads are taken."""
    NEb, NTb = _blocks(g)
    rows = {k: [] for k in ("cell", "slot", "bid", "lidx", "w")}
    for c in [iz * g.Nx + ix for iz in range(1, g.Nz - 1, 3) for ix in range(3, g.Nx - 2, 7)]:
        bids = set()
        while len(bids) < 32:
            bids.add(int(rng.integers(0, NTb)) | (int(rng.integers(NEb // 2, NEb)) << 12))
        for s, b in enumerate(sorted(bids)):
            ok = _on_grid(g, b, LB)
            for lidx in rng.choice(ok, size=min(int(rng.integers(8, 40)), len(ok)), replace=False).tolist():
                rows["cell"].append(c)
                rows["slot"].append(s)
                rows["bid"].append(b)
                rows["lidx"].append(lidx)
                rows["w"].append(float(10 ** rng.uniform(-6, -2)))
    return {k: np.asarray(v) for k, v in rows.items()}


@pytest.mark.parametrize("kib", [8, 4])
@pytest.mark.parametrize("prob", [
    problems.reference_grid_case(Nx=40, Nz=80, E0=60.0, sigma_x=2.5, n_samples=400, max_iterations=300),
    problems.config1(n_samples=400),
], ids=["refgrid60_LB256", "config1_LB128"])
def test_gather_shapes_bit_exact(hip, orc, prob, kib):
    """Every launch shape of the K4 gather (LDS rows per destination; destinations with more
    blocks add into global payloads) reproduces the oracle."""
    g, o = _pair(hip, orc, prob)
    g.set_gather_rows(kib)
    sp, tp = prob.make_spot(), prob.make_transport()
    eg, ag = g.run_spot(sp, tp)
    eo, ao = o.run_spot(sp, tp)
    assert ag.iterations == ao.iterations
    assert np.array_equal(eg, eo)
    assert ag.cell_bin_updates == ao.cell_bin_updates


@pytest.mark.parametrize("kib", [16, 4])
@pytest.mark.parametrize("layout", [problems.LAYOUT_REF, problems.LAYOUT_CFG])
def test_saturated_slot_tables_bit_exact(hip, orc, layout, kib):
    prob = problems.reference_grid_case(Nx=42, Nz=12, layout=layout)
    g, o = _pair(hip, orc, prob)
    g.set_gather_rows(kib)
    rows = _saturating_psi(np.random.default_rng(5), g.grid, g.LB)
    tp = prob.make_transport(sigma_x_initial=2.0)
    for e in (g, o):
        e.psi_load(rows)
    tg, _, ag = g.run_iterations(tp, 2)
    to, _, ao = o.run_iterations(tp, 2)
    bg, rg = g.psi_dump()
    bo, ro = o.psi_dump()
    assert np.array_equal(bg, bo)
    assert _rows_key(rg) == _rows_key(ro)
    assert np.array_equal(tg, to)
    assert ag.cell_bin_updates == ao.cell_bin_updates
    assert ag.W_transport_drop == pytest.approx(ao.W_transport_drop, rel=1e-12, abs=1e-18)
    if layout == problems.LAYOUT_CFG:  # the tables did fill
        assert max(len(np.unique(ro["slot"][ro["cell"] == c])) for c in np.unique(ro["cell"])) == 32


def test_psi_load_rejects_bins_off_the_grid(hip, orc):
    """The 1029-bin energy grid in blocks of 4 ends in a partial block: its bins past N_E are
    rejected by both loaders, the last bins on the grid load."""
    from sm_2d_amd.engine import LibError
    prob = problems.reference_grid_case(Nx=8, Nz=8, layout=problems.LAYOUT_CFG)
    g, o = _pair(hip, orc, prob)
    NEb, NTb = _blocks(g.grid)
    b = (NTb - 1) | ((NEb - 1) << 12)
    ok = _on_grid(g.grid, b, g.LB)
    bad = np.setdiff1d(np.arange(g.LB), ok)
    assert len(ok) > 0 and len(bad) > 0

    def row(lidx):
        return {"cell": np.array([9]), "slot": np.array([0]), "bid": np.array([b], np.uint32),
                "lidx": np.array([lidx]), "w": np.array([1e-3])}

    for e in (g, o):
        e.psi_load(row(int(ok[-1])))
        with pytest.raises(LibError, match="outside the grid"):
            e.psi_load(row(int(bad[0])))


@pytest.mark.parametrize("layout", [problems.LAYOUT_REF, problems.LAYOUT_CFG])
def test_cross_cell_lateral_emission_bit_exact(hip, orc, layout):
    """Cells of 0.05 mm are narrower than the coarse transport's lateral spread, so +z exits emit
    into the +z buckets of listed neighbours (device_bucket.cuh:952-983) throughout the spot: the
    GPU applies them after the phase's own emissions, in source-cell and component order, as the
    oracle does."""
    prob = problems.reference_grid_case(Nx=60, Nz=60, E0=70.0, sigma_x=0.4, n_samples=300,
                                        max_iterations=200, layout=layout)
    prob.grid.update(dx=0.05, dz=0.05, x_min=-1.5)
    g, o = _pair(hip, orc, prob)
    sp, tp = prob.make_spot(), prob.make_transport()
    eg, ag = g.run_spot(sp, tp)
    eo, ao = o.run_spot(sp, tp)
    assert ao.cross_cell_emissions > 100_000
    assert ag.cross_cell_emissions == ao.cross_cell_emissions
    assert ag.iterations == ao.iterations
    assert np.array_equal(eg, eo)
    assert ag.cell_bin_updates == ao.cell_bin_updates
    assert ag.W_transport_drop == pytest.approx(ao.W_transport_drop, rel=1e-12, abs=1e-18)
    assert ag.max_cell_W_closure < 1e-6 and ag.W_fail_iters == 0


def test_cross_cell_routing_matches_oracle(hip, orc):
    """One component on 0.02 mm cells with its neighbours listed: the phase space after one
    iteration equals the oracle's, slot for slot (tests/test_oracle_cross_cell.py pins the
    oracle's routing)."""
    prob = problems.Problem("cross_cell", problems._grid(9, 4, 0.02, 0.02, 36, -0.35, 0.35, problems.LAYOUT_REF),
                            "piecewise", dict(E_fine_on=10.0, E_fine_off=11.0, max_iterations=1), [dict(E0=150.0)])
    g, o = _pair(hip, orc, prob)
    bid = 4 | (314 << 12)
    lidx = 2 + 4 * (1 + 2 * 4)
    cells = [4, 0, 1, 2, 3, 5, 6, 7, 8]
    rows = dict(cell=np.array(cells), slot=np.zeros(9, int), bid=np.full(9, bid), lidx=np.full(9, lidx),
                w=np.array([1.0e-3] + [1e-13] * 8, np.float32))
    tp = prob.make_transport(sigma_x_initial=3.0)
    for e in (g, o):
        e.reset()
        e.psi_load(rows)
    _, _, ag = g.run_iterations(tp, 1)
    _, _, ao = o.run_iterations(tp, 1)
    bg, rg = g.psi_dump()
    bo, ro = o.psi_dump()
    assert np.array_equal(bg, bo)
    assert _rows_key(rg) == _rows_key(ro)
    assert ag.cross_cell_emissions == ao.cross_cell_emissions > 0
