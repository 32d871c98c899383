"""The allocation-free K1 + K2 entry point (sm2d_k1k2_mask_compact_noalloc) against the
allocating one (sm2d_k1k2_mask_compact) on the same phase space, through the C ABI: the same
mask and the same ascending active and coarse lists, bit for bit, on two calls in a row (the
second on a smaller grid, so stale scratch contents of the first would show), in both local-bin
layouts; sizes beyond the scratch limit are rejected before any device work."""
import ctypes as C

import numpy as np
import pytest

from sm_2d_amd import abi, problems
from sm_2d_amd.engine import Engine

from test_gpu_parity import _random_psi

pytestmark = pytest.mark.gpu


def _fn(hip):
    f = hip.lib.sm2d_k1k2_mask_compact_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                  C.c_int, C.c_int, C.c_float, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    return f


def _b_threshold(edges, E, n_e_local):
    """compute_b_E_threshold (k1k6_pipeline.cuh:419-422): the energy block of E's bin."""
    NE = len(edges) - 1
    b = 0 if E < edges[0] else NE - 1 if E >= edges[NE] else int(np.searchsorted(edges, E, side="right")) - 1
    return b // n_e_local


def _check(hip, layout, Nx, Nz, seed, wmin):
    import torch
    prob = problems.reference_grid_case(Nx=Nx, Nz=Nz, layout=layout)
    g = problems.setup_engine(hip, prob, Engine)
    N, LB = g.N, g.LB
    rng = np.random.default_rng(seed)
    rows = _random_psi(rng, g.grid, N, LB, n_cells=min(N // 2, 200), max_slots=6)
    prev = (rng.random(N) < 0.3).astype(np.uint8)
    tp = prob.make_transport(weight_active_min=wmin)
    g.psi_load(rows)
    mask, act, crs = g.mask_compact(tp, prev)
    # the same phase space in the reference's dense layout
    bids = np.full(N * abi.KB, abi.EMPTY, np.uint32)
    vals = np.zeros(N * abi.KB * LB, np.float32)
    for c, s, b, l, w in zip(rows["cell"], rows["slot"], rows["bid"], rows["lidx"], rows["w"]):
        bids[c * abi.KB + s] = b
        vals[(c * abi.KB + s) * LB + l] = w
    dev = torch.device("cuda:0")
    t_bid = torch.from_numpy(bids.view(np.int32)).to(dev)
    t_val = torch.from_numpy(vals).to(dev)
    t_prev = torch.from_numpy(prev).to(dev)
    t_mask = torch.full((N,), 7, dtype=torch.uint8, device=dev)
    t_act = torch.full((N,), -1, dtype=torch.int32, device=dev)
    t_crs = torch.full((N,), -1, dtype=torch.int32, device=dev)
    t_cnt = torch.full((2,), -1, dtype=torch.int32, device=dev)
    edges = g.tables.E_edges
    b_on = _b_threshold(edges, tp.E_fine_on, g.grid.n_e_local)
    b_off = _b_threshold(edges, tp.E_fine_off, g.grid.n_e_local)
    stream = torch.cuda.current_stream().cuda_stream
    rc = _fn(hip)(Nx, Nz, LB, t_bid.data_ptr(), t_val.data_ptr(), t_prev.data_ptr(), t_mask.data_ptr(),
                  b_on, b_off, tp.weight_active_min, t_act.data_ptr(), t_crs.data_ptr(), t_cnt.data_ptr(), stream)
    assert rc == abi.OK
    torch.cuda.synchronize()
    na, nc = (int(v) for v in t_cnt.cpu().numpy())
    assert (na, nc) == (len(act), len(crs))
    assert np.array_equal(t_mask.cpu().numpy(), mask)
    assert np.array_equal(t_act.cpu().numpy()[:na], act)
    assert np.array_equal(t_crs.cpu().numpy()[:nc], crs)
    assert len(act) > 0 and len(crs) > 0
    g.close()


@pytest.mark.parametrize("layout", [problems.LAYOUT_REF, problems.LAYOUT_CFG])
def test_noalloc_mask_compact_equals_allocating_entry_point(hip, layout):
    _check(hip, layout, 60, 50, seed=5, wmin=1e-6)
    _check(hip, layout, 30, 20, seed=6, wmin=1e-12)  # smaller, right after the first


def test_noalloc_rejects_grids_beyond_the_scratch(hip):
    rc = _fn(hip)(4096, 2048, 128, None, None, None, None, 0, 0, 1e-12, None, None, None, None)
    assert rc == abi.ERR_ARG
    assert b"scratch limit" in hip.lib.sm2d_last_error()
