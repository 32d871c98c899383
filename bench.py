"""Throughput of the K1-K6 transport step on one MI355X.

A step is one complete pencil-beam spot of the benchmark workload (injection, every K1..K6
iteration until the phase space is empty, EdepC read-back) through the HIP library.  The
workload is the 150 MeV water spot on 200 x 400 cells of 0.5 mm with 256 log-spaced energy
bins (0.1-250 MeV), 512 angular bins over [-pi/2, pi/2] and 128 local bins per block;
inputs are synthetic (the seeded Gaussian source of the reference injector).

Reported: cell-bin updates/s (value), spots/s, ms per spot, the roofline of the dominant
kernel (transport, HBM-bound: 8 B per cell-bin update = one FP32 read of psi_in + one FP32
write of psi_out, divided by the kernel time from HIP events on the library's stream), the same
for the K4 gather (8 B per non-zero bucket value moved = one FP32 read of the bucket + one FP32
write of psi_out) and the CPU oracle timed on the same host on a bounded sample of the same
workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def _profiles_traffic(name: str):
    """HBM bytes per launch of a kernel measured by rocprofv3 counters, when its summary exists."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            try:
                return float(line.split(":", 1)[1].strip())
            except ValueError:
                return None
    return None


def cpu_baseline(prob, budget_s: float = 20.0) -> dict:
    """The oracle (a serial C++ port of the reference step) on the same spot.  The spot is run
    with a bounded iteration count chosen from a short probe so the sample takes ~budget_s."""
    from oracle import pyoracle
    from sm_2d_amd import problems

    orc = pyoracle.load()
    eng = problems.setup_engine(orc, prob, pyoracle.OracleEngine)
    sp = prob.make_spot()
    probe_iters = 20
    t0 = time.perf_counter()
    _, a = eng.run_spot(sp, prob.make_transport(max_iterations=probe_iters))
    dt = time.perf_counter() - t0
    full = prob.transport["max_iterations"]
    iters = full if dt <= 0 else int(min(full, max(probe_iters, probe_iters * budget_s / dt)))
    t0 = time.perf_counter()
    _, a = eng.run_spot(sp, prob.make_transport(max_iterations=iters))
    dt = time.perf_counter() - t0
    eng.close()
    return {"value": a.cell_bin_updates / dt, "unit": "cell-bin updates/s", "cores": 1,
            "kind": "port",
            "sample": f"first {a.iterations} iterations of the same spot ({a.cell_bin_updates} "
                      f"cell-bin updates in {dt:.2f} s, serial C++ oracle, -O2)"}


def batched_section(hip, n_spots: int, plain_sample: int, mlp=None) -> dict:
    """The 256-spot workload (pencil beams of 70-230 MeV over +-40 mm with sigma 3-6 mm on water
    200 x 600 at 0.5 mm): one batched launch sequence for all spots against the plain path
    (one spot at a time) timed on the first `plain_sample` spots of the same list.  With a
    trained network `mlp`, the same batch again with the learned fine-transport physics
    attached: spots/s against the plain path, and its summed dose against the plain batch's on
    voxels above 10 % of the maximum."""
    import torch
    from sm_2d_amd import problems
    from sm_2d_amd.engine import Engine

    prob = problems.config4(n_spots=n_spots)
    spots = [prob.make_spot(i) for i in range(n_spots)]
    tp = prob.make_transport()
    one = problems.setup_engine(hip, prob, Engine)
    one.run_spot(spots[0], tp)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for sp in spots[:plain_sample]:
        one.run_spot(sp, tp)
    torch.cuda.synchronize()
    dt_plain = time.perf_counter() - t0
    one.close()
    eng = problems.setup_engine(hip, prob, Engine, n_spots=n_spots)
    t0 = time.perf_counter()
    eng.run_batch(spots[:min(8, n_spots)], tp)
    t_warm = time.perf_counter() - t0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    esum, _, a = eng.run_batch(spots, tp)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    eng.set_profiling(True)  # stage times from a second, eagerly launched run
    eng.run_batch(spots, tp)
    tm = eng.timing()
    tuner = eng.tuner_report()
    plain_sps = plain_sample / dt_plain
    batch_sps = n_spots / dt
    learned = None
    if mlp is not None:
        import numpy as np
        eng.set_profiling(False)
        eng.surrogate_attach(mlp, 64)
        eng.run_batch(spots[:min(8, n_spots)], tp)  # captures the chunk graphs again
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        es, _, al = eng.run_batch(spots, tp)
        torch.cuda.synchronize()
        dt_l = time.perf_counter() - t0
        m10 = esum > 0.1 * esum.max()
        rel = np.abs(es - esum)[m10] / esum[m10]
        learned = {"batched_learned_spots_per_s": n_spots / dt_l, "batched_learned_s": dt_l,
                   "speedup_vs_plain_spots_per_s": n_spots / dt_l / plain_sps,
                   "dose_vs_plain_batch": {"total_rel": float(es.sum() / esum.sum() - 1.0),
                                           "voxel_gt10pct_mean_rel": float(rel.mean()),
                                           "voxel_gt10pct_p95_rel": float(np.percentile(rel, 95)),
                                           "voxel_gt10pct_max_rel": float(rel.max()),
                                           "voxels_within_1pct": float((rel <= 0.01).mean())},
                   "audit": {"surrogate_active": al.surrogate_active, "fallback": al.surrogate_fallback,
                             "max_cell_W_closure": al.max_cell_W_closure, "W_fail_iters": al.W_fail_iters}}
    eng.close()
    return {"workload": f"{n_spots} pencil beams, E0 U(70,230) MeV, x0 U(-40,40) mm, sigma_x U(3,6) mm, "
                        "water 200x600 @ 0.5 mm, 256 log E x 512 theta, 128 local bins",
            "batched_spots_per_s": batch_sps, "batched_s": dt,
            "batched_updates_per_s": a.cell_bin_updates / dt,
            "batched_iterations": a.iterations, "batched_cell_bin_updates": a.cell_bin_updates,
            "plain_spots_per_s": plain_sps, "plain_sample_spots": plain_sample,
            "speedup_spots_per_s": batch_sps / plain_sps, "surrogate": False,
            "batched_learned": learned,
            "warm_batch_s": t_warm,
            "stage_ms": {k: getattr(tm, k) for k in ("mask_ms", "transport_ms", "gather_ms", "audit_ms")},
            "transport_tuner": tuner}


def surrogate_section(hip, keep: list | None = None) -> dict:
    """The learned fine-transport physics against the plain physics on a 60 MeV water spot whose
    whole track runs in the fine transport (where the learned widths act): one-time training and
    table build, ms per spot of both, and the dose comparison (total deposited energy, 4 x 4
    block means over blocks above 10 % of the maximum)."""
    import numpy as np
    import torch
    from sm_2d_amd import problems
    from sm_2d_amd.engine import Engine, Mlp
    prob = problems.reference_grid_case(Nx=40, Nz=80, E0=60.0, sigma_x=2.0, n_samples=600,
                                        max_iterations=400)
    prob.transport.update(E_fine_on=300.0, E_fine_off=300.0)
    eng = problems.setup_engine(hip, prob, Engine)
    sp, tp = prob.make_spot(), prob.make_transport()
    eng.run_spot(sp, tp)

    def timed(n=3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            e, a = eng.run_spot(sp, tp)
        torch.cuda.synchronize()
        return e, a, (time.perf_counter() - t0) / n * 1e3

    ep, ap, ms_plain = timed()
    m = Mlp(hip, seed=11)
    t0 = time.perf_counter()
    cfg = dict(n_samples=1 << 21, batch=4096, steps=4000, lr=2e-3, w_mom=4.0, w_att=4.0)
    hist = m.train(eng, n_hist=4, **cfg)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    ev = m.evaluate(eng, 65536, seed=123)
    t0 = time.perf_counter()
    eng.surrogate_attach(m, 64)
    t_attach = time.perf_counter() - t0
    st = eng.surrogate_stats()
    eng.run_spot(sp, tp)
    es, as_, ms_surr = timed()

    def blk(d, b=4):
        nz, nx = d.shape
        return d[: nz // b * b, : nx // b * b].reshape(nz // b, b, nx // b, b).sum((1, 3))

    bp, bs = blk(ep), blk(es)
    mk = bp > 0.1 * bp.max()
    rel = np.abs(bs - bp)[mk] / bp[mk]
    if keep is not None:
        keep.append(m)  # the trained network, for the batched section
    else:
        m.close()
    eng.close()
    flops_train = 3 * 75008 * cfg["batch"] * cfg["steps"]
    return {"workload": "60 MeV water spot, 40 x 80 cells @ 1 mm (reference grid), all energies in "
                        "the fine transport, 600 source samples",
            "train": {"samples": cfg["n_samples"], "steps": cfg["steps"], "batch": cfg["batch"],
                      "seconds": t_train, "loss_first_last": [float(hist[0]), float(hist[-1])],
                      "mfma_tflops_achieved": flops_train / t_train / 1e12,
                      "heldout_rms_ln_sigmaE_sigmaTheta_lnA": [float(v) for v in ev[:3]]},
            "table": {"build_s": t_attach, "nodes_checked": st[0], "nodes_rejected": st[1],
                      "max_dev": st[2], "rms_dev": st[3]},
            "ms_per_spot_plain": ms_plain, "ms_per_spot_learned": ms_surr,
            "speedup": ms_plain / ms_surr,
            "dose": {"total_rel": float(es.sum() / ep.sum() - 1.0),
                     "block4_mean_rel": float(rel.mean()),
                     "block4_p95_rel": float(np.percentile(rel, 95))},
            "audit": {"surrogate_active": as_.surrogate_active, "fallback": as_.surrogate_fallback,
                      "max_cell_W_closure": as_.max_cell_W_closure, "W_fail_iters": as_.W_fail_iters}}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--batch-spots", type=int, default=256)
    ap.add_argument("--plain-sample", type=int, default=8)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 256 MiB device to device three times (HBM counter calibration)")
    ap.add_argument("--only", choices=("batch", "learned"),
                    help="print only that section (for profiling one workload); not the metric line")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = 64 * 1024 * 1024
        src = torch.ones(n, device="cuda", dtype=torch.float32)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(f"copied {4 * n} bytes x 3", flush=True)
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (spots shard "
              f"across GPUs by a longest-first schedule, see DESIGN.md)", file=sys.stderr)
        return 2

    import torch
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.zeros(1, device="cuda:0")
    from sm_2d_amd import problems, runtime
    from sm_2d_amd.engine import Engine

    if args.only:
        hip = runtime.load_library()
        sec = (batched_section(hip, args.batch_spots, args.plain_sample) if args.only == "batch"
               else surrogate_section(hip))
        print(json.dumps({"only": args.only, "section": sec}), flush=True)
        return 0

    prob = problems.config2()
    hip = runtime.load_library()
    eng = problems.setup_engine(hip, prob, Engine)
    sp, tp = prob.make_spot(), prob.make_transport()
    for _ in range(args.warmup):
        eng.run_spot(sp, tp)
    # Timed region: the production launch path (chunks of iterations replayed as HIP graphs).
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    updates = 0
    audits = []
    for _ in range(args.steps):
        _, a = eng.run_spot(sp, tp)
        updates += a.cell_bin_updates
        audits.append(a)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    # Second timed region, same steps, for the per-kernel times: HIP events around every stage of
    # every iteration on the library's stream (graph capture drops event records, so these chunks
    # are launched eagerly; the kernels themselves are the same).
    eng.set_profiling(True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        eng.run_spot(sp, tp)
    torch.cuda.synchronize()
    dt_prof = time.perf_counter() - t1
    tm = eng.timing()
    tuner = eng.tuner_report()
    a = audits[-1]
    # dominant kernel: the coarse-phase transport launch (the fine phase of this workload is empty)
    dom_ms, dom_upd = tm.coarse_transport_ms, tm.coarse_updates
    if tm.fine_transport_ms > tm.coarse_transport_ms:
        dom_ms, dom_upd = tm.fine_transport_ms, tm.fine_updates
    achieved = 8.0 * dom_upd / (dom_ms * 1e-3) / 1e9 if dom_ms > 0 else 0.0
    traffic = _profiles_traffic("transport_config2.txt")
    # K4: the coarse-phase gather (one launch per iteration while profiling)
    g_ms, g_val = tm.coarse_gather_ms, tm.coarse_gather_values
    g_achieved = 8.0 * g_val / (g_ms * 1e-3) / 1e9 if g_ms > 0 else 0.0
    out = {
        "metric": "cell-bin updates/s + spots/s (batched+surrogate vs plain) at <=1% dose diff",
        "value": updates / dt,
        "unit": "cell-bin updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Gaussian source, water phantom)",
        "config": {
            "workload": "plain K1-K6, one 150 MeV spot (sigma_x 3.8 mm, 1% sigma_E, 2 mrad), "
                        "water 200x400 cells @ 0.5 mm, 256 log E bins x 512 theta bins, "
                        "128 local bins (8 theta x 4 E x 4 x_sub)",
            "path": "plain",
            "spots_per_step": 1,
            "iterations_per_spot": a.iterations,
            "cell_bin_updates_per_spot": a.cell_bin_updates,
        },
        "spots_per_s": args.steps / dt,
        "ms_per_spot": dt / args.steps * 1e3,
        "ms_per_spot_eager_with_events": dt_prof / args.steps * 1e3,
        "stage_ms_per_spot": {k: getattr(tm, k) / args.steps for k in
                              ("mask_ms", "transport_ms", "gather_ms", "audit_ms")},
        "roofline": {
            "bound": "hbm",
            "kernel": "k_transport (coarse phase)" if dom_ms == tm.coarse_transport_ms else "k_transport (fine phase)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "algorithmic_bytes_per_update": 8,
            "launches": tm.iterations,
            "avg_launch_us": dom_ms * 1e3 / max(1, tm.iterations),
            "updates_per_launch": dom_upd / max(1, tm.iterations),
        },
        "roofline_k4": {
            "bound": "hbm",
            "kernel": "k_gather (coarse phase)",
            "achieved": g_achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": g_achieved / HBM_PEAK_GBS,
            "traffic": _profiles_traffic("gather_config2.txt"),
            "algorithmic_bytes_per_value": 8,
            "launches": tm.iterations,
            "avg_launch_us": g_ms * 1e3 / max(1, tm.iterations),
            "values_per_launch": g_val / max(1, tm.iterations),
        },
        "transport_tuner": tuner,
        "audit": {"max_cell_W_closure": a.max_cell_W_closure, "W_fail_iters": a.W_fail_iters,
                  "E_fail_iters": a.E_fail_iters, "terminated": a.terminated,
                  "E_boundary_MeV": a.E_boundary, "E_dep_MeV": a.E_dep},
    }
    eng.close()
    trained = []
    if not args.no_surrogate:
        out["learned_physics"] = surrogate_section(hip, keep=trained)
    if not args.no_batch:
        out["batched"] = batched_section(hip, args.batch_spots, args.plain_sample,
                                         mlp=trained[0] if trained else None)
    for m in trained:
        m.close()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(prob, args.cpu_budget)
        out["gpu_over_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
