"""Throughput of the K1-K6 transport step on one MI355X.

Headline workload (value): 256 pencil-beam spots transported together (one batched launch
sequence per step, chunks of iterations replayed as HIP graphs) with the learned fine-transport
physics attached: energies U(70, 230) MeV, lateral offsets U(-40, 40) mm, sigma_x U(3, 6) mm,
1 % energy spread, on water 200 x 600 cells of 0.5 mm, 256 log-spaced energy bins
(0.1-250 MeV), 512 angular bins over [-pi/2, pi/2], 128 local bins per block.  Inputs are
synthetic (the seeded Gaussian source of the reference injector) and resident on the device
side before the timed region; a step ends with every spot's EdepC summed on the host.

A step is one sm2d_run_batch call over all 256 spots.  Reported beside it:
  - the plain path (one spot at a time) over the same 256 spots, and the plain batch;
  - the summed dose of batched+learned against the plain path, and the fine-transport work
    (cell-bin updates and learned substeps) that shows where the learned physics acts;
  - the roofline of the dominant kernel (the coarse transport: 8 B of algorithmic HBM traffic
    per cell-bin update, one FP32 read of psi_in and one FP32 write of psi_out) from HIP events
    on the kernel's own dispatches, and its HBM traffic from rocprofv3 counters (profiles/);
  - the same batch on the reference runner's piecewise 1029-bin energy grid, on water and on
    the seeded slab phantom (400 x 1200 at 0.25 mm), where the fine transport and the learned
    physics run, with the plain path's seed-to-seed dose difference as its noise floor;
  - the 150 MeV single spot on 200 x 400 cells (plain path, the per-spot latency);
  - the CPU oracle (-O3 port of the reference step) on the host cores, on a sample of the spots.
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


def _log(msg: str) -> None:
    """Progress of the sections on stderr (the metric line alone goes to stdout)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def _sync():
    import torch
    torch.cuda.synchronize()


def dose_diff(test, ref) -> dict:
    """Dose difference of `test` against `ref` on the voxels above 10 % of ref's maximum: local
    (|d| / ref) and global (|d| / ref max) relative differences, and the total."""
    import numpy as np
    m = ref > 0.1 * ref.max()
    d = np.abs(test - ref)[m]
    loc = d / ref[m]
    glo = d / ref.max()
    return {"voxels": int(m.sum()), "total_rel": float(test.sum() / ref.sum() - 1.0),
            "local_mean": float(loc.mean()), "local_p95": float(np.percentile(loc, 95)),
            "local_max": float(loc.max()), "local_within_1pct": float((loc <= 0.01).mean()),
            "global_mean": float(glo.mean()), "global_p95": float(np.percentile(glo, 95)),
            "global_max": float(glo.max()), "global_within_1pct": float((glo <= 0.01).mean())}


def _audit(a) -> dict:
    return {"iterations": a.iterations, "terminated": a.terminated,
            "cell_bin_updates": a.cell_bin_updates, "fine_cell_bin_updates": a.fine_cell_bin_updates,
            "surrogate_active": a.surrogate_active, "surrogate_substeps": a.surrogate_substeps,
            "surrogate_fallback": a.surrogate_fallback, "max_cell_W_closure": a.max_cell_W_closure,
            "W_fail_iters": a.W_fail_iters, "max_W_error": a.max_W_error,
            "W_leftover": a.W_leftover, "E_dep_MeV": a.E_dep}


def plain_spots(hip, prob, spots, tp):
    """One spot at a time through sm2d_run_spot (the reference's call pattern); returns the
    summed EdepC in spot order, the seconds and the cell-bin updates."""
    import numpy as np
    from sm_2d_amd import problems
    from sm_2d_amd.engine import Engine
    one = problems.setup_engine(hip, prob, Engine)
    one.run_spot(spots[0], tp)
    _sync()
    esum = np.zeros((prob.grid["Nz"], prob.grid["Nx"]))
    upd = 0
    t0 = time.perf_counter()
    for sp in spots:
        e, a = one.run_spot(sp, tp)
        esum += e
        upd += a.cell_bin_updates
    _sync()
    dt = time.perf_counter() - t0
    one.close()
    return esum, dt, upd


def roofline_from_timing(tm, traffic):
    """Roofline of the coarse transport from a profiled (eagerly launched, event-timed) run."""
    achieved = 8.0 * tm.coarse_updates / (tm.coarse_transport_ms * 1e-3) / 1e9 if tm.coarse_transport_ms else 0.0
    launches = max(1, tm.transport_launches // 2)
    return {"bound": "hbm", "kernel": "k_transport<Layout<8,4,4,1>, coarse, tuned waves> (all coarse launches)",
            "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic, "algorithmic_bytes_per_update": 8, "launches": launches,
            "avg_launch_us": tm.coarse_transport_ms * 1e3 / launches,
            "updates_per_launch": tm.coarse_updates / launches,
            "algorithmic_bytes_per_launch": 8.0 * tm.coarse_updates / launches}


def headline(hip, mlp, args) -> dict:
    """The 256-spot batch with the learned physics attached (the metric's workload)."""
    from sm_2d_amd import problems
    from sm_2d_amd.engine import Engine
    n = args.batch_spots
    prob = problems.config4(n_spots=n)
    spots = [prob.make_spot(i) for i in range(n)]
    tp = prob.make_transport()
    eng = problems.setup_engine(hip, prob, Engine, n_spots=n)
    plain_batch, _, ap = eng.run_batch(spots, tp)  # plain batch (also captures the chunk graphs)
    if mlp is not None:
        eng.surrogate_attach(mlp, 64)
        table = eng.surrogate_stats()
    for _ in range(args.warmup):
        eng.run_batch(spots, tp)
    _sync()
    t0 = time.perf_counter()
    upd = 0
    for _ in range(args.steps):
        esum, _, a = eng.run_batch(spots, tp)
        upd += a.cell_bin_updates
    _sync()
    dt = time.perf_counter() - t0
    # second pass, same batch, eagerly launched with HIP events on the transport dispatches
    eng.set_profiling(True)
    eng.run_batch(spots, tp)
    tm = eng.timing()
    eng.set_profiling(False)
    tuner = eng.tuner_report()
    eng.close()
    out = {"value": upd / dt, "spots_per_s": n * args.steps / dt, "ms_per_step": dt / args.steps * 1e3,
           "ms_per_spot": dt / args.steps / n * 1e3, "audit": _audit(a),
           "roofline": roofline_from_timing(tm, _profiles_traffic("transport_batch256.txt")),
           "stage_ms_per_step": {k: getattr(tm, k) for k in ("mask_ms", "transport_ms", "gather_ms", "audit_ms")},
           "transport_tuner": tuner,
           "learned_physics_attached": mlp is not None,
           "dose_vs_plain_batch": dose_diff(esum, plain_batch), "plain_batch_audit": _audit(ap)}
    if mlp is not None:
        out["learned_table"] = {"nodes_checked": table[0], "nodes_rejected": table[1]}
    if not args.no_plain:
        _log("plain path over all spots")
        ep, dtp, updp = plain_spots(hip, prob, spots, tp)
        out["plain"] = {"spots": n, "seconds": dtp, "spots_per_s": n / dtp, "updates_per_s": updp / dtp}
        out["speedup_spots_per_s_vs_plain"] = out["spots_per_s"] / out["plain"]["spots_per_s"]
        out["dose_vs_plain"] = dose_diff(esum, ep)
        out["plain_equals_batch_bitwise"] = bool((ep == plain_batch).all())
    return out


def fine_grid_variant(hip, mlp, name: str, n: int, plain_sample: int, slabs: bool) -> dict:
    """The batch on the reference runner's piecewise 1029-bin energy grid, where components slow
    down into the fine transport: plain batch, the same with the learned physics, the plain
    batch with other source seeds (the dose noise floor), and a sample of plain single spots."""
    from sm_2d_amd import problems
    from sm_2d_amd.engine import Engine
    prob = problems.config5(n_spots=n) if slabs else problems.config4_fine_grid(n_spots=n)
    spots = [prob.make_spot(i) for i in range(n)]
    reseeded = [prob.make_spot(i) for i in range(n)]
    for sp in reseeded:
        sp.random_seed += 7919
    tp = prob.make_transport()
    eng = problems.setup_engine(hip, prob, Engine, n_spots=n)
    eng.run_batch(spots, tp)
    _sync()
    t0 = time.perf_counter()
    e_plain, _, a_plain = eng.run_batch(spots, tp)
    _sync()
    dt_plain = time.perf_counter() - t0
    e_seed, _, _ = eng.run_batch(reseeded, tp)
    out = {"workload": prob.note, "spots": n, "plain_batch_s": dt_plain,
           "plain_batch_spots_per_s": n / dt_plain, "plain_batch_audit": _audit(a_plain),
           "plain_reseeded_vs_plain": dose_diff(e_seed, e_plain)}
    if mlp is not None:
        eng.surrogate_attach(mlp, 64)
        st = eng.surrogate_stats()
        eng.run_batch(spots, tp)
        _sync()
        t0 = time.perf_counter()
        e_l, _, a_l = eng.run_batch(spots, tp)
        _sync()
        dt_l = time.perf_counter() - t0
        out.update({"learned_batch_s": dt_l, "learned_batch_spots_per_s": n / dt_l,
                    "learned_audit": _audit(a_l), "learned_vs_plain": dose_diff(e_l, e_plain),
                    "learned_table": {"nodes_checked": st[0], "nodes_rejected": st[1]},
                    "learned_over_plain_batch_speed": dt_plain / dt_l})
        # the same network evaluated inside the fine transport for every substep (no table)
        eng.surrogate_enable(2)
        eng.run_batch(spots, tp)
        _sync()
        t0 = time.perf_counter()
        e_i, _, a_i = eng.run_batch(spots, tp)
        _sync()
        dt_i = time.perf_counter() - t0
        out["inferred"] = {"batch_s": dt_i, "spots_per_s": n / dt_i, "audit": _audit(a_i),
                           "vs_plain": dose_diff(e_i, e_plain), "vs_table": dose_diff(e_i, e_l),
                           "table_over_inferred_speed": dt_i / dt_l}
    eng.close()
    if plain_sample:
        _, dts, upds = plain_spots(hip, prob, spots[:plain_sample], tp)
        out["plain_single"] = {"spots": plain_sample, "seconds": dts, "spots_per_s": plain_sample / dts}
        out["plain_batch_over_single"] = out["plain_batch_spots_per_s"] / out["plain_single"]["spots_per_s"]
        if mlp is not None:
            out["learned_batch_over_plain_single"] = out["learned_batch_spots_per_s"] / out["plain_single"]["spots_per_s"]
    return out


def single_spot(hip, steps: int) -> dict:
    """The 150 MeV spot on 200 x 400 cells through the plain path (per-spot latency)."""
    from sm_2d_amd import problems
    from sm_2d_amd.engine import Engine
    prob = problems.config2()
    eng = problems.setup_engine(hip, prob, Engine)
    sp, tp = prob.make_spot(), prob.make_transport()
    eng.run_spot(sp, tp)
    _sync()
    t0 = time.perf_counter()
    upd = 0
    for _ in range(steps):
        _, a = eng.run_spot(sp, tp)
        upd += a.cell_bin_updates
    _sync()
    dt = time.perf_counter() - t0
    eng.set_profiling(True)
    for _ in range(steps):
        eng.run_spot(sp, tp)
    tm = eng.timing()
    eng.close()
    g_ms, g_val = tm.coarse_gather_ms, tm.coarse_gather_values
    g_ach = 8.0 * g_val / (g_ms * 1e-3) / 1e9 if g_ms > 0 else 0.0
    launches = max(1, tm.iterations)
    return {"workload": "plain K1-K6, one 150 MeV spot (sigma_x 3.8 mm, 1% sigma_E, 2 mrad), water "
                        "200x400 @ 0.5 mm, 256 log E x 512 theta, 128 local bins",
            "ms_per_spot": dt / steps * 1e3, "updates_per_s": upd / dt, "iterations": a.iterations,
            "cell_bin_updates": a.cell_bin_updates,
            "roofline": roofline_from_timing(tm, _profiles_traffic("transport_config2.txt")),
            "roofline_k4": {"bound": "hbm", "kernel": "k_gather (coarse phase)", "achieved": g_ach,
                            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": g_ach / HBM_PEAK_GBS,
                            "traffic": _profiles_traffic("gather_config2.txt"),
                            "avg_launch_us": g_ms * 1e3 / launches, "values_per_launch": g_val / launches},
            "stage_ms_per_spot": {k: getattr(tm, k) / steps for k in ("mask_ms", "transport_ms", "gather_ms", "audit_ms")}}


def learned_physics(hip) -> tuple:
    """Trains the network on the GPU (forward and backward kernels on the matrix cores) on
    samples of the plain physics; returns it with the training figures."""
    from sm_2d_amd import problems
    from sm_2d_amd.engine import Engine, Mlp
    prob = problems.reference_grid_case(Nx=40, Nz=80, E0=60.0, sigma_x=2.0, n_samples=600,
                                        max_iterations=400)
    eng = problems.setup_engine(hip, prob, Engine)
    m = Mlp(hip, seed=11)
    cfg = dict(n_samples=1 << 22, batch=4096, steps=4000, lr=2e-3, w_mom=4.0, w_att=4.0)
    _sync()
    t0 = time.perf_counter()
    hist = m.train(eng, n_hist=4, **cfg)
    _sync()
    t_train = time.perf_counter() - t0
    ev = m.evaluate(eng, 65536, seed=123)
    eng.close()
    flops = 3 * 75008 * cfg["batch"] * cfg["steps"]
    return m, {"samples": cfg["n_samples"], "steps": cfg["steps"], "batch": cfg["batch"],
               "seconds": t_train, "loss_first_last": [float(hist[0]), float(hist[-1])],
               "mfma_tflops_achieved": flops / t_train / 1e12,
               "heldout_rms_ln_sigmaE_sigmaTheta_lnA": [float(v) for v in ev[:3]]}


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(budget_s: float, n_spots: int) -> dict:
    """The oracle (a C++ port of the reference step, -O3 -march=x86-64-v3) on the host cores:
    spots of the headline workload run to completion, one thread, then spots on every core the
    process may use (at most 16, the host share of one GPU), one context per thread.  Context
    set-up is outside the timed spans."""
    import threading
    from concurrent.futures import ThreadPoolExecutor
    from oracle import pyoracle
    from sm_2d_amd import problems

    lib = pyoracle.load(fast=True)
    prob = problems.config4(n_spots=n_spots)
    tp = prob.make_transport()

    def run(idx, barrier=None):
        eng = problems.setup_engine(lib, prob, pyoracle.OracleEngine)
        if barrier is not None:
            barrier.wait()
        t0 = time.perf_counter()
        upd = 0
        for i in idx:
            _, a = eng.run_spot(prob.make_spot(i), tp)
            upd += a.cell_bin_updates
        t1 = time.perf_counter()
        eng.close()
        return upd, t0, t1

    u, t0, t1 = run([0])
    t_one = t1 - t0
    k1 = max(1, min(n_spots, int(budget_s / 2 / max(t_one, 1e-3))))
    u1, dt1 = u, t_one
    if k1 > 1:
        u, t0, t1 = run(list(range(1, k1)))
        u1, dt1 = u1 + u, dt1 + (t1 - t0)
    cores = min(16, len(os.sched_getaffinity(0)))
    per = max(1, min(int(budget_s / 2 / max(t_one, 1e-3)), n_spots // cores))
    groups = [[(t * per + j) % n_spots for j in range(per)] for t in range(cores)]
    bar = threading.Barrier(cores)
    with ThreadPoolExecutor(cores) as ex:
        res = list(ex.map(lambda g: run(g, bar), groups))
    un = sum(r[0] for r in res)
    dtn = max(r[2] for r in res) - min(r[1] for r in res)
    return {"value": un / dtn, "unit": "cell-bin updates/s", "cores": cores, "kind": "port",
            "spots_per_s": cores * per / dtn,
            "one_core": {"value": u1 / dt1, "spots_per_s": k1 / dt1, "spots": k1, "seconds": dt1},
            "cpu_model": _cpu_model(), "nproc": os.cpu_count(),
            "sample": f"{cores} threads x {per} spots of the 256-spot workload run to completion "
                      f"({un} cell-bin updates in {dtn:.1f} s); one thread: spots 0..{k1 - 1} "
                      f"({u1} updates in {dt1:.1f} s); oracle -O3 -march=x86-64-v3, one context per thread"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch-spots", type=int, default=256)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=24.0)
    ap.add_argument("--no-surrogate", action="store_true", help="headline without the learned physics")
    ap.add_argument("--no-plain", action="store_true", help="skip the one-spot-at-a-time pass")
    ap.add_argument("--no-variants", action="store_true", help="skip the 1029-bin-grid batches")
    ap.add_argument("--variant-spots", type=int, default=256)
    ap.add_argument("--variant-plain-sample", type=int, default=16)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 256 MiB device to device three times (HBM counter calibration)")
    ap.add_argument("--only", choices=("headline", "single", "fine_grid", "slabs"),
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
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (spots shard across "
              f"GPUs by a longest-first schedule with one all-reduce of the summed dose)", file=sys.stderr)
        return 2

    import torch
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.zeros(1, device="cuda:0")
    from sm_2d_amd import runtime
    hip = runtime.load_library()

    if args.only:
        _log(f"section {args.only}")
        mlp = None
        if args.only != "single" and not args.no_surrogate:
            mlp, _ = learned_physics(hip)
        if args.only == "headline":
            args.no_plain = True
            sec = headline(hip, mlp, args)
        elif args.only == "single":
            sec = single_spot(hip, args.steps)
        else:
            sec = fine_grid_variant(hip, mlp, args.only, args.variant_spots, 0, args.only == "slabs")
        print(json.dumps({"only": args.only, "section": sec}), flush=True)
        return 0

    _log("training the learned physics")
    mlp, train = (None, None) if args.no_surrogate else learned_physics(hip)
    _log("headline batch")
    h = headline(hip, mlp, args)
    out = {
        "metric": "cell-bin updates/s + spots/s (batched+surrogate vs plain) at <=1% dose diff",
        "value": h["value"],
        "unit": "cell-bin updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": h["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Gaussian sources, water phantom)",
        "config": {
            "workload": f"{args.batch_spots} pencil beams batched, learned fine-transport physics "
                        f"attached: E0 U(70,230) MeV, x0 U(-40,40) mm, sigma_x U(3,6) mm, 1% sigma_E, "
                        f"water 200x600 cells @ 0.5 mm, 256 log E bins (0.1-250 MeV) x 512 theta bins "
                        f"over [-pi/2, pi/2], 128 local bins (8 theta x 4 E x 4 x_sub)",
            "spots_per_step": args.batch_spots,
            "path": "batched+learned" if mlp is not None else "batched",
        },
        "spots_per_s": h["spots_per_s"],
        "ms_per_spot": h["ms_per_spot"],
        "roofline": h["roofline"],
        "headline": h,
    }
    if train is not None:
        out["learned_training"] = train
    if not args.no_variants:
        _log("fine-grid variants")
        out["fine_grid_water"] = fine_grid_variant(hip, mlp, "fine_grid", args.variant_spots,
                                                   args.variant_plain_sample, slabs=False)
        _log("fine-grid slab variant")
        out["fine_grid_slabs"] = fine_grid_variant(hip, mlp, "slabs", args.variant_spots,
                                                   args.variant_plain_sample, slabs=True)
    _log("single spot")
    out["single_spot"] = single_spot(hip, args.steps)
    if mlp is not None:
        mlp.close()
    if not args.no_cpu_baseline:
        _log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(args.cpu_budget, args.batch_spots)
        out["gpu_over_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
