"""Short text summaries of rocprofv3 output (the files kept under profiles/).

    python -m sm_2d_amd.profile_summary stats DIR --cmd "COMMAND"
        per-kernel calls / total / average / share from DIR/run_kernel_stats.csv
        (rocprofv3 --kernel-trace --stats --output-format csv -d DIR -o run -- ...)

    python -m sm_2d_amd.profile_summary traffic --fetch DIR --write DIR --cal-fetch DIR
        --cal-write DIR --kernel REGEX --algorithmic-bytes B [--copy-bytes N]
        [--stride S --offset O]
        HBM bytes per dispatch of the kernels matching REGEX from two counter runs
        (rocprofv3 --pmc FETCH_SIZE, then --pmc WRITE_SIZE), each scaled by what the same
        counter reads for a plain device copy of N bytes (bench.py --copy-calibration).
        --stride/--offset keep every S-th matching dispatch from the O-th on, in dispatch
        order (a kernel launched once per phase: stride 2 separates the phases).
"""
from __future__ import annotations

import argparse
import csv
import os
import re
import sys


def _short(name: str) -> str:
    name = name.replace("(anonymous namespace)::", "")
    if name.startswith("void "):
        name = name[5:]
    depth, out = 0, []
    for ch in name:  # drop the argument list, keep template arguments
        if ch == "(" and depth == 0:
            break
        depth += ch == "<"
        depth -= ch == ">"
        out.append(ch)
    return "".join(out)


def stats(d: str, cmd: str) -> str:
    rows = list(csv.DictReader(open(os.path.join(d, "run_kernel_stats.csv"))))
    lines = ["# rocprofv3 kernel statistics", f"# command: {cmd}",
             f"{'kernel':66s} {'calls':>7s} {'total_ms':>10s} {'avg_us':>9s} {'pct':>6s}"]
    for r in rows:
        lines.append(f"{_short(r['Name'])[:66]:66s} {int(r['Calls']):7d} "
                     f"{float(r['TotalDurationNs']) / 1e6:10.3f} {float(r['AverageNs']) / 1e3:9.2f} "
                     f"{float(r['Percentage']):6.2f}")
    return "\n".join(lines)


def _counter(d: str, regex: str, stride: int = 1, offset: int = 0) -> list[float]:
    pat = re.compile(regex)
    rows = []
    for r in csv.DictReader(open(os.path.join(d, "run_counter_collection.csv"))):
        if pat.search(r["Kernel_Name"]):
            rows.append((int(r["Dispatch_Id"]), float(r["Counter_Value"])))
    rows.sort()
    return [v for _, v in rows[offset::stride]]


def traffic(a) -> str:
    copy_kib = a.copy_bytes / 1024.0
    cf = _counter(a.cal_fetch, a.copy_kernel)
    cw = _counter(a.cal_write, a.copy_kernel)
    if not cf or not cw:
        raise SystemExit("no copy dispatch in the calibration runs")
    f_fetch = copy_kib / max(cf)
    f_write = copy_kib / max(cw)
    fk = _counter(a.fetch, a.kernel, a.stride, a.offset)
    wk = _counter(a.write, a.kernel, a.stride, a.offset)
    if not fk or not wk:
        raise SystemExit(f"no dispatch of {a.kernel!r} in the counter runs")
    fetch = sum(fk) / len(fk)
    write = sum(wk) / len(wk)
    bytes_per = (fetch * f_fetch + write * f_write) * 1024.0
    return "\n".join([
        f"# HBM traffic of {a.kernel!r} dispatches",
        f"#   calibration: a {a.copy_bytes} B device copy reads FETCH_SIZE {max(cf):.0f} KiB "
        f"(factor {f_fetch:.3f}) and WRITE_SIZE {max(cw):.0f} KiB (factor {f_write:.3f})",
        f"dispatches: {len(fk)}",
        f"fetch_kib_per_dispatch_raw: {fetch:.1f}",
        f"write_kib_per_dispatch_raw: {write:.1f}",
        f"traffic_bytes_per_launch: {bytes_per:.0f}",
        f"algorithmic_bytes_per_launch: {a.algorithmic_bytes:.0f}",
        f"traffic_over_algorithmic: {bytes_per / max(a.algorithmic_bytes, 1.0):.2f}",
    ])


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="what", required=True)
    s = sub.add_parser("stats")
    s.add_argument("dir")
    s.add_argument("--cmd", required=True)
    t = sub.add_parser("traffic")
    for k in ("fetch", "write", "cal_fetch", "cal_write"):
        t.add_argument("--" + k.replace("_", "-"), dest=k, required=True)
    t.add_argument("--kernel", required=True)
    t.add_argument("--algorithmic-bytes", type=float, required=True)
    t.add_argument("--copy-bytes", type=int, default=256 * 1024 * 1024)
    t.add_argument("--copy-kernel", default="copyBuffer")
    t.add_argument("--stride", type=int, default=1)
    t.add_argument("--offset", type=int, default=0)
    a = ap.parse_args(argv)
    print(stats(a.dir, a.cmd) if a.what == "stats" else traffic(a))
    return 0


if __name__ == "__main__":
    sys.exit(main())
