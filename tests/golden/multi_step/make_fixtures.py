"""Writes the multi-iteration fixtures of tests/test_oracle_golden.py and tests/test_io.py from
the reference's committed dumps:
  - results/debug_ps_raw_iter_{00_initial,01..10_after_K4}.csv -> iter_NN.csv: the columns the
    phase-space loader needs (cell, slot, bid, lidx, weight), with the weight text kept as printed;
  - results/debug_cells_iter_{01..10}_after_K4.csv -> cells_NN.csv and
    results/one_step_case_{A..D}/debug_cells_iter_01_after_K4.csv -> cells_case_X.csv: the
    per-cell summaries of the same phase space, copied unchanged.

Run in a container that holds the reference checkout (the GPU box does not):
    python tests/golden/multi_step/make_fixtures.py [/root/reference]
"""
import csv
import os
import shutil
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
names = ["00_initial"] + [f"{i:02d}_after_K4" for i in range(1, 11)]
for n in names:
    src = os.path.join(REF, "results", f"debug_ps_raw_iter_{n}.csv")
    with open(src) as f, open(os.path.join(HERE, f"iter_{n[:2]}.csv"), "w") as g:
        w = csv.writer(g, lineterminator="\n")
        w.writerow(["cell", "slot", "bid", "lidx", "weight"])
        for r in csv.DictReader(f):
            w.writerow([r["cell"], r["slot"], r["bid"], r["lidx"], r["weight"]])
for i in range(1, 11):
    shutil.copyfile(os.path.join(REF, "results", f"debug_cells_iter_{i:02d}_after_K4.csv"),
                    os.path.join(HERE, f"cells_{i:02d}.csv"))
for c in "ABCD":
    shutil.copyfile(os.path.join(REF, "results", f"one_step_case_{c}", "debug_cells_iter_01_after_K4.csv"),
                    os.path.join(HERE, f"cells_case_{c}.csv"))
print("wrote", len(names) + 14, "files to", HERE)
