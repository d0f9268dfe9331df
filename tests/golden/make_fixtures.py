"""Copies the reference's published result tables into this directory as test
fixtures (inputs and expected outputs only):

* convergence_results.csv  <- data/convergence/results.csv (MMS errors at
  T = 0.1 for q = 1..6, mixed/matched, upwind/central/alternating, N = 5..25;
  CG mass solver, absorbing_2beta with sigma = 0, dt = h/(10 q^2))
* alpha_limit_order.csv    <- data/alpha_limit/order.csv (difference norms of
  the alpha -> 0 study, experiments/alpha_limit/run.jl)
* steepening_<case>_series.csv <- data/steepening/<case>/series.csv (t, energy,
  max|v_h|, min(1 - 2 beta v_h) every 32 steps; experiments/steepening/run.jl,
  cases nl, lin, nl_dbc), whole
* steepening_<case>_recv.csv   <- data/steepening/<case>/recv.csv, every 8th
  step (the rows after steps 8, 16, ..., 18432), prefixed by the step number
* steepening_nl_grid.csv       <- data/steepening/nl/grid.csv, whole
* steepening_nl_snap_0.80.csv  <- data/steepening/nl/snap_0.80.csv, rows and
  columns 3, 11, 19, ... (every 8th, from the 4th), with the 0-based row index
  first and a header line of the 0-based column indices
The values are copied as text, digit for digit.

Run once in the build container: python3 tests/golden/make_fixtures.py
"""
import os
import shutil

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
for src, dst in (("data/convergence/results.csv", "convergence_results.csv"),
                 ("data/alpha_limit/order.csv", "alpha_limit_order.csv")):
    shutil.copyfile(os.path.join(REF, src), os.path.join(HERE, dst))
    print("wrote", dst)

STEEP = "data/steepening"
for case in ("nl", "lin", "nl_dbc"):
    shutil.copyfile(os.path.join(REF, STEEP, case, "series.csv"), os.path.join(HERE, f"steepening_{case}_series.csv"))
    print("wrote", f"steepening_{case}_series.csv")
    with open(os.path.join(REF, STEEP, case, "recv.csv")) as f:
        rows = f.read().splitlines()
    with open(os.path.join(HERE, f"steepening_{case}_recv.csv"), "w") as f:
        for i in range(7, len(rows), 8):
            f.write(f"{i + 1},{rows[i]}\n")
    print("wrote", f"steepening_{case}_recv.csv")
shutil.copyfile(os.path.join(REF, STEEP, "nl", "grid.csv"), os.path.join(HERE, "steepening_nl_grid.csv"))
print("wrote steepening_nl_grid.csv")
with open(os.path.join(REF, STEEP, "nl", "snap_0.80.csv")) as f:
    rows = f.read().splitlines()
sel = list(range(3, len(rows), 8))
with open(os.path.join(HERE, "steepening_nl_snap_0.80.csv"), "w") as f:
    f.write("row," + ",".join(str(c) for c in sel) + "\n")
    for r in sel:
        vals = rows[r].split(",")
        f.write(f"{r}," + ",".join(vals[c] for c in sel) + "\n")
print("wrote steepening_nl_snap_0.80.csv")
