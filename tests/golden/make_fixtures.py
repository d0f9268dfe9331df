"""Copies the reference's published result tables into this directory as test
fixtures (inputs and expected outputs only):

* convergence_results.csv  <- data/convergence/results.csv (MMS errors at
  T = 0.1 for q = 1..6, mixed/matched, upwind/central/alternating, N = 5..25;
  CG mass solver, absorbing_2beta with sigma = 0, dt = h/(10 q^2))
* alpha_limit_order.csv    <- data/alpha_limit/order.csv (difference norms of
  the alpha -> 0 study, experiments/alpha_limit/run.jl)

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
