"""CPU oracle for the EDG Westervelt RHS and RK4 — test infrastructure only.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this package, and only as the checker; the product path in
``edg_westervelt_amd`` never does.

* ``oracle.batched`` — ctypes front end of ``wedg_oracle.c``, the plain-C FP64
  restatement of the reference's batched algorithm (src/rhs.jl, src/solver.jl).
* ``oracle.dense``   — numpy restatement of the element-by-element dense
  formulation of test/reference_rhs.jl (dense solve of the mass systems), a
  structurally different second restatement used to pin the first.
* ``oracle.norms``   — loop restatements of ``_mms_errors`` and
  ``difference_norms`` (src/solver.jl:772-836).

Pinning (each function is checked by at least one of these in tests/):
the dense restatement at <= 1e-12 over the reference's own test matrix
(test/runtests.jl:189-293, 716-732); the initialization known answer
(test/runtests.jl:168-184, 340-357); the closed forms of the boundary fluxes
(test/runtests.jl:295-317); dense mass solves (test/runtests.jl:418-444);
locality and superposition (test/runtests.jl:319-338); the golden errors of
data/convergence/results.csv and test/runtests.jl:208-217.
"""
