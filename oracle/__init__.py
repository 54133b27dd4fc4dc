"""CPU oracle for the Ising hot path.

TEST INFRASTRUCTURE ONLY. Nothing in the product package may import, call, link or execute
anything under ``oracle/``: only the test suite, the smoke check and the CPU-baseline leg of the
benchmark use it, and only as the checker.

* :mod:`oracle.ising_oracle` -- NumPy restatement of the reference algorithm (Philox stream,
  thresholds, layouts, int8 and multi-spin-coded half sweeps, observables), plus two
  full-lattice restatements written straight from the definitions, used to pin the first.
* ``oracle/ising_oracle.c`` -- plain C restatement of the same half sweeps for large and long
  runs, built into ``oracle/libising_oracle.so`` and pinned to the NumPy restatement on small
  cases (:mod:`oracle.c_oracle` loads it).
"""
