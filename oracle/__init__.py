"""CPU FP64 oracle for the Schwarz-preconditioned CG path.

TEST INFRASTRUCTURE ONLY. Nothing in the `dd_solvers` package imports this package; only
`tests/`, `__graft_entry__.smoke()` and the CPU-baseline leg of `bench.py` use it, and only
as the checker (or the timed CPU baseline), never as a step of the GPU path.

The reference (mgr0dzicki/dd-solvers) cannot run here: it needs a CUDA build, cuDSS and
numba. This package restates its algorithm in numpy/scipy, one function per reference
routine, each citing the file:line it follows. Paths are relative to the reference root.

Modules
  kernels  batched GEMV, BSR SpMV, CSR->BSR, CSR sort / duplicate merge, SPD solve
  pcg      CG.solve, including the stopping rule, T0, metadata and the Lanczos estimate
  schwarz  AdditiveSchwarz / HybridSchwarz apply, T0, A_C, R0A, generalised to uneven and
           overlapping subdomains

How each function is pinned (reference fixtures, closed forms, brute force) is listed in
the docstring of each module; tests/test_oracle.py holds the checks.
"""
