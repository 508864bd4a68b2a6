"""Test infrastructure: the CPU oracle of the shift-invert path.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this package, and
only as the checker (or, for the CPU baseline, as the timed CPU restatement); the product path
(hsolve_amd) never imports it. Contents:
  yee_exact     closed-form discrete spectra of homogeneous Yee/FIT pencils (pins inertia counts
                and eigenvalues at every size);
  inertia       Sylvester counts from an independent sparse LDL^T (SuperLU with diagonal pivots)
                and from dense symmetric eigenvalues;
  evaluate      the reference's independent evaluator rules (backward error, matching, ties);
  eigs_cpu      a scipy restatement of the whole solve (SuperLU shift-invert Lanczos + exact
                inertia certificate), checked against the reference's golden spectra;
  cpu_path      the CPU baseline the bench times: C++/OpenMP multifrontal LDL^T with Bunch-Kaufman
                pivots (cpu_ldlt.cpp, OpenBLAS kernels) over the engine's host analysis, the same
                block Krylov-Schur driver on host tensors, one certified spectrum slice end to end;
                pinned by tests/test_cpu_path.py (inertia = SuperLU and closed form, solve backward
                error, a slice passing the evaluator).
Each function states the reference file:line it follows, and how it is pinned.
"""
