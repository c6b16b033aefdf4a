"""CPU oracle for the pairwise Gaussian-kernel path - TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this package, and only as the checker (or, for the CPU baseline, as the thing
timed on the host). The product (flash-sd-kde_amd/) never imports it.

Modules:
  reference_np  FP32 numpy restatement of the reference's flash_sd_kde/reference.py
                (op order kept: norm expansion, clamp at 0, exp, row sums).
  fp64          FP64 direct-difference sums, FP64 log-sum-exp log-densities
                (restating experiments/error_suite_16d/reference.py:32-82), FP64
                score accumulators and debias, and the error metrics.

Pinning (what checks each function; see tests/test_oracle.py):
  reference_np.*      golden vectors written by the reference's own reference.py
                      (tests/golden/make_golden.py) + FP64 brute force.
  fp64.kde_sums / log_density / score_sums
                      closed forms (n = 1, identical points, two points at
                      distance r), and agreement with the golden FP32 vectors
                      within FP32 rounding.
  fp64.loo_loglik     brute force over j != i on tiny sets (independent loop).
"""
