"""CPU oracle for the demap + min-sum hot path.

TEST INFRASTRUCTURE.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker or the timed CPU baseline -- never as part of the product path.

Contents:
  minsum_oracle.c / ldpc.py  FP32 min-sum with exactly the semantics of the
                             upstream compiled kernel lab/ldpc_kernel.c:65-146,
                             the alpha ladder of lab/m9_fast.py:843-891, and a
                             pure-Python scalar restatement for tiny codes.
  demap.py                   FP64 max-log demapper (lab/m6_bicm.py:213-225),
                             the sigma^2 estimator, CSI weighting and the LDM
                             Gaussian-folded core demap (lab/m15_ldm_llr.py:77-85).
  sp_ref.py                  float64 log-tanh sum-product (lab/m16_margin.py:
                             325-390) and the near-miss rescue policy
                             (lab/m44_ldm.py:2527-2555).
  ti.py                      twisted-block (hybrid) and convolutional time
                             de-interleavers (lab/m6_tbi.py:47-96,
                             lab/m10_cti.py:108-152, lab/m44_ldm.py:1187-1304).
  cpe.py                     per-symbol common-phase-error correction
                             (lab/m6_cells.py:176-209), dummy-cell SNR
                             (lab/m16_margin.py:303-318), per-cell noise
                             weights (lab/m9_fast.py:1096-1115).
  Makefile                   builds liboracle_minsum.so and, where the upstream
                             tree is present, _ref/ldpc_kernel.so from the
                             upstream's own source file.

Pinning (what each function is checked against) is listed in the module
docstrings and exercised by tests/test_oracle.py.
"""
