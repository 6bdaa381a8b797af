"""CPU oracle for the TT-SVD / TR-SVD hot path (test infrastructure, FP64).

Only tests, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg import this package;
the engine never does.  See ``ttsvd.py`` for the reference file:line each function restates.

Parity status of each function (how it is pinned, in ``tests/test_oracle.py``):

* ``philox.philox4x32_10`` -- the three published Random123 known-answer vectors.
* ``philox.log_f32``/``sincos_quadrant``/``normal4`` -- against numpy log/sin/cos (within 2e-6)
  and by moment checks; bit-level agreement with the GPU is checked on the GPU.
* ``ttsvd.gram_eig_step``/``rsvd_step``/``tt_svd``/``tr_svd`` -- closed forms: numpy FP64 SVD
  singular values, exact-rank inputs reconstruct to roundoff, orthonormal cores, the randomized
  step equals the exact step when rank(A) <= ell; brute-force loops on tiny shapes.
* ``ttsvd.permute_231``/``permute_312`` -- brute-force index loops of the reference formulas and
  the inverse-pair identity.
* ``ttsvd.delta_rank`` -- hand-computed spectra.
"""
