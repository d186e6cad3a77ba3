"""FP64 CPU oracle for the Toeplitz subspace normal operator and the subspace
backprojection of MRISubspaceRecon.jl (repo MagneticResonanceImaging/
MRFingerprintingRecon.jl).

TEST INFRASTRUCTURE ONLY.  Only the tests, the smoke check and the benchmark's
CPU-baseline leg import this package, and only as the checker (or, for
``cpu_port``, as the timed CPU baseline).  The product path never calls it.

The reference is Julia and cannot run here (no julia, no Julia packages, no
network), so it cannot be compiled into ``oracle/_ref``; every function below
is a restatement that cites the reference line it follows.  What pins each
function (tests/test_oracle_*.py):

* ``nudft``: exact sums; pinned by direct per-sample evaluation, the inverse
  property of the Cartesian DFT and linearity.
* ``nufft64``: ES-kernel NUFFT; pinned against ``nudft`` to 1e-11.
* ``toeplitz.kmask_indcs``: footprint rule of NonuniformFFTs.jl's sigma = 1
  Gaussian spreading (half-support 4 cells, cell = floor(x / dx) with
  x = mod(-2 pi k, 2 pi)).  That library is not vendored in the reference, so
  the rule is recalled, not read: PARITY UNPINNED against the reference for the
  exact footprint (pinned only by invariants: containment of all sample cells,
  point symmetry for point-symmetric trajectories, size bounds).  Effect size
  (``kmask_sensitivity.py``, FP64 CG on the same data and unmasked kernel):
  changing the rule to half-support 3 or 5, or to a symmetric 9-cell footprint,
  moves the config-2 solution (256^2, 20 coils, K = 5, 25 iterations) by
  3.3e-5 to 4.7e-5 and the config-1 solution (64^2, 8 coils, K = 4, 20
  iterations) by 3.1e-4 to 4.4e-4; the other edge convention for samples on a
  cell edge changes no cell at either size.
* ``toeplitz.toeplitz_kernel`` (complex U): pinned by brute force (the masked
  operator equals A^H A from dense NUDFT matrices when the mask covers the
  grid), the Hermitian-per-voxel invariant and the K = 1 classic Toeplitz case.
* ``toeplitz.toeplitz_kernel`` (real U): restated as Re(FFT(p)).  The
  reference's FFTW brfft over NonuniformFFTs' real-plan half spectrum can only
  differ on Nyquist planes of the PSF, which the unmasked operator never uses;
  that Nyquist-plane convention is PARITY UNPINNED.
* ``toeplitz.apply_normal``: pinned by the full-Cartesian identity
  A^H A = prod(N) I, self-adjointness, positive semi-definiteness and brute
  force A^H A x from dense NUDFT matrices.
* ``toeplitz.backprojection``: pinned against a brute-force per-voxel sum and
  the fully sampled Cartesian inverse (test/backprojection_cart.jl:19-66).
* ``cg.cg``: IterativeSolvers.jl 0.9 recurrence (third-party, not vendored);
  pinned by convergence to numpy.linalg.solve and equality with a textbook CG
  on dense Hermitian matrices.
"""
