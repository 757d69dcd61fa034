"""FP64 CPU oracle for the thick_ptycho forward-model hot path.

TEST INFRASTRUCTURE ONLY.  This package restates, in numpy/scipy complex128,
the reference algorithms of bryceshirley/thick_ptycho that the HIP library
replaces (Crank-Nicolson slice stepping with per-slice LU, its adjoint, the
angular-spectrum multislice propagator, far-field intensities and the PWE
reconstruction gradient).  Each function cites the reference file:line it
follows.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker or the timed CPU baseline.
The product package ``thick_ptycho_amd`` never imports it and has no CPU
fallback.

Pinning (how each function is checked independently of its own formula):
  * operators / stepping: the reference's analytic 2nd-order convergence tests
    (2D and 3D, all boundary types), scipy ``solve_banded`` on the same
    tridiagonal systems, a hand-written Thomas solve, Neumann cos-mode
    eigenvector amplification in closed form, free-space norm conservation,
    equality with the whole-system ``full_lu`` solve.
  * adjoint modes: inner-product identities against the assembled step
    operators and ``full_lu`` with ``trans='H'``; central finite differences
    for the exact gradient.
  * multislice: plane-wave eigenvectors with closed-form eigenvalues, constant
    object phase, Parseval.
  * far field: ``numpy.fft`` identities (Parseval, single-mode delta).
  * reference gradient (``compute_grad_least_squares``): no golden values;
    pinned where the object is uniform along z, where its adjoint coincides
    with the exact one and it must equal a closed form in the finite-difference
    pinned exact adjoint states (tests/test_oracle_gradient.py); the z-varying
    case rests on the restatement.  The exact gradient is pinned by finite
    differences (amplitude and exit-wave losses); the reduced-domain gradient
    is the sum of per-window exact gradients, each finite-difference pinned.
  * parallel-in-time GMRES (``pint``): the preconditioner against a sparse
    solve of the assembled alpha-circulant matrix, the solution against the
    whole-system SuperLU solve in both modes.
  * 3PIE (``pie``): modulus-projection invariants, back-propagation as the
    conjugate transpose of the dense propagation matrix, falling loss.
  * learned preconditioner and BiCGSTAB (``learned``): ``correlate2d``,
    complex arithmetic, finite differences, ``spsolve`` and scipy's bicgstab.
  * streaming 3D reconstruction (``recon3d``): SuperLU on the assembled slice
    operators and the finite-difference pinned ``exact_gradient_3d``.
"""
