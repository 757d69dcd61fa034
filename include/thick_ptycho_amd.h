/*
 * thick_ptycho_amd — C ABI of the MI355X (gfx950) kernels for thick-sample
 * ptychography forward models: Crank–Nicolson slice stepping of the paraxial
 * wave equation (per-slice tridiagonal LU), its adjoint, and the
 * angular-spectrum multislice propagator.
 *
 * Conventions
 *   - Every function returns int: TP_OK (0) or a negative TP_ERR_* code;
 *     tp_last_error() returns a message for the calling thread.
 *   - All pointers named d_* are device pointers owned by the caller; complex
 *     values are interleaved (re, im): complex64 = 2 x float, complex128 =
 *     2 x double.  The library never allocates device memory.
 *   - Work is enqueued on `stream` (a hipStream_t; NULL = default stream) and
 *     is asynchronous; calls are capturable into HIP graphs.
 *
 * Reference interfaces replaced (paths relative to bryceshirley/thick_ptycho):
 *   tp_tri_factor   <- PWEIterativeLUSolver._construct_solve_cache /
 *                      _build_lu_factors (thick_ptycho/forward_model/pwe/solvers/
 *                      iterative_lu.py:66-133): spla.splu of A0 - D_j per slice.
 *   tp_tri_sweep    <- PWEIterativeLUSolver._solve_single_probe_impl
 *                      (iterative_lu.py:138-197) driven by
 *                      BaseForwardModelSolver.solve (forward_model/base/
 *                      base_solver.py:84-154): u_j = LU_j^{-1}(B_j u_{j-1} + b).
 *   tp_tri_step     <- one iteration j of the loop at iterative_lu.py:185-193
 *                      for every probe (the "plain" one-launch-per-slice form).
 *   tp_tri_sweep with gfwd/grad set
 *                   <- the adjoint sweep plus gradient accumulation of
 *                      ReconstructorPWE.compute_grad_least_squares
 *                      (thick_ptycho/reconstruction/pwe_reconstructor.py:100-164).
 *   tp_amplitude_seed
 *                   <- ReconstructorBase.get_error_in_exit_wave /
 *                      _apply_phase_retrieval_constraint
 *                      (thick_ptycho/reconstruction/base_reconstructor.py:87-207).
 *   tp_ms_transmission, tp_ms_sweep
 *                   <- MSForwardModelSolver._solve_single_probe
 *                      (thick_ptycho/forward_model/multislice/ms_solver.py:90-159)
 *                      with _propagate / _get_propagation_kernels (:49-78).
 *   tp_farfield     <- BaseForwardModelSolver.get_farfield_intensities
 *                      (forward_model/base/base_solver.py:213-249), noise off.
 *   tp_band_factor  <- the same _construct_solve_cache / _build_lu_factors
 *                      (iterative_lu.py:66-133) for two transverse axes, where
 *                      A0 - D_j is the block-tridiagonal matrix of
 *                      operator_matrices.py:75-126 (bandwidth nx).
 *   tp_band_sweep   <- _solve_single_probe_impl (iterative_lu.py:138-197) for
 *                      two transverse axes, all probes in one launch.
 *   tp_grad3d_accumulate, tp_amplitude_seed2d
 *                   <- the gradient of compute_grad_least_squares
 *                      (pwe_reconstructor.py:100-164) for two transverse axes
 *                      (exact adjoint, per-probe 2D far field).
 *   tp_ms2d_kernel, tp_ms2d_sweep, tp_farfield2d
 *                   <- no counterpart (the reference's multislice raises for
 *                      two transverse axes, ms_solver.py:28-31): the 2D-FFT
 *                      generalisation of MSForwardModelSolver and a per-probe
 *                      2D far field.
 *   tp_bicgstab_solve, tp_krylov_diag, tp_cn_finish, tp_cnn_*
 *                   <- no counterpart in the reference: an iterative
 *                      (learned-preconditioned) alternative to the per-slice
 *                      LU solve of iterative_lu.py:185-193 in 3D, with the
 *                      banded factor above as its fallback.
 *   tp_window_grad_accumulate
 *                   <- ReconstructorPWE.accumulate_gradient (pwe_reconstructor.py:144-164),
 *                      the window sum of reduced-domain gradients.
 *   tp_pie_adjoint_transmission, tp_pie_update (with tp_ms_sweep, conj_h)
 *                   <- ReconstructorMS._apply_backprop_updates / _update_object_n
 *                      (reconstruction/ms_reconstructor.py:109-190), 3PIE.
 *   tp_pint_dots, tp_pint_combine
 *                   <- the Arnoldi projections inside scipy's gmres as called
 *                      at full_pint.py:240-248.
 *   tp_pint_matvec  <- _pintobj_matvec_exact (forward_model/pwe/solvers/utils/
 *                      _pint_utils.py:27-54), the operator of PWEFullPinTSolver
 *                      (full_pint.py:71-275); its alpha-circulant preconditioner
 *                      runs on tp_tri_factor / tp_tri_sweep.
 */
#ifndef THICK_PTYCHO_AMD_H
#define THICK_PTYCHO_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TP_ABI_VERSION 4

#define TP_OK 0
#define TP_ERR_INVALID (-1)     /* bad argument or unsupported size      */
#define TP_ERR_HIP (-2)         /* a HIP runtime call failed              */
#define TP_ERR_SINGULAR (-3)    /* zero pivot in a per-slice factor       */
#define TP_ERR_UNSUPPORTED (-4) /* size/mode not built into this library  */

/* Maximum number of scans (objects of possibly different width) one
 * tp_tri_sweep / tp_ms_sweep launch walks together. */
#define TP_MAX_SCANS 8

int tp_abi_version(void);
const char* tp_last_error(void);

/* ------------------------------------------------------------------------
 * Tridiagonal CN stepping (one transverse axis)
 *
 * Every mode solves A_s u_s = B_s u_{s-1} + b_s with tridiagonal A_s and
 * B_s = diag(sigma_s) - A_s (A0 + B0 = 2I for the CN pair, so sigma = 2 unless
 * the right-hand operator uses a different slice than the left one), so a step
 * is u_s = A_s^{-1}(sigma u_{s-1} + b_s) - u_{s-1}.
 *
 * Rows are split into 64 lanes of `chunk` consecutive rows (chunk = 1, 2, 4,
 * 8, 16 or 32, n <= 64*chunk); n_pad = 64*chunk.  Coefficient buffer
 * (`d_coef`): complex64 [steps][6][n_pad], row i stored at
 * (i % chunk) * 64 + i / chunk (tp_tri_coef_bytes gives its size).
 * ------------------------------------------------------------------------ */

size_t tp_tri_coef_bytes(int n, int steps, int chunk);

/* Factor, for every step s = 1..steps, the LHS tridiagonal
 *     A(s) = tridiag(lhs_sub, lhs_diag + lhs_sign * c_l(s), lhs_sup)
 * where c_l(s) = C[:, lhs_col0 + col_step*(s-1)] (a column outside [0, ncol)
 * contributes 0; conj_c conjugates C first).  With with_sigma != 0 the
 * per-row sigma(s) = diag A(s) + rhs_diag + rhs_sign * c_r(s),
 * c_r(s) = C[:, rhs_col0 + col_step*(s-1)], is stored too; otherwise sigma = 2
 * is implied.  The caller guarantees rhs_sub = -lhs_sub and rhs_sup = -lhs_sup.
 * d_C: complex128, row-major (n, ncol) = (n, nz-1) exactly as returned by
 *      create_object_contribution(...).reshape(-1, nz-1).
 * d_base: complex128 [6][n] = lhs_sub, lhs_diag, lhs_sup, rhs_sub, rhs_diag, rhs_sup
 *      (sub[0] and sup[n-1] are ignored).
 * One wavefront factors one step: pivots in FP64 (no pivot search: the CN
 * operators are strictly diagonally dominant), products stored in complex64.
 * d_status (optional, caller-zeroed int[2]): [0] is set to 1 if a pivot is
 * exactly 0; [1] is raised (atomic max) to the coupling reach of the factors:
 * the largest number of rows of a chunk, counted from its first row forward or
 * its last row backward, over which the carry from the neighbouring chunk
 * keeps a factor above 2^-32 (|re| + |im| of prod(-l) resp. prod(q) in FP64).
 * A reach below `chunk` lets tp_tri_sweep drop the 64-lane carry scans (see
 * tp_tri_scan.reach).  Factor sets that share one d_status get the reach of
 * their union. */
int tp_tri_factor(const void* d_C, int n, int ncol, const void* d_base,
                  double lhs_sign, double rhs_sign, int conj_c,
                  int lhs_col0, int rhs_col0, int col_step, int steps,
                  int chunk, int with_sigma, void* d_coef, int* d_status, void* stream);

/* The arguments of one tp_tri_factor call. */
typedef struct tp_tri_factor_job {
  const void* C;
  const void* base;
  void* coef;
  int* status;
  double lhs_sign, rhs_sign;
  int n, ncol, steps, conj_c, lhs_col0, rhs_col0, col_step, with_sigma;
} tp_tri_factor_job;

/* Up to TP_MAX_SCANS independent tp_tri_factor calls (one chunk for all) in one
 * launch, as the factorisations of every (scan, direction) of a pass
 * (_construct_solve_cache per scan and mode, iterative_lu.py:66-133). */
int tp_tri_factor_batch(const tp_tri_factor_job* jobs, int njobs, int chunk, void* stream);

/* One scan of a multi-scan sweep. All strides are in complex elements. */
typedef struct tp_tri_scan {
  const void* coef;       /* d_coef of this scan (step 1 first)                 */
  const void* u0;         /* initial planes: complex64, u0 + p*u0_pstride       */
  const void* rhs;        /* optional extra RHS per step, or NULL:
                             rhs + p*rhs_pstride + (s-1)*rhs_sstride, complex64 */
  void* out;              /* full field or NULL: plane s at out + p*out_pstride +
                             (out_z0 + out_zdir*s)*out_zstride, s = 0..steps     */
  void* ckpt;             /* checkpoints or NULL: plane s/ckpt_every for
                             s % ckpt_every == 0 at ckpt + p*ckpt_pstride +
                             (s/ckpt_every)*ckpt_zstride                         */
  void* exitp;            /* last plane or NULL: exitp + p*exit_pstride          */
  int64_t u0_pstride, rhs_pstride, rhs_sstride;
  int64_t out_pstride, out_zstride, ckpt_pstride, ckpt_zstride, exit_pstride;
  int n, steps, probes;
  int out_z0, out_zdir, ckpt_every;
  int sigma;              /* coef holds sigma (factored with with_sigma != 0)   */
  int rhs_steps;          /* rhs applies to steps 1..rhs_steps (0 = every step) */
  /* Gradient accumulation (persistent sweep only).  After step s the new state
   * lambda is combined with two planes of `gfwd`, a = plane gfwd_z0 + gfwd_zdir*(s-1)
   * and b = plane gfwd_z0 + gfwd_zdir*s, and
   *     grad[row] += sum over probes of conj(lambda) * (a + b),
   *     row = grad_z0 + grad_zdir*(s-1)
   * (b is left out on row 0 when grad_skip_row0 != 0).  grad: complex64 rows of
   * n values, row stride grad_zstride; accumulated with atomics (caller zeroes). */
  const void* gfwd;
  void* grad;
  int64_t gfwd_pstride, gfwd_zstride, grad_zstride;
  int gfwd_z0, gfwd_zdir, grad_z0, grad_zdir, grad_skip_row0;
  /* 0: every probe of the scan steps through `coef`.  Otherwise probe p has its
   * own factors at coef + p*coef_pstride (complex elements, even), as in the
   * reference's reduced-domain mode where each scan position sees its own window
   * of the object; persistent sweep only, with probes_per_block = 1 and no
   * gradient accumulation. */
  int64_t coef_pstride;
  /* 0: `coef` was factored with the launch's chunk.  Otherwise the chunk it was
   * factored with, which may be half the launch's (launch chunk 16 or 32, no
   * gradient accumulation, n <= 64*chunk): a narrow scan then steps with half
   * as many rows per lane beside the wide ones of the same persistent launch.
   * The plain one-step form takes 0 only. */
  int chunk;
  /* Optional (NULL = full carries): d_status + 1 of the factorisation(s) that
   * wrote `coef` (all of them, for per-probe factors).  When *reach is below
   * the scan's chunk (8 rows or more) and has_offd is set, tp_tri_sweep takes
   * the near-carry step:
   * the carry into a chunk is the neighbouring chunk's local end value, applied
   * to the first / last *reach rows only, with no 64-lane scan; the terms it
   * leaves out are below 2^-32 of the carried values.  Read on the device when
   * the sweep starts, so a sweep queued behind its factorisation needs no host
   * synchronisation.  tp_tri_step takes the same near-carry step (reading the
   * word at every launch); tp_tri_sweep_wide ignores it. */
  const int* reach;
  /* The LHS off-diagonals of the factored operator, which the near-carry step
   * needs (it eliminates with the stored 1/piv and these instead of staging the
   * multipliers): with has_offd != 0, sub is uniform on rows 2 .. n-2 and sup
   * on rows 1 .. n-3, every one of them is purely imaginary, and offd holds
   * complex64 {sub interior, sup interior, sub[1], sub[n-1], sup[0], sup[n-2]}
   * as 12 floats whose real parts are zero (the step reads the imaginary parts;
   * every mode's CN operator with real k, dx, dz has this form, n >= 4).
   * has_offd = 0: no near-carry step. */
  float offd[12];
  int has_offd;
} tp_tri_scan;

/* Persistent batched sweep: one launch walks all `steps` slices of every probe
 * of every scan (nscans <= TP_MAX_SCANS, all with n <= 64*chunk and the same
 * sigma flag).  One wavefront per probe; the per-step coefficients are staged
 * in LDS and shared by the `probes_per_block` probes (1, 2, 4 or 8) of one
 * workgroup. */
int tp_tri_sweep(const tp_tri_scan* scans, int nscans, int chunk, int probes_per_block, void* stream);

/* Plain form: advance every probe of every scan by the single step `s`
 * (1-based), reading plane s-1 from `out` (or u0 when s == 1, which also
 * writes plane 0) and writing plane s.  Requires `out`. */
int tp_tri_step(const tp_tri_scan* scans, int nscans, int chunk, int s, void* stream);

/* Wide grids: each probe's rows are split over `waves` wavefronts (1..16) of
 * one workgroup, 64 * waves lanes of `chunk` rows (chunk = 8 or 16), so
 * n <= 64 * waves * chunk (16384 rows at chunk 16 and 16 wavefronts).  The
 * coefficient buffer is complex64 [steps][6][n_pad] with n_pad = 64*waves*chunk
 * and row i at (i % chunk) * 64 * waves + i / chunk (with waves = 1 the layout
 * above); tp_tri_coef_bytes_wide gives its size.
 * Replaces the same reference interfaces as tp_tri_factor / tp_tri_sweep
 * (iterative_lu.py:66-197) for the widths of the reference's own examples
 * (N = 12001 and 7400: notebooks/2D_examples/1_generate_simulated_data.ipynb,
 * 4.pwe_parallel_in_time_forward.ipynb), which the one-wavefront kernels
 * (n <= 2048) do not hold. */
size_t tp_tri_coef_bytes_wide(int n, int steps, int chunk, int waves);

/* tp_tri_factor for the wide layout: one workgroup of 64 * waves lanes per step.
 * d_status[1] takes the coupling reach as in tp_tri_factor (tp_tri_sweep_wide
 * carries across every chunk whatever it is). */
int tp_tri_factor_wide(const void* d_C, int n, int ncol, const void* d_base,
                       double lhs_sign, double rhs_sign, int conj_c,
                       int lhs_col0, int rhs_col0, int col_step, int steps,
                       int chunk, int waves, int with_sigma, void* d_coef, int* d_status,
                       void* stream);

/* Persistent sweep for the wide layout: one probe per workgroup; every field of
 * tp_tri_scan applies (rhs, out/ckpt/exit, gradient accumulation with global
 * atomics, per-probe factors through coef_pstride) except a half chunk: scans
 * must leave `chunk` 0 or equal to the launch's. */
int tp_tri_sweep_wide(const tp_tri_scan* scans, int nscans, int chunk, int waves, void* stream);

/* ------------------------------------------------------------------------
 * Multislice (angular spectrum) and far field, n = 64 * r with r = 1 .. 32
 * ------------------------------------------------------------------------ */

/* T[z][x] = exp(i k (n[x, z] - n_medium) dz) for z = 0..nz-2, complex64.
 * d_n: complex128 row-major (n, nz). */
int tp_ms_transmission(const void* d_n, int n, int nz, double k, double dz, double n_medium,
                       void* d_T, void* stream);

/* Propagator kernel H(kx) * (1/n) in natural frequency order, complex64 [n]. */
int tp_ms_kernel(int n, double dx, double k, double dz, void* d_H, void* stream);

typedef struct tp_ms_scan {
  const void* T;          /* complex64 [steps][n] transmission per slice          */
  const void* H;          /* complex64 [n] = H / n                                */
  const void* u0;         /* complex64 initial planes, u0 + p*u0_pstride          */
  void* out;              /* full field or NULL (plane z at out + p*out_pstride + z*out_zstride) */
  void* exitp;            /* last plane or NULL                                   */
  int64_t u0_pstride, out_pstride, out_zstride, exit_pstride;
  int n, steps, probes, conj_h;  /* conj_h: back-propagation with conj(H)     */
} tp_ms_scan;

/* Persistent multislice: psi <- ifft(fft(psi * T_z) * H) for z = 0..steps-1,
 * one wavefront per probe. */
int tp_ms_sweep(const tp_ms_scan* scans, int nscans, void* stream);

/* Plain form: one slice z (0-based) for every probe, reading plane z from
 * `out` (u0 when z == 0, also writing plane 0) and writing plane z+1. */
int tp_ms_step(const tp_ms_scan* scans, int nscans, int z, void* stream);

/* Amplitude-misfit seed for the adjoint: U = fft(exit) per probe,
 *   loss[p]  = sum_k (|U_k| - a_k)^2,
 *   seed[p]  = scale * sum_k W_k exp(+2 pi i k x / n),  W = (|U| - a) U / |U| (0 where U = 0),
 * i.e. scale = 1 gives dJ/d conj(exit) and scale = 1/n the reference's exit-wave
 * error exit - ifft(a U/|U|).  exit, seed: complex64 [probes][n]; a: float32
 * [probes][n] (sqrt of the measured intensities); loss: float32 [probes]. */
int tp_amplitude_seed(const void* d_exit, const void* d_sqrt_i, int n, int probes, double scale, void* d_seed,
                      void* d_loss, void* stream);

/* Far-field intensities |fft(exit)|^2 (unnormalised, no shift), float32
 * [probes][n], for exit waves complex64 [probes][n] (row stride ld). */
int tp_farfield(const void* d_exit, int64_t ld, int n, int probes, void* d_intensity, void* stream);

/* ------------------------------------------------------------------------
 * Multislice, far field and amplitude seed at any length n (1 <= n <= 16384)
 *
 * Replace the same reference interfaces as tp_ms_sweep / tp_farfield /
 * tp_amplitude_seed (ms_solver.py:49-159, base/base_solver.py:213-249,
 * base_reconstructor.py:87-207), which numpy computes at any n (the
 * reference's 2D reconstruction notebook uses n = 751).  Each operation is a
 * circular convolution on F = tp_spectral_len(n) points (a power of two,
 * F >= 2n - 1, at least 64), one workgroup per probe with the transform in
 * LDS: multislice as the convolution with h = ifft_n(H), a length-n DFT by
 * Bluestein's chirp-z identity.  For n > 8192 (F = 32768) the calls need a
 * workspace of tp_spectral_work_bytes(n, probes) bytes; below it that size is 0.
 * ------------------------------------------------------------------------ */
int tp_spectral_len(int n);
size_t tp_spectral_work_bytes(int n, int probes);
/* Bytes of the FP64 set-up scratch the two table builders below take. */
size_t tp_spectral_setup_bytes(int n);

/* Multislice convolution spectrum S (complex64 [F], 1/F folded in) for the
 * propagator H(kx) of tp_ms_kernel: S = FFT_F(h on (-n, n)) / F with
 * h = ifft_n(H), all in FP64 direct sums.  d_setup: tp_spectral_setup_bytes(n). */
int tp_ms_kernel_any(int n, double dx, double k, double dz, void* d_S, void* d_setup, void* stream);

/* Bluestein tables: chirp w_m = exp(i pi m^2 / n) (complex64 [n]) and
 * S = FFT_F(w on (-n, n)) / F (complex64 [F]). */
int tp_chirp_tables(int n, void* d_chirp, void* d_S, void* d_setup, void* stream);

/* tp_ms_sweep at any n: every scan shares n, and H points at the spectrum S of
 * tp_ms_kernel_any (conj_h uses conj(S), the spectrum of ifft_n(conj(H))).
 * One probe per workgroup; d_work: tp_spectral_work_bytes(n, total probes). */
int tp_ms_sweep_any(const tp_ms_scan* scans, int nscans, void* d_work, void* stream);

/* tp_farfield at any n, with the tables of tp_chirp_tables. */
int tp_farfield_any(const void* d_exit, int64_t ld, int n, int probes, const void* d_chirp, const void* d_S,
                    void* d_intensity, void* d_work, void* stream);

/* tp_amplitude_seed at any n (the forward DFT and the unnormalised inverse by
 * Bluestein), with the tables of tp_chirp_tables. */
int tp_amplitude_seed_any(const void* d_exit, const void* d_sqrt_i, int n, int probes, double scale,
                          const void* d_chirp, const void* d_S, void* d_seed, void* d_loss, void* d_work,
                          void* stream);

/* Strided batched DFT of any length n (Bluestein, tables of tp_chirp_tables):
 * X_k = sum_j x_j exp(-2 pi i j k / n) (inverse != 0: exp(+...), unnormalised)
 * for `count` sequences x_j = in[t*in_dist + j*in_stride] (complex64), written to
 * out[t*out_dist + k*out_stride]: complex64, or |X_k|^2 as float32 with
 * power != 0.  Strides in elements; d_work: tp_spectral_work_bytes(n, count).
 * Builds the reference's two-axis far field (base/base_solver.py:237-242,
 * numpy fft2 over the flattened (scans, N) exit waves) and the per-probe fft2
 * of planes that are not powers of two. */
int tp_dft_any(const void* d_in, int64_t in_stride, int64_t in_dist, void* d_out, int64_t out_stride,
               int64_t out_dist, int n, int count, int inverse, int power, const void* d_chirp, const void* d_S,
               void* d_work, void* stream);

/* x[t*dist + k1*B + j2] *= exp(sign 2 pi i j2 k1 / (A B)) for t < count (FP64
 * phases): the twiddles between the passes of a four-step DFT of length A*B
 * (lengths above 16384, e.g. a 256 x 256 plane flattened). */
int tp_twiddle_mul(void* d_x, int A, int B, int count, int64_t dist, int sign, void* stream);

/* ------------------------------------------------------------------------
 * Refinement of the forward sweep (residual and seed beyond complex64)
 *
 * The exit plane of the amplitude-misfit gradient (pwe_reconstructor.py:100-164,
 * base_reconstructor.py:87-207) is needed beyond complex64: the loss amplifies
 * exit-plane errors in far-field bins many orders of magnitude below the peak.
 * A refinement step is the residual of the FP64 recurrence at the current
 * iterate, then a complex64 sweep of the same factors on it (tp_tri_sweep with
 * rhs); the iterate is the sum of the complex64 planes of the sweep and its
 * corrections.  These two calls are the library's FP64 arithmetic besides the
 * factor pivots and the Krylov scalars.
 * ------------------------------------------------------------------------ */

/* r_s = 2 v_{s-1} - A_s (v_{s-1} + v_s) in FP64 for local steps s = 1..steps
 * (sigma = 2 modes, B_s = 2I - A_s), v = u + u1 + u2 (complex64 plane sets with
 * the same strides; u1, u2 may be NULL), A(s) as in tp_tri_factor at global step
 * s_first + s: tridiag(lhs_sub, lhs_diag + lhs_sign * c_l, lhs_sup) from the
 * FP64 d_C and d_base.  Plane t at u + p*u_pstride + t*u_zstride (t = 0..steps);
 * r: complex64, step s at r + p*r_pstride + (s-1)*r_sstride. */
int tp_tri_residual(const void* d_u, const void* d_u1, const void* d_u2, int64_t u_pstride,
                    int64_t u_zstride, int probes, int n, int steps, const void* d_C, int ncol,
                    const void* d_base, double lhs_sign, int conj_c, int lhs_col0, int col_step,
                    int s_first, void* d_r, int64_t r_pstride, int64_t r_sstride, void* stream);

/* tp_amplitude_seed for an exit plane x = x0 + x1 + x2 (complex64 each, row
 * stride ld; x1, x2 may be NULL) and amplitudes sqrt_i + sqrt_i_lo (float32;
 * sqrt_i_lo may be NULL), both FFTs in FP64; n a power of two from 64 to 8192. */
int tp_amplitude_seed_refined(const void* d_x0, const void* d_x1, const void* d_x2, int64_t ld,
                              const void* d_sqrt_i, const void* d_sqrt_i_lo, int n, int probes,
                              double scale, void* d_seed, void* d_loss, void* stream);

/* tp_tri_residual for two transverse axes (the block-tridiagonal A_s of
 * tp_band_factor, operator_matrices.py:75-126): d_tri complex128 [3][3][nx]
 * (block type first / interior / last; sub, diag, sup), d_cpl complex128
 * [ny][2] (f_j, e_j), C element (i, col) at d_C[i * c_istride + col * c_cstride]
 * (complex128).  Planes of nx * ny values; probes, steps <= 65535. */
int tp_band_residual(const void* d_u, const void* d_u1, const void* d_u2, int64_t u_pstride,
                     int64_t u_zstride, int probes, int nx, int ny, int steps, const void* d_C,
                     int64_t c_istride, int64_t c_cstride, int ncol, const void* d_tri, const void* d_cpl,
                     double lhs_sign, int conj_c, int lhs_col0, int col_step, int s_first, void* d_r,
                     int64_t r_pstride, int64_t r_sstride, void* stream);

/* tp_amplitude_seed2d (nx rows of ny values per probe, fft2 over both axes) for
 * an exit plane x0 + x1 + x2 (complex64, probe stride ld; x1, x2 may be NULL)
 * and amplitudes sqrt_i + sqrt_i_lo (float32; sqrt_i_lo may be NULL), both
 * transforms in FP64; nx, ny powers of two from 8 to 2048.  d_work:
 * tp_amplitude_seed2d_refined_work_bytes(nx, ny, probes). */
size_t tp_amplitude_seed2d_refined_work_bytes(int nx, int ny, int probes);
int tp_amplitude_seed2d_refined(const void* d_x0, const void* d_x1, const void* d_x2, int64_t ld,
                                const void* d_sqrt_i, const void* d_sqrt_i_lo, int nx, int ny, int probes,
                                double scale, void* d_seed, void* d_loss, void* d_work, void* stream);

/* ------------------------------------------------------------------------
 * Object filters of the one-axis reconstruction loop
 * (ReconstructorPWE.tv_denoise_amp / gaussian_lowpass_phase /
 * global_object_update, thick_ptycho/reconstruction/pwe_reconstructor.py:370-419).
 * Arrays are row-major (ny, nx); for the (x, z) object ny = x, nx = z.
 * ------------------------------------------------------------------------ */

/* A = |n|, Phi = arg(n) (float64) of a complex128 object. */
int tp_object_polar(const void* d_n, int ny, int nx, void* d_A, void* d_Phi, void* stream);

/* n = A exp(i Phi), complex128. */
int tp_object_join(const void* d_A, const void* d_Phi, int ny, int nx, void* d_n, void* stream);

/* In place: n_iter Chambolle-type TV iterations on the float64 array u with
 * periodic differences and tau = 1/4 (lam > 0), then u - lam div(p).
 * d_work: 3 * ny * nx float64. */
int tp_tv_denoise(void* d_u, int ny, int nx, double lam, int n_iter, double eps, void* d_work, void* stream);

/* tp_tv_denoise without a caller work buffer and without any allocation.  No
 * counterpart in the reference: it is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops).  The work space
 * is a file-scope device array of 3 * TP_TV_NOALLOC_MAX_POINTS float64 (48 MiB),
 * one per GPU the library runs on, held until the process ends; planes of more
 * than TP_TV_NOALLOC_MAX_POINTS points return TP_ERR_INVALID.  Every call on a
 * GPU shares that GPU's array: calls that overlap on one GPU (other host
 * threads, other streams) race and must be serialised by the caller. */
#define TP_TV_NOALLOC_MAX_POINTS (1 << 21)
#define TP_NOALLOC_MAX_DEVICES 64
int tp_tv_denoise_noalloc(void* d_u, int ny, int nx, double lam, int n_iter, double eps, void* stream);

/* In place on a complex64 spectrum X (ny, nx) in FFT order: X *= fftshift(W),
 * W = exp(-(qx^2 + qy^2) / (2 sigma^2)), q = linspace(-1/2, 1/2, n, endpoint=False). */
int tp_lowpass_weight(void* d_X, int ny, int nx, double sigma, void* stream);

/* ------------------------------------------------------------------------
 * CN stepping with two transverse axes (N = nx * ny unknowns per slice)
 *
 * The per-slice operator is block tridiagonal: ny diagonal blocks A_jj
 * (tridiagonal nx x nx: x stencil plus the object term) coupled by scaled
 * identities, A(j, j-1) = f_j I and A(j, j+1) = e_j I.  Flat index j*nx + i.
 * tp_band_factor forms the block LU, S_0 = A_00, S_j = A_jj - f_j e_{j-1} S_{j-1}^{-1},
 * in FP64 and keeps each Schur inverse W_j = S_j^{-1} on the band
 * |col - row| <= w (exact band entries, from the banded LDU of S_j and the
 * Takahashi recurrences); entries beyond the band are dropped, and
 * d_edge[s] = max |W_j(i, i +- w)| / max |W_j(i, i)| over the slice's blocks
 * measures what the band leaves out (the caller widens w until it is below
 * single-precision rounding).  Band coefficients: complex64
 * [steps][ny][w+1][nx][2]: diagonal d of W_j at pair d/2, slot d%2, i.e.
 * W_j(i, i + d - w) at ((d/2)*nx + i)*2 + d%2 (d = 2w+1 is a zero pad).
 * ------------------------------------------------------------------------ */

int64_t tp_band_coef_bytes(int nx, int ny, int steps, int w);

/* d_C: complex128 (N, ncol) row-major (create_object_contribution(...).reshape(-1, nz-1)).
 * d_tri: complex128 [3][3][nx]: the x-stencil diagonals (sub, diag, sup) of the
 *   first block, the interior blocks and the last block of the LHS base operator.
 * d_cpl: complex128 [ny][2] = (f_j, e_j) of the LHS base operator.
 * Block j of A(s) = tri[type(j)] + lhs_sign * diag(c_l(s)) on its diagonal, with
 * c_l(s) = C[j*nx + i, lhs_col0 + col_step*(s-1)] (0 outside [0, ncol); conj_c
 * conjugates).  with_sigma != 0 also stores d_sigma complex64 [steps][N],
 * sigma = 2 + lhs_sign * c_l(s) + rhs_sign * c_r(s), c_r from rhs_col0 (the
 * caller guarantees that the RHS base operator is 2I minus the LHS base).
 * w in {4, 6, 8, 12, 16, 24, 32}, w < nx, nx * (2w+1) * 16 bytes <= 160 KiB.  One wavefront per slice.
 * d_edge: float32 [steps]; d_status: caller-zeroed int, set to 1 on a zero pivot. */
int tp_band_factor(const void* d_C, int nx, int ny, int ncol, const void* d_tri, const void* d_cpl,
                   double lhs_sign, double rhs_sign, int conj_c, int lhs_col0, int rhs_col0, int col_step,
                   int steps, int w, int with_sigma, void* d_W, void* d_sigma, float* d_edge, int* d_status,
                   void* stream);

/* One scan of a 3D sweep; strides in complex elements, planes hold N values. */
typedef struct tp_band_scan {
  const void* W;          /* tp_band_factor band coefficients                     */
  const void* cpl;        /* complex64 [ny][2] = (f_j, e_j) of the factored LHS   */
  const void* sigma;      /* tp_band_factor d_sigma, or NULL for sigma = 2         */
  const void* u0;         /* initial planes, u0 + p*u0_pstride                    */
  void* out;              /* full field or NULL: plane s at out + p*out_pstride +
                             (out_z0 + out_zdir*s)*out_zstride, s = 0..steps     */
  void* exitp;            /* last plane or NULL: exitp + p*exit_pstride          */
  void* ckpt;             /* checkpoints or NULL (as tp_tri_scan)                */
  const void* rhs;        /* extra RHS b_s or NULL: rhs + p*rhs_pstride + (s-1)*rhs_sstride */
  void* work;             /* two planes per probe (work + p*work_pstride), required
                             when out is NULL                                    */
  int64_t u0_pstride, out_pstride, out_zstride, exit_pstride, ckpt_pstride, ckpt_zstride;
  int64_t rhs_pstride, rhs_sstride, work_pstride;
  int nx, ny, w, steps, probes;
  int out_z0, out_zdir, ckpt_every;
  int rhs_steps;          /* rhs applies to steps 1..rhs_steps (0 = every step) */
  int reserved;
} tp_band_scan;

/* Persistent 3D sweep over every probe of every scan (all with the same w,
 * w in {4, 6, 8, 12, 16, 24, 32}, nx <= 256, and the same sigma form): one workgroup per
 * probes_per_block (1, 2 or 4) probes, which share each load of the band coefficients,
 * u_s = A(s)^{-1}(sigma u_{s-1} + b_s) - u_{s-1} (sigma = 2 unless given). */
int tp_band_sweep(const tp_band_scan* scans, int nscans, int probes_per_block, void* stream);
/* tp_band_sweep with a choice of schedule: 0 (the default) prefetches several
 * blocks' coefficients into registers, one workgroup per CU; 1 prefetches one
 * block and fits two workgroups on a CU (built for w <= 8 and 1 or 2 probes per
 * workgroup; TP_ERR_UNSUPPORTED otherwise).  Results are identical. */
int tp_band_sweep_ex(const tp_band_scan* scans, int nscans, int probes_per_block, int schedule, void* stream);

/* Band factors of A(s)^H from those of A(s) (replaces a second tp_band_factor for
 * the adjoint modes of iterative_lu.py:90-128, whose LHS is the conjugate
 * transpose of the forward one): W'_j = W_j^H on the band; with reverse_slices,
 * output slot a holds slice steps-1-a.  Same layout and size as tp_band_factor. */
int tp_band_conj_transpose(const void* d_W, int nx, int ny, int steps, int w, int reverse_slices,
                           void* d_WH, void* stream);

/* The sigma planes of tp_band_factor alone (with_sigma output):
 * sigma_s = 2 + lhs_sign C[:, lhs_col0 + s col_step] + rhs_sign C[:, rhs_col0 + s col_step]. */
int tp_band_sigma(const void* d_C, int nx, int ny, int ncol, double lhs_sign, double rhs_sign, int conj_c,
                  int lhs_col0, int rhs_col0, int col_step, int steps, void* d_sigma, void* stream);

/* ------------------------------------------------------------------------
 * Two transverse axes: multislice with a 2D FFT per slice and per-probe 2D
 * far field (the reference has neither: its multislice raises for 3D,
 * ms_solver.py:28-31).  Planes complex64 (nx, ny) row-major, nx and ny powers
 * of two from 64 to 512.
 * ------------------------------------------------------------------------ */
/* H(ia, ib) = exp(i (kz - k) dz) / (nx ny), kz = sqrt(max(k^2 - kx^2 - ky^2, 0)),
 * kx = 2 pi fftfreq(nx, dx)[ia], ky = 2 pi fftfreq(ny, dy)[ib]; complex64 stored
 * transposed, [ny][nx] (element (ia, ib) at ib*nx + ia), the order in which
 * tp_ms2d_sweep's second pass reads it. */
int tp_ms2d_kernel(int nx, int ny, double dx, double dy, double k, double dz, void* d_H, void* stream);
/* psi <- ifft2(fft2(psi * T_z) * H) (conj(H) when conj_h) for z = 0..steps-1 on
 * every probe.  d_T: complex64 [steps][nx*ny] (tp_ms_transmission with
 * n = nx*ny); d_u0: [probes][nx*ny]; d_out: full field [probes][steps+1][nx*ny]
 * or NULL; d_exit: [probes][nx*ny] or NULL; d_work: 3 * probes * nx * ny complex64. */
int tp_ms2d_sweep(const void* d_T, const void* d_H, const void* d_u0, void* d_out, void* d_exit, int nx, int ny,
                  int steps, int probes, int conj_h, void* d_work, void* stream);
/* |fft2(exit)|^2 per probe (unnormalised, no shift), float32 [probes][nx][ny];
 * d_work: probes * nx * ny complex64. */
int tp_farfield2d(const void* d_exit, int nx, int ny, int probes, void* d_intensity, void* d_work, void* stream);

/* Loss seed of the two-axis amplitude misfit: U = fft2(exit) per probe,
 *   loss[p] += sum (|U| - a)^2   (caller zeroes loss, float32 [probes]),
 *   seed[p]  = scale * unnormalised ifft2((|U| - a) U / |U|)  (scale = 1: dJ/d conj(exit)).
 * a: float32 [probes][nx][ny]; d_work: 2 * probes * nx * ny complex64. */
int tp_amplitude_seed2d(const void* d_exit, const void* d_sqrt_i, int nx, int ny, int probes, double scale,
                        void* d_seed, void* d_loss, void* d_work, void* stream);
/* Per-slice plane products summed over probes (the multislice gradient's
 * accumulation, the adjoint of ms_solver.py:143-158 with respect to the
 * transmission; no counterpart in the reference, which has no MS gradient):
 * G[z][f] += sum_p conj(a[p*a_pstride + (a_z0 + a_zstep z) N + f]) *
 *                    b[p*b_pstride + (b_z0 + b_zstep z) N + f],  z < nplanes. */
int tp_plane_conj_dot(const void* d_a, int64_t a_pstride, int a_z0, int a_zstep, const void* d_b,
                      int64_t b_pstride, int b_z0, int b_zstep, int probes, int nplanes, int64_t N,
                      void* d_G, void* stream);

/* G[s-1][f] += sum_p conj(lam_p[L-s+1][f]) (u_p[s][f] + u_p[s-1][f]) for s = 1..L (= steps):
 * the per-slice gradient accumulation of two transverse axes from full fields,
 * u (forward) and lam (exact adjoint in sweep order), both complex64
 * [probes][steps+1][N]; G complex64 [steps][N]. */
int tp_grad3d_accumulate(const void* d_u, const void* d_lam, int probes, int steps, int64_t N, void* d_G,
                         void* stream);

/* ------------------------------------------------------------------------
 * Learned preconditioner: complex-linear 4-layer 3x3 CNN on the (ny, nx)
 * plane of a 3D step, 1 -> 16 -> 16 -> 16 -> 1 complex channels (2 -> 32 ->
 * 32 -> 32 -> 2 real), no biases or activations.  Parameters: complex64
 * [tp_cnn_param_count()], layer l laid out [9 taps][OC][IC] (tap = 3*dy + dx,
 * (dy, dx) in {0,1,2}^2 reading pixel (y+dy-1, x+dx-1)); layers in order
 * (OC, IC) = (16,1), (16,16), (16,16), (1,16).  They are expanded into a BF16
 * weight buffer of tp_cnn_weight_bytes() for the MFMA kernels (activations
 * BF16, accumulation FP32).  Each plane is divided by its RMS before the
 * network and multiplied back after it, so M is linear.
 * ------------------------------------------------------------------------ */
int tp_cnn_param_count(void);
int64_t tp_cnn_weight_bytes(void);
int64_t tp_cnn_workspace_bytes(int P, int nx, int ny, int train);
/* d_wbuf <- BF16 expansion of d_params (caller-allocated, zeroed here). */
int tp_cnn_expand(const void* d_params, void* d_wbuf, void* stream);
/* d_out = M(d_in) for P planes, complex64 (P, ny*nx) row-major (flat y*nx + x). */
int tp_cnn_apply(const void* d_wbuf, const void* d_in, void* d_out, int P, int nx, int ny, void* d_work,
                 void* stream);
/* One training step on pairs (in, target): loss = mean |M_n(in/s) - target/s|^2
 * (s = RMS of each input plane) -> *d_loss (double); real weight gradients by
 * back-propagation through the same kernels; d_grad (optional) receives the
 * complex-parameter gradient (dL/da, dL/db per weight a + ib); with step >= 1
 * an Adam update (lr, beta1, beta2, eps; bias correction by `step`) of
 * d_params with moments d_m, d_v (complex64 [count], caller-zeroed before the
 * first step) and re-expansion into d_wbuf.  step = 0 only computes. */
int tp_cnn_train_step(void* d_params, void* d_wbuf, void* d_m, void* d_v, const void* d_in, const void* d_target,
                      int P, int nx, int ny, int step, double lr, double beta1, double beta2, double eps,
                      void* d_work, double* d_loss, void* d_grad, void* stream);

/* ------------------------------------------------------------------------
 * BiCGSTAB for A(s) x = b on all P planes of a slice (3D CN step)
 * ------------------------------------------------------------------------ */
/* d_diag (complex64 [N]) <- diagonal of A(s): tri diagonal of the block type +
 * lhs_sign * C[:, col] (conj_c conjugates; col outside [0, ncol) adds 0).
 * d_tri: complex128 [3][3][nx] as for tp_band_factor. */
int tp_krylov_diag(const void* d_C, int nx, int ny, int ncol, const void* d_tri, double lhs_sign, int conj_c,
                   int col, void* d_diag, void* stream);
int64_t tp_bicgstab_workspace_bytes(int P, int nx, int ny, int max_iter, int precond);
/* Right-preconditioned BiCGSTAB (precond 0 none, 1 Jacobi, 2 the CNN with
 * weights d_cnn_w) from the initial guess in d_x; stops a plane at
 * ||b - A x|| <= tol ||b||.  d_tri32 (complex64 [3][3][nx]) and d_cpl32
 * (complex64 [ny][2]) hold the base operator off-diagonals.  Inner products in
 * FP64.  d_done (int [P], optional): k >= 1 converged after k iterations, 0 not
 * converged within max_iter, -1 breakdown.  check_every > 0 synchronises the
 * stream every that many iterations and stops when every plane is done (not
 * capturable); 0 runs exactly max_iter iterations without host synchronisation. */
int tp_bicgstab_solve(const void* d_diag, const void* d_tri32, const void* d_cpl32, int nx, int ny, int P,
                      const void* d_b, void* d_x, int precond, const void* d_cnn_w, double tol, int max_iter,
                      int check_every, void* d_work, int* d_done, void* stream);
/* out = xscale * x - uprev over `total` complex64 values (the CN update). */
int tp_cn_finish(int64_t total, const void* d_x, const void* d_uprev, void* d_out, double xscale, void* stream);

/* ---------------------------------------------------------------- parallel in time
 * Whole space-z system of one transverse axis applied matrix-free, for every
 * probe: (M x)_j = (A - diag C_j) x_j - (B + diag C_j) x_{j-1}, x_{-1} = 0, and
 * y = M x, or y = b - M x when d_b is non-null (b may alias y; x may not).
 * Replaces _pintobj_matvec_exact (forward_model/pwe/solvers/utils/_pint_utils.py:27-54)
 * in its forward form; the reference's adjoint-mode system passes A^H, B^H and the
 * conjugated, slice-reversed C in the same form (full_pint.py:148-160).
 *   d_ops: A sub, diag, sup, B sub, diag, sup (6 x N complex; sub[i] multiplies
 *          x[i-1], sup[i] multiplies x[i+1]);  d_C: (L, N) complex, row j = C_j,
 *          shared by every probe (c_pstride = 0) or one block per probe at
 *          p * c_pstride (c_pstride >= L * N: reduced-domain windows);
 *   d_x, d_b, d_y: (P, L, N) complex.
 * Complex128 throughout when fp64 = 1, complex64 when 0. */
int tp_pint_matvec(const void* d_ops, const void* d_C, int64_t c_pstride, const void* d_x, const void* d_b, void* d_y,
                   int N, int L, int P, int fp64, void* stream);
/* Arnoldi projections of a batch of probes, the Krylov-basis work of the GMRES
 * call at full_pint.py:240-248 (scipy.sparse.linalg.gmres).  V: (P, rows, len)
 * complex64 with probe stride v_pstride and row stride v_rstride; w: (P, len).
 * tp_pint_dots writes partial[(p * nblk + b) * 33 + i] = block b's share of
 *   sum_e conj(V[p, i, e]) w[p, e] for i < nrows (nrows <= 32) and of
 *   sum_e |w[p, e]|^2 (real part) in slot nrows; the caller sums over b.
 * tp_pint_combine writes out[p, e] = w[p, e] - sum_{i < nrows} V[p, i, e] c[p, i]
 *   (w null: 0; w may alias out). */
/* Reduced-domain gradient accumulation, the per-scan window sum of
 * ReconstructorPWE.accumulate_gradient (reconstruction/pwe_reconstructor.py:144-164)
 * for the exact gradient.  u, lam: (S, L + 1, W) complex64, the forward planes
 * of every scan and its exact-adjoint planes in stepping order (plane k holds
 * lambda_{L-k+1}); win: S int32 first rows of the windows (a_s + W <= nx);
 * G: (L, nx) complex64, overwritten with
 *   G[j-1][x] = sum_{s: a_s <= x < a_s + W} conj(lam^s_j) (u^s_j + u^s_{j-1}) at row x - a_s. */
int tp_window_grad_accumulate(const void* d_u, const void* d_lam, const int* d_win, int S, int L, int W, int nx,
                              void* d_G, void* stream);

/* ---------------------------------------------------------------- multislice 3PIE
 * ReconstructorMS._apply_backprop_updates / _update_object_n
 * (reconstruction/ms_reconstructor.py:109-190) for one scan position, split in
 * three launches on the unchanged object: tp_pie_adjoint_transmission, a
 * tp_ms_sweep with conj_h = 1 over the nz factors it writes (starting from the
 * corrected exit wave), and tp_pie_update.
 * n: complex128 rows of the scan's window (N rows x nz, row stride ld >= nz).
 * tp_pie_adjoint_transmission: Tadj (nz, N) complex64, row 0 = 1,
 *   row j = conj(exp(i k (n[:, nz-j] - n_medium) dz)).
 * tp_pie_update: psi (nz, N) forward planes, back (nz + 1, N) the sweep's planes;
 *   for every slice z: x = conj(T_z) back[nz - z],
 *   n[:, z] += step conj(psi_z) (x - psi_z) / (max |psi_z|^2 + eps) / (i k dz). */
int tp_pie_adjoint_transmission(const void* d_n, int64_t ld, int N, int nz, double k, double dz, double n_medium,
                                void* d_Tadj, void* stream);
int tp_pie_update(void* d_n, int64_t ld, int N, int nz, const void* d_psi, const void* d_back, double k, double dz,
                  double n_medium, double step, double eps, void* stream);

int tp_pint_dots(const void* d_V, int64_t v_pstride, int64_t v_rstride, int nrows, const void* d_w,
                 int64_t w_pstride, int64_t len, int P, int nblk, void* d_partial, void* stream);
int tp_pint_combine(const void* d_V, int64_t v_pstride, int64_t v_rstride, int nrows, const void* d_c, int c_pstride,
                    const void* d_w, int64_t w_pstride, void* d_out, int64_t out_pstride, int64_t len, int P,
                    void* stream);

#ifdef __cplusplus
}
#endif
#endif /* THICK_PTYCHO_AMD_H */
