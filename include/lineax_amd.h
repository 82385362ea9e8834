/*
 * lineax_amd — C ABI of the MI355X (gfx950) solver kernels behind the
 * lineax-compatible host layer in lineax_amd/.
 *
 * Conventions shared by every entry point:
 *   - `stream` is a hipStream_t passed as void* (0 = the legacy default
 *     stream).  Every call is asynchronous and stream-ordered: it enqueues
 *     kernels and returns; it never synchronises and never allocates device
 *     memory.  Scratch memory is caller-provided (`*_workspace_size`).
 *   - Matrices are row-major.  Batched arguments carry an element stride
 *     between consecutive systems; a stride of 0 broadcasts one operand
 *     (e.g. one factorisation reused across many right-hand sides).
 *   - `dtype` is LXA_F32 or LXA_F64 and applies to every floating operand
 *     unless an argument says otherwise.
 *   - The return value is LXA_OK or an error code; lxa_last_error() gives a
 *     message for the calling thread.  Numerical outcomes (zero pivots,
 *     non-positive pivots, refinement status) are never errors: they are
 *     written per system to device `info` arrays and mapped to lineax's
 *     RESULTS codes by the host.
 *
 * Reference citations are to patrick-kidger/lineax (lineax/...).
 */
#ifndef LINEAX_AMD_H_
#define LINEAX_AMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LXA_OK 0
#define LXA_EINVAL 1       /* invalid argument (shape, stride, dtype, size) */
#define LXA_EHIP 2         /* a HIP runtime call or kernel launch failed    */
#define LXA_EUNSUPPORTED 3 /* valid request outside what is implemented     */

#define LXA_F32 0
#define LXA_F64 1

/* RESULTS codes, in lineax's declaration order (lineax/_solution.py:52-68). */
#define LXA_RESULT_SUCCESSFUL 0
#define LXA_RESULT_MAX_STEPS_REACHED 1
#define LXA_RESULT_SINGULAR 2
#define LXA_RESULT_BREAKDOWN 3
#define LXA_RESULT_STAGNATION 4
#define LXA_RESULT_CONLIM 5
#define LXA_RESULT_NONFINITE_INPUT 6

const char* lxa_version(void);
const char* lxa_last_error(void);

/* ------------------------------------------------------------------------
 * LU (lineax/_solver/lu.py)
 * ---------------------------------------------------------------------- */

/* getrf with partial pivoting on `batch` matrices of order n.
 * Replaces jax.scipy.linalg.lu_factor(operator.as_matrix()) at
 * lineax/_solver/lu.py:53.  `a` is overwritten with the unit lower factor
 * (strictly below the diagonal) and the upper factor.  piv[k] (0-based) is
 * the row interchanged with row k at step k, as jax.scipy's lu_factor
 * returns.  info[i] = k+1 for the first exactly-zero pivot of system i, else
 * 0 (the factorisation still completes; the solve then yields inf/NaN,
 * which lineax maps to RESULTS.singular at lineax/_solve.py:104-113). */
int lxa_lu_factor(void* stream, int dtype, int64_t batch, int64_t n, void* a,
                  int64_t a_batch_stride, int32_t* piv, int32_t* info);

/* getrs: solves A x = b (trans = 0) or A^T x = b (trans = 1) in place in b
 * ([batch, n], stride b_batch_stride).  Replaces jax.scipy.linalg.lu_solve
 * at lineax/_solver/lu.py:64 (trans flag from lu.py:60-61, flipped by
 * LU.transpose at lu.py:68-81). */
int lxa_lu_solve(void* stream, int dtype, int64_t batch, int64_t n,
                 const void* lu, int64_t lu_batch_stride, const int32_t* piv,
                 int64_t piv_batch_stride, void* b, int64_t b_batch_stride,
                 int trans);

/* getrf + getrs fused for small orders (n <= 64), one launch: x = A^-1 b
 * for `batch` systems, A read-only ([batch, n, n], stride a_batch_stride),
 * b [batch, n] (stride b_batch_stride), x [batch, n] (stride
 * x_batch_stride).  The factors and pivots (lu [batch, n, n] contiguous,
 * piv [batch, n], 0-based as lxa_lu_factor) are written only when lu / piv
 * are non-null (both or neither); info as lxa_lu_factor (may be null).
 * Equals lxa_lu_factor followed by lxa_lu_solve (trans = 0) bit for bit.
 * Replaces the lu_factor / lu_solve pair of lineax/_solver/lu.py:53, 64 for
 * a vmapped solve that forms its state and solves once (the first compute
 * of a batch of small operators). */
int lxa_lu_small_solve(void* stream, int dtype, int64_t batch, int64_t n,
                       const void* a, int64_t a_batch_stride, const void* b,
                       int64_t b_batch_stride, void* x, int64_t x_batch_stride,
                       void* lu, int32_t* piv, int32_t* info);

/* ------------------------------------------------------------------------
 * Cholesky (lineax/_solver/cholesky.py)
 * ---------------------------------------------------------------------- */

/* potrf, upper (A = U^T U), on `batch` matrices of order n.  Replaces
 * jax.scipy.linalg.cho_factor(matrix) at lineax/_solver/cholesky.py:60.
 * Only the upper triangle of A is read.  negate != 0 factors -A (the NSD
 * branch, cholesky.py:58-59).  On success the upper triangle holds U and the
 * strict lower triangle zeros; on a non-positive (or NaN) pivot the whole
 * matrix is set to NaN and info[i] = k+1, matching JAX's NaN-filled
 * factor so that the solve reports RESULTS.singular. */
int lxa_chol_factor(void* stream, int dtype, int64_t batch, int64_t n, void* a,
                    int64_t a_batch_stride, int negate, int32_t* info);

/* potrs: solves U^T U x = b in place in b.  Replaces
 * jax.scipy.linalg.cho_solve at lineax/_solver/cholesky.py:74; negate != 0
 * negates the solution (cholesky.py:75-76). */
int lxa_chol_solve(void* stream, int dtype, int64_t batch, int64_t n,
                   const void* u, int64_t u_batch_stride, void* b,
                   int64_t b_batch_stride, int negate);

/* Uneven batch of Cholesky solves: system i has order n[i] in [1, 256], its
 * float32 matrix at a + a_offset[i] (row-major, leading dimension n[i];
 * only the upper triangle is read) and its float32 right-hand side at
 * b + b_offset[i]; its solution is written at x + b_offset[i] in x_dtype.
 * Systems are taken in the order given by order[0..count) (a permutation or
 * a subset of 0..count-1 of length `count`).
 *
 * flags & LXA_UNEVEN_PERSISTENT: one persistent launch sized to the device;
 *   workgroups pull systems from a device work queue in `order` (callers
 *   sort it largest-first for load balance).  Otherwise one workgroup per
 *   listed system.
 * max_refine > 0: after the FP32 factorisation and solve, FP64 residuals
 *   r = b - A x are formed and x (kept in FP64) is corrected with FP32
 *   solves until the normwise backward error
 *   eta = |r|_inf / (|A|_inf |x|_inf + |b|_inf) <= tol, at most max_refine
 *   residuals.  x_dtype must then be LXA_F64.
 *
 * Outputs per system: info (0 ok, k+1 non-positive pivot -> x is NaN, -1
 * order outside [1, max_n] -> system skipped),
 * iters (residuals formed), eta (last backward error, -1 if none formed).
 * iters and eta may be NULL.  Semantically each system is
 * lx.linear_solve(lx.MatrixLinearOperator(A_i, positive_semidefinite_tag),
 * b_i, lx.Cholesky()) (lineax/_solver/cholesky.py:43-78), evaluated for the
 * exact float64 value of the float32 data when max_refine > 0. */
#define LXA_UNEVEN_PERSISTENT 1
/* flags & LXA_UNEVEN_SORTED (with LXA_UNEVEN_PERSISTENT): `order` lists the
 * systems by n descending.  Each tile class's launch then takes its own
 * contiguous range of the order instead of sweeping all of it.  The claim is
 * checked on the device first; an order that is not sorted is swept. */
#define LXA_UNEVEN_SORTED 2
/* Bytes of workspace the persistent launch needs (three work-queue counters
 * and a flag: 16 bytes whatever the count); lxa_chol_uneven_solve refuses a
 * smaller workspace with LXA_EINVAL. */
int lxa_chol_uneven_workspace_size(int64_t count, size_t* bytes);
int lxa_chol_uneven_solve(void* stream, int64_t count, int max_n,
                          const int32_t* n, const int64_t* a_offset,
                          const float* a, const int64_t* b_offset,
                          const float* b, void* x, int x_dtype,
                          const int32_t* order, int flags, int max_refine,
                          double tol, int32_t* info, int32_t* iters,
                          double* eta, void* workspace,
                          size_t workspace_bytes);

/* The autotuner's hook (no counterpart in the reference):
 * lxa_chol_uneven_solve with the tile-class boundaries given per call
 * instead of taken from the library's autotune table: orders <= max4
 * (<= 64) on 4 x 4 blocks of one wave, <= max8 (<= 128) on 8 x 8 blocks of
 * four waves, larger on 16 x 16 blocks.  lineax_amd/tune_uneven.py times each
 * class per order bucket with it and writes the table;
 * lxa_chol_uneven_tuned_classes reports the table the library was built
 * with. */
int lxa_chol_uneven_tune_solve(void* stream, int64_t count, int max_n,
                               const int32_t* n, const int64_t* a_offset,
                               const float* a, const int64_t* b_offset,
                               const float* b, void* x, int x_dtype,
                               const int32_t* order, int flags, int max_refine,
                               double tol, int32_t* info, int32_t* iters,
                               double* eta, void* workspace,
                               size_t workspace_bytes, int max4, int max8);
int lxa_chol_uneven_tuned_classes(int* max4, int* max8);

/* lxa_chol_uneven_solve without a workspace argument, for callers that
 * cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  It has no counterpart in the
 * reference.  Same arguments, semantics and results as lxa_chol_uneven_solve;
 * the work-queue counters live in a 256-byte device array inside the
 * library (one per GPU, held until the process ends), whose address is
 * looked up on the first call on each device.  Calls allocate nothing (no
 * hipMalloc, hipFree or allocating library routine) and refuse a request
 * whose counters would exceed that array (LXA_EINVAL).  Each persistent
 * launch clears its counter with a fill kernel on `stream` first.  All calls
 * on one GPU share the array: calls that overlap on one GPU (different
 * streams, or host threads without ordering) race, and must be serialised
 * by the caller. */
int lxa_chol_uneven_solve_noalloc(void* stream, int64_t count, int max_n,
                                  const int32_t* n, const int64_t* a_offset,
                                  const float* a, const int64_t* b_offset,
                                  const float* b, void* x, int x_dtype,
                                  const int32_t* order, int flags,
                                  int max_refine, double tol, int32_t* info,
                                  int32_t* iters, double* eta);

/* ------------------------------------------------------------------------
 * Operators and result mapping
 * ---------------------------------------------------------------------- */

/* y = A x (trans = 0, A is [m, n]) or y = A^T x (trans = 1), batched,
 * FP64 accumulation.  Replaces jnp.matmul(self.matrix, vector,
 * precision=HIGHEST) in MatrixLinearOperator.mv, lineax/_operator.py:265-269
 * (trans = 1 serves the transposed operator, _operator.py:274-277, on the
 * same storage).  trans = 1 splits the rows over workgroups and adds the
 * per-chunk sums in a fixed order; it needs lxa_gemv_workspace_size bytes
 * of device scratch (0 for trans = 0 and for small problems). */
int lxa_gemv_workspace_size(int dtype, int64_t batch, int64_t m, int64_t n, int trans, size_t* bytes);
int lxa_gemv(void* stream, int dtype, int64_t batch, int64_t m, int64_t n,
             const void* a, int64_t a_batch_stride, const void* x,
             int64_t x_batch_stride, void* y, int64_t y_batch_stride,
             int trans, void* work, size_t work_bytes);

/* The RESULTS post-processing of lineax/_solve.py:104-123, per system:
 * result[i] == successful and any non-finite entry in solution row i
 * -> singular; then result[i] == singular and any non-finite entry in
 * vector row i -> nonfinite_input.  solution/vector are [rows, len]. */
int lxa_solve_results(void* stream, int dtype, int64_t rows, int64_t len,
                      const void* solution, int64_t solution_row_stride,
                      const void* vector, int64_t vector_row_stride,
                      int32_t* result);


/* ------------------------------------------------------------------------
 * Grid operators (matrix-free FunctionLinearOperator.mv of the Krylov
 * workloads, lineax/_operator.py:664-732)
 * ---------------------------------------------------------------------- */

/* y = A x for the five-point stencil with constant coefficients
 * coeffs = {center, west (j-1), east (j+1), south (i-1), north (i+1)} on an
 * nx x ny grid (j fastest) with zero Dirichlet boundary, for `rows` grid
 * functions.  in/out dtypes may differ (FP64 accumulation).  x and y must
 * not overlap (rows are read ahead of the rows written). */
int lxa_stencil5_apply(void* stream, int in_dtype, int out_dtype, int64_t rows, int64_t nx, int64_t ny,
                       const double* coeffs, const void* x, int64_t x_row_stride, void* y,
                       int64_t y_row_stride);

/* ------------------------------------------------------------------------
 * Krylov building blocks (lineax/_solver/cg.py, lineax/_solver/gmres.py)
 *
 * A batch of right-hand sides is `rows` vectors of length n, row-major.
 * Each row has its own device state; finished rows are frozen (lineax under
 * jax.vmap).  The operator's matvec is applied by the caller between these
 * calls (it can be any operator: the GEMV kernel, a stencil kernel, or user
 * code).  Reductions are FP64, two-phase and deterministic.  `ydtype` is
 * the dtype of y and b; `vdtype` that of the working vectors.  Supported
 * pairs (y, vectors): (F64, F64), (F32, F32), (F64, F32).
 * ---------------------------------------------------------------------- */

typedef struct {
  double gamma;     /* CG: <z, r>                                       */
  double alpha;     /* CG: step length (NaN when the breakdown guard fires) */
  double beta;      /* CG: gamma_new / gamma                            */
  double ip;        /* CG: <A p, p>                                     */
  double norm_r;    /* solver norm of r / (atol + rtol |b|)             */
  double norm_diff; /* solver norm of diff / (atol + rtol |y|)          */
  int32_t step;     /* CG iterations / GMRES cycles incl. the first pass */
  int32_t active;   /* 1 while the row's loop condition holds           */
  int32_t result;   /* RESULTS code, written by lxa_*_result            */
  int32_t pad;
} lxa_krylov_row; /* 64 bytes */

typedef struct {
  double r_norm;    /* |r|_2 at the start of the cycle                  */
  double step_norm; /* |A v_j|_2 of the current Arnoldi step             */
  double scale;     /* 1/|.| of the last normalised vector (0 on breakdown) */
  double r_min;     /* smallest residual norm seen (stagnation test)     */
  int32_t inner;    /* Arnoldi steps done in this cycle                  */
  int32_t columns;  /* basis columns written in this cycle               */
  int32_t breakdown;       /* breakdown in this cycle                    */
  int32_t inner_done;      /* inner loop finished for this row           */
  int32_t carry_breakdown; /* breakdown of the last cycle                */
  int32_t deferred;        /* breakdown of the cycle before (gmres.py:145-158) */
  int32_t stagnation;      /* cycles without a decrease of the residual  */
  int32_t pad;
} lxa_gmres_row; /* 64 bytes */

typedef struct {
  double rtol, atol;
  double rcond;        /* CG guard: 2 eps(dtype) size (lineax/_misc.py:30-38) */
  double eps;          /* GMRES breakdown scale: eps(dtype)                   */
  int64_t max_steps;   /* 10 * size when max_steps_given == 0                  */
  int32_t max_steps_given;
  int32_t norm_kind;   /* 0 max_norm (default), 1 two_norm, 2 rms_norm        */
  int32_t restart;     /* GMRES: min(restart, size), at most 128              */
  int32_t stagnation_iters;
} lxa_krylov_params;

/* Bytes of FP64 partial-sum scratch for a batch (GMRES: pass its restart). */
int lxa_krylov_partials_size(int64_t rows, int64_t n, int restart, size_t* bytes);

/* CG, one iteration = alpha, update (or update + residual on stabilising
 * steps), [preconditioner gamma], finish.  Replaces the lax.while_loop of
 * lineax/_solver/cg.py:162-211; the result mapping is cg.py:213-222. */
int lxa_cg_begin(void* stream, int vdtype, int64_t rows, int64_t n, const void* z, const void* r,
                 const lxa_krylov_params* prm, lxa_krylov_row* state, double* partials);
int lxa_cg_alpha(void* stream, int vdtype, int64_t rows, int64_t n, const void* q, const void* p,
                 const lxa_krylov_params* prm, lxa_krylov_row* state, double* partials);
int lxa_cg_update(void* stream, int ydtype, int vdtype, int64_t rows, int64_t n, const void* p, const void* q,
                  void* y, void* r, const void* b, int stable, int fuse_gamma, const lxa_krylov_params* prm,
                  lxa_krylov_row* state, double* partials);
int lxa_cg_residual(void* stream, int ydtype, int vdtype, int64_t rows, int64_t n, const void* b, const void* t,
                    void* r, int fuse_gamma, const lxa_krylov_params* prm, lxa_krylov_row* state,
                    double* partials);
int lxa_cg_precond_gamma(void* stream, int ydtype, int vdtype, int64_t rows, int64_t n, const void* b,
                         const void* z, const void* r, const lxa_krylov_params* prm, lxa_krylov_row* state,
                         double* partials);
int lxa_cg_finish(void* stream, int vdtype, int64_t rows, int64_t n, const void* z, void* p,
                  const lxa_krylov_params* prm, lxa_krylov_row* state, double* partials);
int lxa_cg_result(void* stream, int64_t rows, const lxa_krylov_params* prm, lxa_krylov_row* state);

/* ---------------------------------------------------------------------------
 * LSMR (lineax/_solver/lsmr.py:94-359, damp 0, two_norm) on device-resident
 * per-row state, so rows stop independently.  u and v are stored unscaled:
 * the bidiagonalisation's unit vectors are u / beta and v / alpha (beta,
 * alpha of the row state; a zero norm divides by 1, as lsmr.py does).
 * Sequence: [caller: t = A x0] begin (u = b - t, beta, |b|), [caller:
 * z = A^T u] init (v, alpha, the recurrence's initial scalars, h = v /
 * alpha, hbar = 0); per iteration: [caller: w = A v] step_u (u = w / alpha
 * - alpha u / beta, the new beta), [caller: z = A^T u] step_v (when beta is
 * nonzero v = z / beta - beta v / alpha and the new alpha; the Givens
 * recurrence of lsmr.py:240-312 and the estimates of |r|, |A|, cond(A),
 * |A^T r|), step_x (hbar, x, h updates and |x|, then the stopping rules of
 * lsmr.py:316-329); finally result (lsmr.py:344-357).  Partials: as
 * lxa_krylov_partials_size(rows, max(m, n), 0).
 * ------------------------------------------------------------------------- */
typedef struct {
  double alpha, beta;
  double zetabar, alphabar, rho, rhobar, cbar, sbar;
  double betadd, betad, rhodold, tautildeold, thetatilde, zeta, delta;
  double normA2, maxrbar, minrbar, condA, normr, normAr, normb, normx, normA;
  double c_hbar, c_x, c_h; /* coefficients of this step's vector updates */
  int32_t itn, istop, active, result;
} lxa_lsmr_row; /* 232 bytes */

typedef struct {
  double rtol, atol, conlim;
  int64_t max_steps; /* min(m, n) * 10 when max_steps_given == 0 */
  int32_t max_steps_given, pad;
} lxa_lsmr_params;

int lxa_lsmr_begin(void* stream, int ydtype, int vdtype, int64_t rows, int64_t m, const void* b, const void* t,
                   void* u, const lxa_lsmr_params* prm, lxa_lsmr_row* state, double* partials);
int lxa_lsmr_init(void* stream, int vdtype, int64_t rows, int64_t n, const void* z, void* v, void* h, void* hbar,
                  const lxa_lsmr_params* prm, lxa_lsmr_row* state, double* partials);
int lxa_lsmr_step_u(void* stream, int vdtype, int64_t rows, int64_t m, const void* w, void* u,
                    const lxa_lsmr_params* prm, lxa_lsmr_row* state, double* partials);
int lxa_lsmr_step_v(void* stream, int vdtype, int64_t rows, int64_t n, const void* z, void* v,
                    const lxa_lsmr_params* prm, lxa_lsmr_row* state, double* partials);
int lxa_lsmr_step_x(void* stream, int ydtype, int vdtype, int64_t rows, int64_t n, const void* v, void* h,
                    void* hbar, void* x, const lxa_lsmr_params* prm, lxa_lsmr_row* state, double* partials);
int lxa_lsmr_result(void* stream, int64_t rows, const lxa_lsmr_params* prm, lxa_lsmr_row* state);

/* ---------------------------------------------------------------------------
 * Thin singular value decomposition of a batch of dense matrices, A = U
 * diag(s) V^T with s descending (jax.scipy.linalg.svd(full_matrices=False)
 * in lineax/_solver/svd.py:51), by one-sided Jacobi rotations: one
 * workgroup per matrix.  a: batch matrices m x n row-major (a_stride
 * elements apart); u: [batch][m][k], s: [batch][k], vt: [batch][k][n],
 * k = min(m, n) <= 512; work: lxa_svd_workspace bytes; sweeps[b] (may be
 * NULL) receives the Jacobi sweeps used, or max_sweeps + 1 when the last
 * allowed sweep still rotated (not converged).  Singular values are sorted
 * descending with NaN above every number and ties by column, so every
 * output slot is written.
 * ------------------------------------------------------------------------- */
int lxa_svd_workspace(int dtype, int64_t batch, int64_t m, int64_t n, size_t* bytes);
int lxa_svd(void* stream, int dtype, int64_t batch, int64_t m, int64_t n, const void* a, int64_t a_stride, void* u,
            void* s, void* vt, void* work, int max_sweeps, int32_t* sweeps);

/* BiCGStab with right preconditioning (lineax/_solver/bicgstab.py:78-205;
 * its body 143-167, cond_fun 121-140, RESULTS 185-197).  Sequence:
 * [caller: t = A y0] begin (r = r0 = b - t, the initial carry alpha = omega
 * = rho = 1); per iteration: direction (p = r + beta (p - omega v); first = 1
 * on the first iteration, where p = r), [caller: x = M p, v = A x], alpha,
 * s (s = r - alpha v), [caller: z = M s, t = A z], omega, update (y += alpha
 * x + omega z, r = s - omega t, step, tolerance norms and cond_fun); finally
 * result.  Row state: gamma = rho, alpha, ip = omega, beta = <r0, r> of the
 * current residual.  prm->eps selects the breakdown test: 0 tests for exact
 * zeros of omega, alpha, rho (lineax with jax_enable_x64), > 0 for values
 * below it (lineax's 1e-16 otherwise).  Partials as lxa_krylov_partials_size
 * with restart 0. */
int lxa_bicgstab_begin(void* stream, int ydtype, int vdtype, int64_t rows, int64_t n, const void* b, const void* t,
                       void* r, void* r0, const lxa_krylov_params* prm, lxa_krylov_row* state, double* partials);
int lxa_bicgstab_direction(void* stream, int vdtype, int64_t rows, int64_t n, const void* r, void* p, const void* v,
                           int first, lxa_krylov_row* state);
int lxa_bicgstab_alpha(void* stream, int vdtype, int64_t rows, int64_t n, const void* r0, const void* v,
                       const lxa_krylov_params* prm, lxa_krylov_row* state, double* partials);
int lxa_bicgstab_s(void* stream, int vdtype, int64_t rows, int64_t n, const void* r, const void* v, void* s,
                   lxa_krylov_row* state);
int lxa_bicgstab_omega(void* stream, int vdtype, int64_t rows, int64_t n, const void* s, const void* t,
                       const lxa_krylov_params* prm, lxa_krylov_row* state, double* partials);
int lxa_bicgstab_update(void* stream, int ydtype, int vdtype, int64_t rows, int64_t n, const void* x, const void* z,
                        const void* s, const void* t, const void* r0, void* y, void* r, const void* b,
                        const lxa_krylov_params* prm, lxa_krylov_row* state, double* partials);
int lxa_bicgstab_result(void* stream, int64_t rows, const lxa_krylov_params* prm, lxa_krylov_row* state);

/* Batched FP64 products with float32 operators of uneven shapes, for FP64
 * refinement residuals: per system, trans = 0: y = b - A x (y = A x when b
 * is 0), trans = 1: y = A^T x.  A float32 row-major with leading dimension
 * lda; x, y float64; b float32.  Addresses are device pointers.  lineax
 * does not refine, so this has no single counterpart there: it is the matmul
 * of lineax/_operator.py:265-269 evaluated as an FP64 residual, for the
 * request loop's refinement. */
typedef struct {
  int64_t a, x, b, y;
  int32_t m, n, lda, pad;
} lxa_gemv_desc; /* 48 bytes */
int lxa_gemv_uneven(void* stream, int64_t count, const lxa_gemv_desc* descs, int trans, int max_m, int max_n);

/* Fused CG for a constant-coefficient five-point stencil operator (the
 * matrix-free FunctionLinearOperator of lineax/_operator.py:664-732 whose mv
 * is lxa_stencil5_apply) with the identity preconditioner.  One iteration of
 * lineax/_solver/cg.py:169-206 is two launches:
 *   lxa_cg_stencil_direction: p_new = r + beta p_old (p_new = r when first),
 *     <A p_new, p_new> and alpha with the rcond guard (cg.py:172-178);
 *   lxa_cg_stencil_update: y += alpha p_new; r -= alpha A p_new, or, when
 *     stable, r = b - A y (cg.py:187-198); gamma, beta, step, both stopping
 *     norms and the loop condition (cg.py:150-167, 202-206).
 * The per-row scalars are finished on the device by a one-block-per-row
 * kernel after each sweep.  Rows are grid functions [rows][nx][ny] (row
 * stride nx*ny); coeffs = (center, west, east, south, north) as in
 * lxa_stencil5_apply.  p_new must not alias r or p_old (ping-pong two
 * direction buffers; p_old is read even when first, so pass a valid buffer).
 * Stabilising steps read y around each tile, so they write y_new to
 * y_out != y_in; plain steps pass y_out == y_in (in place); a finished row's
 * iterate is copied across.  b may be stored in float32 (bdtype) when that is
 * exact.  Dtypes (y, vectors, b): (f64, f32, f32|f64), (f64, f64, f64), (f32,
 * f32, f32).  Workspace: FP64 partials of the size lxa_cg_stencil_workspace
 * returns. */
int lxa_cg_stencil_workspace(int64_t rows, int64_t nx, int64_t ny, size_t* partial_bytes);
int lxa_cg_stencil_direction(void* stream, int vdtype, int64_t rows, int64_t nx, int64_t ny, const double* coeffs,
                             const void* r, const void* p_old, void* p_new, int first,
                             const lxa_krylov_params* prm, lxa_krylov_row* state, double* partials);
int lxa_cg_stencil_update(void* stream, int ydtype, int vdtype, int bdtype, int64_t rows, int64_t nx, int64_t ny,
                          const double* coeffs, const void* p, const void* y_in, void* y_out, void* r,
                          const void* b, int stable,
                          const lxa_krylov_params* prm, lxa_krylov_row* state, double* partials);

/* GMRES: init, then per restart cycle: cycle_begin (normalise r into basis
 * column 0), for j < restart: [caller: w = M A v_j] arnoldi(j), then update
 * (Hessenberg least squares by Householder QR, y += V z), [caller:
 * t = A y], residual (r = b - t, then optionally the caller preconditions
 * r and calls residual again with write = 0 for the norms), cycle_end.  The
 * first cycle is lineax's dummy pass: skip cycle_begin/arnoldi/update and
 * call cycle_end with first_pass = 1.  Replaces gmres.py:143-320 and the
 * result mapping gmres.py:217-237.  Basis: [rows][restart + 1][n];
 * hessenberg: [rows][restart][restart + 1] FP64; proj: [rows][restart + 1];
 * z: [rows][restart]. */
int lxa_gmres_init(void* stream, int64_t rows, lxa_krylov_row* state, lxa_gmres_row* gstate);
int lxa_gmres_cycle_begin(void* stream, int vdtype, int64_t rows, int64_t n, const void* r, void* basis,
                          const lxa_krylov_params* prm, lxa_krylov_row* state, lxa_gmres_row* gstate,
                          double* hessenberg, double* partials);
int lxa_gmres_arnoldi(void* stream, int vdtype, int64_t rows, int64_t n, int j, void* basis, const void* w,
                      const lxa_krylov_params* prm, lxa_krylov_row* state, lxa_gmres_row* gstate,
                      double* hessenberg, double* proj, double* partials);
int lxa_gmres_update(void* stream, int ydtype, int vdtype, int64_t rows, int64_t n, const void* basis, void* y,
                     const lxa_krylov_params* prm, lxa_krylov_row* state, lxa_gmres_row* gstate,
                     const double* hessenberg, double* z, double* partials);

/* As lxa_gmres_update, for a basis stored as columns v_k / colscale[k]
 * ([rows][restart + 1] FP64, lxa_gmres_stencil_step's layout). */
int lxa_gmres_update_scaled(void* stream, int ydtype, int vdtype, int64_t rows, int64_t n, const void* basis,
                            void* y, const lxa_krylov_params* prm, lxa_krylov_row* state, lxa_gmres_row* gstate,
                            const double* hessenberg, double* z, const double* colscale, double* partials);

/* Fused Arnoldi steps for a constant-coefficient five-point stencil operator
 * (coeffs as in lxa_stencil5_apply) with no preconditioner, grids of width
 * ny <= 1024, restart <= 32.  Replaces lxa_gmres_arnoldi's classical
 * Gram-Schmidt step (lineax/_solver/gmres.py:355-398) with one sweep per
 * step that reads the basis once:
 *   mode 0, j = 0: t = A v_0, its projections;
 *   mode 1, 1 <= j < restart: u_j = w'_{j-1} (the orthogonalised previous
 *     step), stored unnormalised as basis column j; t = A u_j; the
 *     projections of step j.  Closes step j-1 (|w'|, breakdown test,
 *     Hessenberg row j-1);
 *   mode 2, j = restart: closes the last step.
 * w_in / w_out: the raw previous / next w ([rows][nx*ny], ping-pong).
 * coefs ([rows][33] FP64), colscale ([rows][restart + 1] FP64), proj
 * ([rows][restart + 1] FP64) carry the step to step state; basis, the
 * Hessenberg buffer and the row states are lxa_gmres_*'s.  Begin a cycle
 * with lxa_gmres_cycle_begin, end it with lxa_gmres_update_scaled. */
int lxa_gmres_stencil_workspace(int64_t rows, int64_t nx, int64_t ny, size_t* partial_bytes, size_t* coef_bytes);
int lxa_gmres_stencil_step(void* stream, int vdtype, int64_t rows, int64_t nx, int64_t ny, const double* coeffs,
                           int mode, int j, const lxa_krylov_params* prm, void* basis, const void* w_in, void* w_out,
                           double* coefs, double* colscale, double* proj, double* hessenberg, lxa_krylov_row* state,
                           lxa_gmres_row* gstate, double* partials);
int lxa_gmres_residual(void* stream, int ydtype, int vdtype, int64_t rows, int64_t n, const void* b,
                       const void* t, void* r, int write, const lxa_krylov_params* prm, lxa_krylov_row* state,
                       double* partials);
int lxa_gmres_cycle_end(void* stream, int64_t rows, int64_t n, int first_pass, const lxa_krylov_params* prm,
                        lxa_krylov_row* state, lxa_gmres_row* gstate, double* partials);
int lxa_gmres_result(void* stream, int64_t rows, const lxa_krylov_params* prm, lxa_krylov_row* state,
                     const lxa_gmres_row* gstate);

/* ---------------------------------------------------------------------------
 * Matrix-core update C <- C - A B (float32 storage, row-major), the
 * Schur-complement update of the blocked LU (the trailing-matrix GEMM inside
 * LAPACK getrf that jax.scipy.linalg.lu_factor reaches from
 * lineax/_solver/lu.py:53).  mode selects the product precision: float32,
 * bfloat16 / float16 inputs rounded to nearest, or split bfloat16 (bf16x3);
 * FP32 accumulation.
 * ------------------------------------------------------------------------- */
#define LXA_GEMM_F32 0
#define LXA_GEMM_BF16 1
#define LXA_GEMM_F16 2
#define LXA_GEMM_BF16X3 3 /* split bf16: hi*hi + hi*lo + lo*hi, ~16-bit operands */
int lxa_gemm_sub(void* stream, int mode, int64_t m, int64_t n, int64_t k, const float* a, int64_t lda,
                 const float* b, int64_t ldb, float* c, int64_t ldc);
/* C_s = A_s B_s for `batch` systems (< 65536; strides in elements, 0
 * broadcasts an operand), same tiles and modes.  Replaces the materialised
 * product of lineax's Normal solver (lineax/_solver/normal.py:105-120 via
 * ComposedLinearOperator.as_matrix, _operator.py:1125-1192) for float32
 * operators: the Gram matrix A^T A (A A^T when wide).  With LXA_GEMM_F32,
 * mode may add LXA_GEMM_TRANS_A (A_s stored k x m: element (i, l) at
 * a[l * lda + i], lda >= m) and / or LXA_GEMM_TRANS_B (B_s stored n x k:
 * element (l, j) at b[j * ldb + l], ldb >= k), so a transposed operand is
 * read in place, without a copy. */
#define LXA_GEMM_TRANS_A 256
#define LXA_GEMM_TRANS_B 512
int lxa_gemm_batched(void* stream, int mode, int64_t batch, int64_t m, int64_t n, int64_t k,
                     const float* a, int64_t lda, int64_t a_stride, const float* b,
                     int64_t ldb, int64_t b_stride, float* c, int64_t ldc,
                     int64_t c_stride);

/* ---------------------------------------------------------------------------
 * Large dense LU (one system of order n in the thousands), the factor and
 * solve behind lineax/_solver/lu.py:53 (lu_factor) and :64 (lu_solve) for
 * large operators.  The system is randomised by a two-sided depth-2 random
 * butterfly transform A' = U^T A V (lxa_rbt_matrix; n divisible by 4: pad
 * with an identity block), A' is factored with pivoting only inside each
 * 128 x 128 diagonal block, and
 * solutions are refined against A by the caller (GMRES preconditioned with
 * these factors, FP64 residuals from lxa_gemv_f32a).
 *
 * lxa_lu_large_factor: blocked LU in place, P A = L U, each 128-column
 *   block column pivoting over its diagonal block and the 128 rows below it
 *   (so every multiplier is bounded by 1, as with partial pivoting, but no
 *   device-wide arg-max is needed; L unit lower below the diagonal, U on
 *   and above); piv[i] (n entries, 0-based, i <= piv[i] < i + 256) is the
 *   row swapped with row i, as getrf records it; linv / uinv receive the inverses of the
 *   128 x 128 diagonal blocks of L and U (ceil(n/128) each, row-major, dense);
 *   mode (LXA_GEMM_*) is the precision of the outer Schur-complement
 *   products (panels are float32).  info[0] (zeroed by the caller) receives
 *   1 + the index of the first exactly-zero pivot.
 * lxa_lu_large_trsv: kind 0: L y = x, 1: U y = x, 2: U^T y = x, 3: L^T y = x
 *   for nrhs <= 8 vectors x[r * ldx + i] in place; inv is linv (kinds 0, 3)
 *   or uinv (kinds 1, 2); sync is 16-byte aligned caller scratch of
 *   lxa_lu_large_trsv_workspace(n) int32 words (cleared by the call; word 1
 *   is set if a block gave up waiting for the block it depends on).  One
 *   persistent launch; the solved blocks pass between workgroups through
 *   the workspace as (tag, value) words.
 * lxa_rbt_matrix: A <- U^T A V with du, dv of 2n diagonal entries each
 *   (outer level d[0..n), inner level d[n..2n)).
 * lxa_rbt_vectors: form 1 applies U^T (or V^T) to nrhs vectors, form 0
 *   applies V (or U) with the given diagonals.
 * lxa_gemv_f32a: y = A x (b == NULL) or y = b - A x with float32 A [m, k],
 *   x / b / y float32 or float64 (nrhs vectors, row strides ld*), FP64
 *   accumulation: the refinement residual and the GMRES products.
 * ------------------------------------------------------------------------- */
int lxa_lu_large_block(void);
int lxa_lu_large_factor(void* stream, int mode, int64_t n, float* a, int64_t lda, float* linv, float* uinv,
                        int32_t* piv, int32_t* info);
/* The factorisation's row permutation: perm[i] = the original row at
 * position i once the swap sequence piv has been applied in order (getrf's
 * row interchanges, laswp, composed once).  piv must be windowed as
 * lxa_lu_large_factor writes it: piv[j] in [j, j + 256).  lxa_lu_large_permute then
 * applies it to nrhs vectors x[r * ldx + i]: inverse = 0 gives y = P x
 * (before the L solve), inverse = 1 gives y = P^T x (after the L^T solve of
 * a transposed system); y must not alias x. */
int lxa_lu_large_perm(void* stream, int64_t n, const int32_t* piv, int32_t* perm);
int lxa_lu_large_permute(void* stream, int64_t n, int nrhs, const float* x, int64_t ldx, float* y, int64_t ldy,
                         const int32_t* perm, int inverse);
int lxa_lu_large_trsv_workspace(int64_t n, int64_t* words);
int lxa_lu_large_trsv(void* stream, int kind, int64_t n, int nrhs, const float* lu, int64_t lda, const float* inv,
                      float* x, int64_t ldx, int32_t* sync);
int lxa_rbt_matrix(void* stream, int64_t n, float* a, int64_t lda, const float* du, const float* dv);
int lxa_rbt_vectors(void* stream, int dtype, int64_t n, int nrhs, void* x, int64_t ldx, const float* d, int form);
int lxa_gemv_f32a(void* stream, int64_t m, int64_t k, int nrhs, const float* a, int64_t lda, int x_dtype,
                  const void* x, int64_t ldx, const void* b, int64_t ldb, int y_dtype, void* y, int64_t ldy);

/* ---------------------------------------------------------------------------
 * Householder QR over an uneven batch, and triangular solves.
 *
 * Replaces the arithmetic of lineax/_solver/qr.py: `jnp.linalg.qr(mode="raw")`
 * (qr.py:62, LAPACK geqrf), `jll.ormqr` (qr.py:89, 90) and
 * `jsp.linalg.solve_triangular` (qr.py:83, 91; lineax/_solver/triangular.py).
 *
 * Each system is described by one lxa_qr_desc in device memory.  The matrix
 * is row-major m x n at a + a_offset with row stride lda.  After
 * lxa_qr_factor it holds R on and above the diagonal and the Householder
 * vectors below it (unit leading entry implied), as geqrf's raw output does
 * (transposed to row-major); tau receives min(m, n) scalars at tau_offset
 * and tblk the compact-WY blocks: for panel p (columns [p*nb, p*nb + nb)),
 * an upper-triangular nb x nb T_p at t_offset + p*nb*nb, row-major, with
 * H_{p*nb} ... H_{p*nb+nb-1} = I - V_p T_p V_p^T.  nb = lxa_qr_block(dtype).
 * w_offset addresses lxa_qr_work_elems() elements of scratch per system.
 * Right-hand sides for lxa_qr_apply / lxa_trsv: nrhs vectors per system at
 * b + b_offset + r*ldb.
 * ------------------------------------------------------------------------- */
typedef struct {
  int64_t a_offset;
  int64_t tau_offset;
  int64_t t_offset;
  int64_t w_offset;
  int64_t b_offset;
  int32_t m, n, lda, ldb;
} lxa_qr_desc; /* 56 bytes */

/* Panel width of the blocked factorisation: 16 for float32, 8 for float64. */
int lxa_qr_block(int dtype);
/* Scratch elements one system of order m x n needs in lxa_qr_factor. */
int lxa_qr_work_elems(int dtype, int64_t m, int64_t n, int64_t* elems);
/* geqrf of every system (qr.py:62).  max_m / max_n / max_k bound m, n and
 * min(m, n) over the batch (they size the launch grids). */
int lxa_qr_factor(void* stream, int dtype, int64_t count, const lxa_qr_desc* desc, int max_m, int max_n,
                  int max_k, void* a, void* tau, void* tblk, void* work);
/* Blocked Cholesky over an uneven batch of float32 SPD systems of order
 * n <= 1024 (the lxa_qr_desc fields used: a_offset, b_offset, n, lda), one
 * workgroup per system: the cho_factor / cho_solve pair of
 * lineax/_solver/cholesky.py:60, 74 for batches whose orders differ.
 * factor: upper triangle in place (the strict lower triangle is left as is);
 * info[s] = k + 1 at a non-positive pivot (factor then NaN), -1 if n > 1024.
 * solve: x = (U^T U)^-1 b in place at b + b_offset (one vector per system). */
int lxa_chol_blocked_factor(void* stream, int64_t count, const lxa_qr_desc* desc, float* a, int32_t* info);
int lxa_chol_blocked_solve(void* stream, int64_t count, const lxa_qr_desc* desc, const float* u, float* b);

/* One panel step of lxa_qr_factor for each system at its own column
 * kcol[s] (a multiple of lxa_qr_block(dtype); systems with kcol[s] >=
 * min(m, n) are skipped): the continuous-batching form, where systems that
 * joined the batch at different times advance together.  max_rows >= m - kcol
 * and max_trail_cols >= n - kcol - 1 over the systems with columns left.
 * The caller advances kcol by the block after the call. */
int lxa_qr_factor_step(void* stream, int dtype, int64_t count, const lxa_qr_desc* desc, const int32_t* kcol,
                       int max_rows, int max_trail_cols, void* a, void* tau, void* tblk, void* work);
/* b <- Q^T b (trans = 1, ormqr with transpose=True, qr.py:89) or b <- Q b
 * (trans = 0, the min-norm branch's ormqr, qr.py:86) for vectors of length m. */
int lxa_qr_apply(void* stream, int dtype, int64_t count, const lxa_qr_desc* desc, int max_m, int max_k,
                 const void* a, const void* tblk, int nrhs, void* b, int trans);
/* In-place triangular solve with the leading k x k triangle of a, k =
 * min(m, n) (solve_triangular, qr.py:83, 91; triangular.py:81-83):
 * lower = 0 reads the upper triangle, 1 the lower; trans = 1 solves with the
 * transpose; unit_diag = 1 takes the diagonal as ones.  The first k entries
 * of each right-hand side are overwritten.  Supports k <= 8192. */
int lxa_trsv(void* stream, int dtype, int64_t count, const lxa_qr_desc* desc, int max_k, const void* a, int nrhs,
             void* b, int lower, int trans, int unit_diag);

#ifdef __cplusplus
}
#endif

#endif /* LINEAX_AMD_H_ */
