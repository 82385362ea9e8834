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
int lxa_chol_uneven_workspace_size(int64_t count, size_t* bytes);
int lxa_chol_uneven_solve(void* stream, int64_t count, int max_n,
                          const int32_t* n, const int64_t* a_offset,
                          const float* a, const int64_t* b_offset,
                          const float* b, void* x, int x_dtype,
                          const int32_t* order, int flags, int max_refine,
                          double tol, int32_t* info, int32_t* iters,
                          double* eta, void* workspace,
                          size_t workspace_bytes);

/* ------------------------------------------------------------------------
 * Operators and result mapping
 * ---------------------------------------------------------------------- */

/* y = A x (trans = 0, A is [m, n]) or y = A^T x (trans = 1), batched.
 * Replaces jnp.matmul(self.matrix, vector, precision=HIGHEST) in
 * MatrixLinearOperator.mv, lineax/_operator.py:265-269. */
int lxa_gemv(void* stream, int dtype, int64_t batch, int64_t m, int64_t n,
             const void* a, int64_t a_batch_stride, const void* x,
             int64_t x_batch_stride, void* y, int64_t y_batch_stride,
             int trans);

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
 * functions.  in/out dtypes may differ (FP64 accumulation). */
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
int lxa_gmres_residual(void* stream, int ydtype, int vdtype, int64_t rows, int64_t n, const void* b,
                       const void* t, void* r, int write, const lxa_krylov_params* prm, lxa_krylov_row* state,
                       double* partials);
int lxa_gmres_cycle_end(void* stream, int64_t rows, int64_t n, int first_pass, const lxa_krylov_params* prm,
                        lxa_krylov_row* state, lxa_gmres_row* gstate, double* partials);
int lxa_gmres_result(void* stream, int64_t rows, const lxa_krylov_params* prm, lxa_krylov_row* state,
                     const lxa_gmres_row* gstate);

#ifdef __cplusplus
}
#endif

#endif /* LINEAX_AMD_H_ */
