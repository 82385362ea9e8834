/*
 * ORACLE — test infrastructure only.  CPU restatement, in double precision,
 * of the dense LAPACK algorithms that lineax reaches through JAX on its CPU
 * backend.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg may load this library; the product path never does.
 *
 * Reference call sites (patrick-kidger/lineax):
 *   lineax/_solver/lu.py:53       jax.scipy.linalg.lu_factor  -> getrf
 *   lineax/_solver/lu.py:64       jax.scipy.linalg.lu_solve   -> getrs
 *   lineax/_solver/cholesky.py:60 jax.scipy.linalg.cho_factor -> potrf ('U')
 *   lineax/_solver/cholesky.py:74 jax.scipy.linalg.cho_solve  -> potrs
 *   lineax/_solver/qr.py:62       jnp.linalg.qr(mode="raw")   -> geqrf
 *   lineax/_solver/qr.py:86,89    jax.lax.linalg.ormqr        -> ormqr
 *   lineax/_solver/qr.py:83,90    solve_triangular            -> trsv
 * The algorithms restated are LAPACK 3.x's unblocked dgetf2, dgetrs, dpotf2,
 * dpotrs, dgeqr2/dlarfg, dorm2r and dtrsv (published reference LAPACK).
 * Row-major storage throughout; pivots are 0-based interchange sequences as
 * jax.scipy.linalg.lu_factor returns them.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define A_(i, j) a[(int64_t)(i) * lda + (j)]

/* dgetf2: returns info (0, or k+1 for the first exactly-zero pivot).  The
 * pivot is the first entry of largest magnitude (idamax); the multipliers
 * are formed by division. */
int lxo_getrf(int n, double* a, int lda, int32_t* piv) {
  int info = 0;
  for (int k = 0; k < n; ++k) {
    int p = k;
    double best = fabs(A_(k, k));
    for (int i = k + 1; i < n; ++i) {
      double v = fabs(A_(i, k));
      if (v > best) {
        best = v;
        p = i;
      }
    }
    piv[k] = p;
    if (A_(p, k) != 0.0) {
      if (p != k) {
        for (int j = 0; j < n; ++j) {
          double t = A_(k, j);
          A_(k, j) = A_(p, j);
          A_(p, j) = t;
        }
      }
      double d = A_(k, k);
      for (int i = k + 1; i < n; ++i) A_(i, k) /= d;
    } else if (info == 0) {
      info = k + 1;
    }
    for (int i = k + 1; i < n; ++i) {
      double l = A_(i, k);
      if (l == 0.0) continue;
      for (int j = k + 1; j < n; ++j) A_(i, j) -= l * A_(k, j);
    }
  }
  return info;
}

/* dgetrs: trans 0 solves A x = b, trans 1 solves A^T x = b, in place. */
void lxo_getrs(int n, const double* a, int lda, const int32_t* piv, double* b, int trans) {
  if (trans == 0) {
    for (int k = 0; k < n; ++k) {
      int q = piv[k];
      if (q != k) {
        double t = b[k];
        b[k] = b[q];
        b[q] = t;
      }
    }
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < i; ++j) b[i] -= A_(i, j) * b[j];
    for (int i = n - 1; i >= 0; --i) {
      for (int j = i + 1; j < n; ++j) b[i] -= A_(i, j) * b[j];
      b[i] /= A_(i, i);
    }
  } else {
    for (int i = 0; i < n; ++i) {
      for (int j = 0; j < i; ++j) b[i] -= A_(j, i) * b[j];
      b[i] /= A_(i, i);
    }
    for (int i = n - 1; i >= 0; --i)
      for (int j = i + 1; j < n; ++j) b[i] -= A_(j, i) * b[j];
    for (int k = n - 1; k >= 0; --k) {
      int q = piv[k];
      if (q != k) {
        double t = b[k];
        b[k] = b[q];
        b[q] = t;
      }
    }
  }
}

/* dpotf2, uplo 'U': A = U^T U, reading only the upper triangle.  On a
 * non-positive or NaN pivot returns k+1 (LAPACK leaves the rest untouched;
 * JAX then replaces the whole factor by NaN, which the caller does). */
int lxo_potrf_upper(int n, double* a, int lda) {
  for (int j = 0; j < n; ++j) {
    double ajj = A_(j, j);
    for (int k = 0; k < j; ++k) ajj -= A_(k, j) * A_(k, j);
    if (!(ajj > 0.0)) {
      A_(j, j) = ajj;
      return j + 1;
    }
    ajj = sqrt(ajj);
    A_(j, j) = ajj;
    for (int i = j + 1; i < n; ++i) {
      double s = A_(j, i);
      for (int k = 0; k < j; ++k) s -= A_(k, j) * A_(k, i);
      A_(j, i) = s / ajj;
    }
  }
  return 0;
}

/* dpotrs, uplo 'U': solves U^T U x = b in place. */
void lxo_potrs_upper(int n, const double* a, int lda, double* b) {
  for (int i = 0; i < n; ++i) {
    for (int k = 0; k < i; ++k) b[i] -= A_(k, i) * b[k];
    b[i] /= A_(i, i);
  }
  for (int i = n - 1; i >= 0; --i) {
    for (int k = i + 1; k < n; ++k) b[i] -= A_(i, k) * b[k];
    b[i] /= A_(i, i);
  }
}

/* dlarfg: H (alpha; x) = (beta; 0), H = I - tau v v^T, v = (1; x_out).
 * Reference LAPACK also rescales when |beta| < safmin; inputs here are
 * well-scaled, so that branch is not restated. */
static void larfg(int m, double* alpha, double* x, int64_t incx, double* tau) {
  double xnorm = 0.0;
  for (int i = 0; i < m; ++i) xnorm = hypot(xnorm, x[i * incx]);
  if (xnorm == 0.0) {
    *tau = 0.0;
    return;
  }
  double beta = -copysign(hypot(*alpha, xnorm), *alpha);
  *tau = (beta - *alpha) / beta;
  double scal = 1.0 / (*alpha - beta);
  for (int i = 0; i < m; ++i) x[i * incx] *= scal;
  *alpha = beta;
}

/* dgeqr2 on an m x n matrix: R in the upper triangle, reflectors below. */
void lxo_geqrf(int m, int n, double* a, int lda, double* tau) {
  int k = m < n ? m : n;
  for (int i = 0; i < k; ++i) {
    larfg(m - i - 1, &A_(i, i), &A_(i + 1 < m ? i + 1 : i, i), lda, &tau[i]);
    if (i + 1 < n && tau[i] != 0.0) {
      double aii = A_(i, i);
      A_(i, i) = 1.0;
      for (int j = i + 1; j < n; ++j) {
        double w = 0.0;
        for (int r = i; r < m; ++r) w += A_(r, i) * A_(r, j);
        w *= tau[i];
        for (int r = i; r < m; ++r) A_(r, j) -= A_(r, i) * w;
      }
      A_(i, i) = aii;
    }
  }
}

/* dorm2r on a vector c of length m: trans 1 applies Q^T (H_0 first),
 * trans 0 applies Q (H_{k-1} first). */
void lxo_ormqr_vec(int m, int k, const double* a, int lda, const double* tau, double* c, int trans) {
  for (int s = 0; s < k; ++s) {
    int i = trans ? s : k - 1 - s;
    double w = c[i];
    for (int r = i + 1; r < m; ++r) w += A_(r, i) * c[r];
    w *= tau[i];
    c[i] -= w;
    for (int r = i + 1; r < m; ++r) c[r] -= A_(r, i) * w;
  }
}

/* dtrsv with the upper triangle of an n x n block (leading dim lda):
 * trans 0 solves R x = b, trans 1 solves R^T x = b. */
void lxo_trsv_upper(int n, const double* a, int lda, double* b, int trans) {
  if (trans == 0) {
    for (int i = n - 1; i >= 0; --i) {
      for (int j = i + 1; j < n; ++j) b[i] -= A_(i, j) * b[j];
      b[i] /= A_(i, i);
    }
  } else {
    for (int i = 0; i < n; ++i) {
      for (int j = 0; j < i; ++j) b[i] -= A_(j, i) * b[j];
      b[i] /= A_(i, i);
    }
  }
}

/* The uneven-batch Cholesky solve in float64 on float32 data: each system
 * i is lx.linear_solve(MatrixLinearOperator(A_i), b_i, Cholesky()) for the
 * exact float64 value of its data.  info[i] = potrf status (x NaN on
 * failure).  Systems are independent; `threads` > 1 splits them over
 * OpenMP threads when the library is built with OpenMP. */
void lxo_chol_uneven(int64_t count, const int32_t* ns, const int64_t* a_off, const float* a32,
                     const int64_t* b_off, const float* b32, double* x, int32_t* info, int threads) {
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 4) num_threads(threads > 0 ? threads : 1)
#endif
  for (int64_t s = 0; s < count; ++s) {
    int n = ns[s];
    double* a = (double*)malloc(sizeof(double) * (size_t)n * (size_t)n);
    const float* src = a32 + a_off[s];
    for (int64_t t = 0; t < (int64_t)n * n; ++t) a[t] = (double)src[t];
    double* xs = x + b_off[s];
    for (int i = 0; i < n; ++i) xs[i] = (double)b32[b_off[s] + i];
    int st = lxo_potrf_upper(n, a, n);
    info[s] = st;
    if (st != 0) {
      for (int i = 0; i < n; ++i) xs[i] = NAN;
    } else {
      lxo_potrs_upper(n, a, n, xs);
    }
    free(a);
  }
  (void)threads;
}
