// Singular value decomposition of a batch of dense matrices by one-sided
// (Hestenes) Jacobi rotations: the factorisation behind lineax's SVD solver
// (lineax/_solver/svd.py:49-53, jax.scipy.linalg.svd with
// full_matrices=False).
//
// One workgroup per matrix.  G is A (m >= n) or A^T (m < n), R x C with
// C = min(m, n), stored column-major in the caller's workspace next to V
// (C x C, starts as I).  A sweep visits every column pair once, in the
// round-robin order that splits each round into C/2 disjoint pairs; a team
// of threads handles a pair: the three inner products (FP64), and when the
// pair is not orthogonal to working precision the rotation that makes it so,
// applied to both G and V.  Sweeps stop when one applies no rotation.  Then
// s_c = |g_c|, U = G diag(1/s), the columns ranked by descending s.  The
// workspace keeps G and V L2-resident for the sizes this targets (C <= 512).
#include <limits>

#include "lxa_common.h"

namespace lxa {
namespace {

constexpr int kSvdThreads = 256;
constexpr int kSvdMaxCols = 512;

template <typename T>
__global__ __launch_bounds__(kSvdThreads) void svd_jacobi_kernel(int m, int n, const T* __restrict__ a,
                                                                  int64_t a_stride, T* __restrict__ u,
                                                                  T* __restrict__ sv, T* __restrict__ vt,
                                                                  T* __restrict__ work, int max_sweeps,
                                                                  int32_t* __restrict__ sweeps_out) {
  __shared__ double s_norm[kSvdMaxCols];
  __shared__ int s_rank[kSvdMaxCols];
  __shared__ int s_rot;
  const int64_t mat = blockIdx.x;
  const bool wide = m < n;
  const int R = wide ? n : m, C = wide ? m : n;
  const int tid = threadIdx.x;
  const T* A = a + mat * a_stride;
  T* G = work + mat * ((int64_t)R * C + (int64_t)C * C);
  T* V = G + (int64_t)R * C;
  for (int64_t idx = tid; idx < (int64_t)R * C; idx += kSvdThreads) {
    const int c = (int)(idx / R), r = (int)(idx % R);
    G[idx] = wide ? A[(int64_t)c * n + r] : A[(int64_t)r * n + c];
  }
  for (int64_t idx = tid; idx < (int64_t)C * C; idx += kSvdThreads) V[idx] = (idx / C == idx % C) ? T(1) : T(0);
  __syncthreads();

  // pairs per round and the team (a power of two, within a wave) per pair
  const int N = C + (C & 1);
  const int P = N / 2;
  int ts = 1;
  while (ts * 2 * P <= kSvdThreads && ts * 2 <= kWave) ts *= 2;
  const int team = tid / ts, tl = tid % ts, teams = kSvdThreads / ts;
  const double eps = static_cast<double>(std::numeric_limits<T>::epsilon());
  int sweep = 0;
  while (sweep < max_sweeps) {
    if (tid == 0) s_rot = 0;
    __syncthreads();
    int rotated = 0;
    for (int rnd = 0; rnd < N - 1; ++rnd) {
      for (int pr = team; pr < P; pr += teams) {
        // player 0 stays, the others rotate by one position per round
        const int k1 = pr, k2 = N - 1 - pr;
        int p = k1 == 0 ? 0 : ((k1 - 1 + rnd) % (N - 1)) + 1;
        int q = ((k2 - 1 + rnd) % (N - 1)) + 1;
        if (p > q) {
          const int t = p;
          p = q;
          q = t;
        }
        if (q >= C) continue;  // the padding column of an odd C (uniform across the team)
        T* gp = G + (int64_t)p * R;
        T* gq = G + (int64_t)q * R;
        double al = 0.0, be = 0.0, ga = 0.0;
        for (int r = tl; r < R; r += ts) {
          const double x = static_cast<double>(gp[r]), y = static_cast<double>(gq[r]);
          al = fma(x, x, al);
          be = fma(y, y, be);
          ga = fma(x, y, ga);
        }
        for (int off = ts / 2; off > 0; off >>= 1) {
          al += __shfl_xor(al, off, kWave);
          be += __shfl_xor(be, off, kWave);
          ga += __shfl_xor(ga, off, kWave);
        }
        if (ga != 0.0 && fabs(ga) > eps * sqrt(al) * sqrt(be)) {
          // the rotation [c s; -s c] that zeroes the pair's inner product
          const double zeta = (be - al) / (2.0 * ga);
          const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
          const double cs = 1.0 / sqrt(1.0 + t * t), sn = cs * t;
          for (int r = tl; r < R; r += ts) {
            const double x = static_cast<double>(gp[r]), y = static_cast<double>(gq[r]);
            gp[r] = static_cast<T>(cs * x - sn * y);
            gq[r] = static_cast<T>(sn * x + cs * y);
          }
          T* vp = V + (int64_t)p * C;
          T* vq = V + (int64_t)q * C;
          for (int r = tl; r < C; r += ts) {
            const double x = static_cast<double>(vp[r]), y = static_cast<double>(vq[r]);
            vp[r] = static_cast<T>(cs * x - sn * y);
            vq[r] = static_cast<T>(sn * x + cs * y);
          }
          rotated = 1;
        }
      }
      __syncthreads();  // the next round pairs these columns with others
    }
    if (rotated) s_rot = 1;
    __syncthreads();
    ++sweep;
    if (!s_rot) break;
  }

  // singular values, their descending order, and the factors
  for (int c = team; c < C; c += teams) {
    const T* gc = G + (int64_t)c * R;
    double nn = 0.0;
    for (int r = tl; r < R; r += ts) nn = fma(static_cast<double>(gc[r]), static_cast<double>(gc[r]), nn);
    for (int off = ts / 2; off > 0; off >>= 1) nn += __shfl_xor(nn, off, kWave);
    if (tl == 0) s_norm[c] = sqrt(nn);
  }
  __syncthreads();
  // Descending order with NaN above every number and ties broken by column
  // index: a total order, so the ranks are a permutation and every output
  // slot is written.
  for (int c = tid; c < C; c += kSvdThreads) {
    const double v = s_norm[c];
    const bool vnan = isnan(v);
    int rk = 0;
    for (int j = 0; j < C; ++j) {
      const double w = s_norm[j];
      const bool wnan = isnan(w);
      const bool above = wnan ? (!vnan || j < c) : (!vnan && (w > v || (w == v && j < c)));
      rk += above ? 1 : 0;
    }
    s_rank[c] = rk;
    sv[mat * C + rk] = static_cast<T>(v);
  }
  __syncthreads();
  // tall: U[r][k] = g_c[r] / s_c (m x C), Vt[k][j] = v_c[j] (C x n);
  // wide: A = V diag(s) (G diag(1/s))^T, so U[j][k] = v_c[j] (m x C) and
  // Vt[k][r] = g_c[r] / s_c (C x n).
  T* Um = u + mat * (int64_t)m * C;
  T* Vm = vt + mat * (int64_t)C * n;
  for (int64_t idx = tid; idx < (int64_t)R * C; idx += kSvdThreads) {
    const int c = (int)(idx / R), r = (int)(idx % R);
    const double sc = s_norm[c];
    const T val = sc > 0.0 ? static_cast<T>(static_cast<double>(G[idx]) / sc) : T(0);
    const int k = s_rank[c];
    if (!wide)
      Um[(int64_t)r * C + k] = val;
    else
      Vm[(int64_t)k * n + r] = val;
  }
  for (int64_t idx = tid; idx < (int64_t)C * C; idx += kSvdThreads) {
    const int c = (int)(idx / C), j = (int)(idx % C);
    const int k = s_rank[c];
    if (!wide)
      Vm[(int64_t)k * n + j] = V[idx];
    else
      Um[(int64_t)j * C + k] = V[idx];
  }
  // sweeps run, or max_sweeps + 1 when the last sweep still rotated (not converged)
  if (tid == 0 && sweeps_out) sweeps_out[mat] = s_rot ? max_sweeps + 1 : sweep;
}

}  // namespace
}  // namespace lxa

extern "C" int lxa_svd_workspace(int dtype, int64_t batch, int64_t m, int64_t n, size_t* bytes) {
  LXA_REQUIRE(bytes && batch >= 0 && m >= 0 && n >= 0, "lxa_svd_workspace: bad argument");
  LXA_REQUIRE(dtype == LXA_F32 || dtype == LXA_F64, "lxa_svd_workspace: dtype must be LXA_F32 or LXA_F64");
  const int64_t R = m > n ? m : n, C = m < n ? m : n;
  *bytes = static_cast<size_t>(batch) * static_cast<size_t>(R * C + C * C) * (dtype == LXA_F64 ? 8 : 4);
  return LXA_OK;
}

extern "C" int lxa_svd(void* stream, int dtype, int64_t batch, int64_t m, int64_t n, const void* a, int64_t a_stride,
                       void* u, void* s, void* vt, void* work, int max_sweeps, int32_t* sweeps) {
  LXA_REQUIRE(dtype == LXA_F32 || dtype == LXA_F64, "lxa_svd: dtype must be LXA_F32 or LXA_F64");
  LXA_REQUIRE(batch >= 0 && batch <= 0x7fffffff && m >= 0 && n >= 0 && m <= (1 << 20) && n <= (1 << 20),
              "lxa_svd: bad batch / shape");
  LXA_REQUIRE(max_sweeps >= 1, "lxa_svd: max_sweeps must be positive");
  const int64_t C = m < n ? m : n;
  if (C > lxa::kSvdMaxCols)
    return lxa::set_error(LXA_EUNSUPPORTED, "lxa_svd: min(m, n) must be <= %d (got %lld)", lxa::kSvdMaxCols,
                          (long long)C);
  if (batch == 0 || C == 0) return LXA_OK;
  LXA_REQUIRE(a && u && s && vt && work && a_stride >= m * n, "lxa_svd: null pointer or bad stride");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == LXA_F64)
    hipLaunchKernelGGL(lxa::svd_jacobi_kernel<double>, dim3(static_cast<unsigned>(batch)), dim3(lxa::kSvdThreads), 0,
                       st, (int)m, (int)n, static_cast<const double*>(a), a_stride, static_cast<double*>(u),
                       static_cast<double*>(s), static_cast<double*>(vt), static_cast<double*>(work), max_sweeps,
                       sweeps);
  else
    hipLaunchKernelGGL(lxa::svd_jacobi_kernel<float>, dim3(static_cast<unsigned>(batch)), dim3(lxa::kSvdThreads), 0,
                       st, (int)m, (int)n, static_cast<const float*>(a), a_stride, static_cast<float*>(u),
                       static_cast<float*>(s), static_cast<float*>(vt), static_cast<float*>(work), max_sweeps,
                       sweeps);
  LXA_LAUNCH_CHECK("lxa_svd");
  return LXA_OK;
}
