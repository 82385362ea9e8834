// Uneven batch of Cholesky solves (orders 1..256) on the matrix cores.
//
// One workgroup of WAVES wavefronts owns one system at a time.  The matrix
// (order n <= 16 NB, padded with an identity block to a multiple of 16) is
// held as 16 x 16 blocks in v_mfma_f32_16x16x4f32 accumulator layout: lane
// l = 16 g + q holds rows 4g..4g+3 of column q of its block.  Only the upper
// triangle of blocks is kept.  Block rows are owned by waves in pairs
// (p, NB-1-p), which hold NB + 1 blocks together, so every wave holds the
// same number of blocks: a compile-time register array.
//
// Right-looking factorisation A = U^T U in 16-row panels K:
//   1. the owner of block row K writes it to LDS (Praw, row-major);
//   2. one wave factors the diagonal block with lane q holding column q
//      (row broadcasts by DPP), and inverts it: V = U_KK^-1, kept for the
//      triangular solves;
//   3. all waves solve the panel, U_KJ = V^T A_KJ, on the matrix cores and
//      store it as operand fragments (PF);
//   4. the owner keeps U_KJ (and V in place of U_KK) in its registers; every
//      wave subtracts U_KI^T U_KJ from its trailing blocks (four MFMAs per
//      block).
// Three barriers per panel.  After the factorisation the registers hold U
// off the diagonal and the inverted diagonal blocks, so the triangular solves
// are block steps (a 16 x 16 product per block, no scalar chain; the
// inverted diagonal blocks stay in LDS): forward
// U^T y = c pushes each solved block down its block row; backward U x = y
// gathers each block row.  With refinement, residuals r = b - A x are formed
// in FP64 from A's upper triangle in global memory, block by block, and x
// (FP64, LDS) is corrected with FP32 solves until the normwise backward
// error meets the tolerance.
//
// Work arrives from a device queue (persistent launch: workgroups pull the
// next system of their class; the order is sorted largest first, and each
// class's launch takes its contiguous range of it, found by binary search)
// or one system per workgroup.
#include <type_traits>

#include "lxa_common.h"

namespace lxa {

int chol_mfma_launch(hipStream_t s, int64_t count, int max_n, const int32_t* n, const int64_t* a_offset,
                     const float* a, const int64_t* b_offset, const float* b, void* x, int x_dtype,
                     const int32_t* order, bool persistent, bool sorted, int max_refine, double tol, int32_t* info,
                     int32_t* iters, double* eta, unsigned int* queue, int max4, int max8);

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

__device__ __forceinline__ f4 mfma4(float a, float b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Lane K of this lane's row of 16 (DPP row_newbcast).
template <int K>
__device__ __forceinline__ float row_bcast(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x150 + K, 0xf, 0xf, false));
}

template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false));
}

template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}

// Sum over the 16 lanes of a row, returned in every lane of the row
// (quad swaps, then half-row and row mirrors).
__device__ __forceinline__ float row_sum16(float v) {
  v += dpp_f<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_f<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_f<0x141>(v);  // row_half_mirror
  v += dpp_f<0x140>(v);  // row_mirror
  return v;
}

__device__ __forceinline__ double row_sum16(double v) {
  v += dpp_d<0xB1>(v);
  v += dpp_d<0x4E>(v);
  v += dpp_d<0x141>(v);
  v += dpp_d<0x140>(v);
  return v;
}

// Sum over the four rows of 16 (lanes q, q+16, q+32, q+48), in every lane.
// v_permlane16_swap / v_permlane32_swap of a register with itself leave
// (value of lane l & ~16, value of lane l | 16) (resp. bit 32) in lane l: the
// lane's own value and its partner's, so their sum is v + partner, exactly as
// an exchange through LDS would give, in two VALU instructions.
__device__ __forceinline__ float pair_sum16(float v) {
  const auto s = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
  return __uint_as_float(s[0]) + __uint_as_float(s[1]);
}
__device__ __forceinline__ float pair_sum32(float v) {
  const auto s = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false, false);
  return __uint_as_float(s[0]) + __uint_as_float(s[1]);
}
__device__ __forceinline__ double pair_sum16(double v) {
  const auto lo = __builtin_amdgcn_permlane16_swap(static_cast<unsigned>(__double2loint(v)),
                                                   static_cast<unsigned>(__double2loint(v)), false, false);
  const auto hi = __builtin_amdgcn_permlane16_swap(static_cast<unsigned>(__double2hiint(v)),
                                                   static_cast<unsigned>(__double2hiint(v)), false, false);
  return __hiloint2double(static_cast<int>(hi[0]), static_cast<int>(lo[0])) +
         __hiloint2double(static_cast<int>(hi[1]), static_cast<int>(lo[1]));
}
__device__ __forceinline__ double pair_sum32(double v) {
  const auto lo = __builtin_amdgcn_permlane32_swap(static_cast<unsigned>(__double2loint(v)),
                                                   static_cast<unsigned>(__double2loint(v)), false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap(static_cast<unsigned>(__double2hiint(v)),
                                                   static_cast<unsigned>(__double2hiint(v)), false, false);
  return __hiloint2double(static_cast<int>(hi[0]), static_cast<int>(lo[0])) +
         __hiloint2double(static_cast<int>(hi[1]), static_cast<int>(lo[1]));
}

__device__ __forceinline__ float group_sum4(float v) { return pair_sum32(pair_sum16(v)); }
__device__ __forceinline__ double group_sum4(double v) { return pair_sum32(pair_sum16(v)); }

__host__ __device__ constexpr int round16(int n) { return (n + 15) & ~15; }

// Wave index within the workgroup, as a scalar (uniform) value, so that
// block-ownership tests compile to scalar branches.
__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }

// Loads of A go through a buffer resource covering the n x n matrix: an
// offset past its end reads 0 with no fault, so entries outside the upper
// triangle (or the matrix) are read as 0 without a branch per element.
constexpr int kOutside = 0x7ffffff0;  // a byte offset beyond any matrix of order <= 256
constexpr int kRsrcWord3 = 0x00020000;  // raw 32-bit buffer, gfx9 resource format

__device__ __forceinline__ __amdgpu_buffer_rsrc_t matrix_rsrc(const float* A, int n) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(A), 0, n * n * 4, kRsrcWord3);
}

// A(i, j) for j >= i inside the matrix, else 0.
__device__ __forceinline__ float upper_at(__amdgpu_buffer_rsrc_t r, int n, int i, int j) {
  const bool in = i < n && j < n && j >= i;
  // the builtin returns the 32 bits of the element
  return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, in ? (i * n + j) * 4 : kOutside, 0, 0));
}

// LDS carve-up of one workgroup (byte offsets).  The residual's per-wave
// column sums overlay the panel buffers, which are idle after the
// factorisation.
template <int NB, int WAVES>
struct Lay {
  static constexpr int kThreads = 64 * WAVES;
  static constexpr int kPairs = NB / 2;
  static constexpr int kPPW = kPairs / WAVES;  // row pairs per wave
  static constexpr int kTPP = NB + 1;          // blocks per row pair
  static constexpr int kN = 16 * NB;           // largest order
  static constexpr int kLdp = kN + 4;          // Praw row stride (floats)
  static constexpr size_t kPraw = 0;                                    // [16][kLdp] f32
  static constexpr size_t kPF = kPraw + 16 * kLdp * 4;                  // [NB][64][4] f32
  static constexpr size_t kPanelEnd = kPF + NB * 256 * 4;
  static constexpr size_t kColPart = 0;                                 // [WAVES][kN] f64 (overlay)
  static constexpr size_t kColAbs = kColPart + WAVES * kN * 8;          // [WAVES][kN] f64 (overlay)
  static constexpr size_t kVfrag = kPanelEnd;                           // [64][4] f32
  static constexpr size_t kVall = kVfrag + 256 * 4;                     // [NB][16][16] f32: V_K = U_KK^-1
  static constexpr size_t kX = kVall + NB * 256 * 4;                    // [kN] f64 solution
  static constexpr size_t kR = kX + kN * 8;                             // [kN] f64 residual
  static constexpr size_t kRowAbs = kR + kN * 8;                        // [kN] f64 row sums of |A|
  static constexpr size_t kC = kRowAbs + kN * 8;                        // [kN] f32 solve vector
  static constexpr size_t kS = kC + kN * 4;                             // [WAVES][16] f32 scratch row
  static constexpr size_t kRed = kS + WAVES * 16 * 4;                   // [WAVES][4] f64 block maxima
  static constexpr size_t kInt = kRed + WAVES * 4 * 8;                  // ints: item, fail, range
  static constexpr size_t kBytes = kInt + 64;
  static_assert(kColAbs + WAVES * kN * 8 <= kPanelEnd, "column sums must fit the panel overlay");
  static_assert(kPPW >= 1 && kPairs % WAVES == 0, "row pairs must divide among the waves");
};

// Block row pair of the t-th pair slot of wave w (a snake, so that each
// wave's rows spread over the matrix).
template <int WAVES>
__device__ __forceinline__ int pair_of(int w, int t) {
  return t * WAVES + ((t & 1) ? (WAVES - 1 - w) : w);
}

// Block (I, J) of slot k of row pair p: row p holds blocks p..NB-1 in slots
// 0..NB-1-p, row NB-1-p holds NB-1-p..NB-1 in slots NB-p..NB.
template <int NB>
__device__ __forceinline__ void slot_block(int p, int k, int& I, int& J) {
  if (k < NB - p) {
    I = p;
    J = p + k;
  } else {
    I = NB - 1 - p;
    J = k - 1;
  }
}

template <int NB, int WAVES>
using Tiles = f4[Lay<NB, WAVES>::kPPW][Lay<NB, WAVES>::kTPP];

struct Shared {
  float* Praw;
  float* PF;
  float* Vfrag;
  float* Vall;
  double* X;
  double* R;
  double* RowAbs;
  double* ColPart;
  double* ColAbs;
  float* C;
  float* S;
  double* Red;
  int* Int;
};

template <int NB, int WAVES>
__device__ __forceinline__ Shared carve(unsigned char* base) {
  using L = Lay<NB, WAVES>;
  Shared s;
  s.Praw = reinterpret_cast<float*>(base + L::kPraw);
  s.PF = reinterpret_cast<float*>(base + L::kPF);
  s.ColPart = reinterpret_cast<double*>(base + L::kColPart);
  s.ColAbs = reinterpret_cast<double*>(base + L::kColAbs);
  s.Vfrag = reinterpret_cast<float*>(base + L::kVfrag);
  s.Vall = reinterpret_cast<float*>(base + L::kVall);
  s.X = reinterpret_cast<double*>(base + L::kX);
  s.R = reinterpret_cast<double*>(base + L::kR);
  s.RowAbs = reinterpret_cast<double*>(base + L::kRowAbs);
  s.C = reinterpret_cast<float*>(base + L::kC);
  s.S = reinterpret_cast<float*>(base + L::kS);
  s.Red = reinterpret_cast<double*>(base + L::kRed);
  s.Int = reinterpret_cast<int*>(base + L::kInt);
  return s;
}

template <int WAVES>
__device__ __forceinline__ void wg_barrier() {
  if constexpr (WAVES == 1) {
    wave_lds_sync();
  } else {
    lds_barrier();
  }
}

// Factors and inverts the diagonal block in Praw columns col0..col0+15 (one
// wave; lane q = l % 16 works on column q, the four rows of 16 lanes
// redundantly).  Writes V = U_KK^-1 as MFMA A-operand fragments of V^T
// (Vfrag[l][kc] = V[4kc + l/16][l%16]) and row-major (Vm, which the
// triangular solves read later).  Returns the
// 1-based index of the first non-positive pivot within the block, or 0.
template <int LDP>
__device__ __forceinline__ int diag_factor_invert(const Shared& s, int col0, float* Vm) {
  const int l = threadIdx.x & 63, q = l & 15;
  float c[16];
#pragma unroll
  for (int p = 0; p < 16; ++p) c[p] = s.Praw[p * LDP + col0 + q];
  float rs[16];  // reciprocal pivots, wave-uniform (scalar registers)
  int fail = 0;
  static_for<0, 16>([&](auto S_) {
    constexpr int st = decltype(S_)::value;
    const float d = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(c[st]), st));
    if (!(d > 0.0f) && fail == 0) fail = st + 1;
    rs[st] = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(__builtin_amdgcn_rsqf(d))));
    c[st] = (q >= st) ? c[st] * rs[st] : c[st];  // U[st][q]
    static_for<st + 1, 16>([&](auto P_) {
      constexpr int p = decltype(P_)::value;
      c[p] = fmaf(-row_bcast<p>(c[st]), c[st], c[p]);  // A[p][q] -= U[st][p] U[st][q]
    });
  });
  // V = U^-1 by columns, in place: U v = e_q by back substitution.  Step p
  // needs row p of U (lane k's c[p], k > p) and v[k] = c[k] for k > p, then
  // c[p] is free for v[p].  The row's broadcasts are tied to the previous
  // step's result, so the compiler cannot issue all of them up front (which
  // would hold them all live).
  static_for<0, 16>([&](auto PP_) {
    constexpr int p = 15 - decltype(PP_)::value;
    float cp = c[p];
    if constexpr (p < 15) asm volatile("" : "+v"(cp) : "v"(c[p + 1]));
    float t = (q == p) ? 1.0f : 0.0f;
    static_for<p + 1, 16>([&](auto K_) {
      constexpr int k = decltype(K_)::value;
      t = fmaf(-row_bcast<k>(cp), c[k], t);
    });
    c[p] = t * rs[p];
  });
  if (l < 16) {
#pragma unroll
    for (int g2 = 0; g2 < 4; ++g2)
      *reinterpret_cast<f4*>(s.Vfrag + (q + 16 * g2) * 4) = f4{c[g2], c[4 + g2], c[8 + g2], c[12 + g2]};
#pragma unroll
    for (int p = 0; p < 16; ++p) Vm[p * 16 + q] = c[p];
  }
  return fail;
}

// Loads the upper triangle of A (order n, row-major) into the block
// registers; padding is the identity.  Entries below the diagonal of a
// diagonal block are zero.
template <int NB, int WAVES>
__device__ __forceinline__ void load_tiles(Tiles<NB, WAVES>& T, const float* __restrict__ A, int n, int nk) {
  using L = Lay<NB, WAVES>;
  const __amdgpu_buffer_rsrc_t rs = matrix_rsrc(A, n);
  const int w = wave_id(), l = threadIdx.x & 63, g = l >> 4, q = l & 15;
  // Every load is issued first, straight into the tiles (blocks past the
  // matrix read the out-of-range offset, which returns 0), and the identity
  // padding is added afterwards: with the add next to each load the
  // compiler waited for every load before issuing the next.
#pragma unroll
  for (int t = 0; t < L::kPPW; ++t) {
    const int p = pair_of<WAVES>(w, t);
#pragma unroll
    for (int k = 0; k < L::kTPP; ++k) {
      int I, J;
      slot_block<NB>(p, k, I, J);
      const int j = 16 * J + q;
#pragma unroll
      for (int r = 0; r < 4; ++r) T[t][k][r] = upper_at(rs, n, 16 * I + 4 * g + r, j);
    }
  }
#pragma unroll
  for (int t = 0; t < L::kPPW; ++t) {
    const int p = pair_of<WAVES>(w, t);
#pragma unroll
    for (int k = 0; k < L::kTPP; ++k) {
      int I, J;
      slot_block<NB>(p, k, I, J);
      if (I == J && 16 * J + 16 > n) {  // uniform: the diagonal block holding the padding
        const int j = 16 * J + q;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = 16 * I + 4 * g + r;
          if (i == j && i >= n) T[t][k][r] += 1.0f;
        }
      }
    }
  }
  (void)nk;
}

// The right-looking factorisation (see the file comment).  Returns 0 or the
// 1-based index of the first non-positive pivot; uniform over the workgroup.
template <int NB, int WAVES>
__device__ __forceinline__ int factor_tiles(Tiles<NB, WAVES>& T, int nk, const Shared& s) {
  using L = Lay<NB, WAVES>;
  const int w = wave_id(), l = threadIdx.x & 63, g = l >> 4, q = l & 15;
  for (int K = 0; K < nk; ++K) {
    // Per-lane LDS offsets, redefined (opaquely) every panel: otherwise the
    // compiler hoists one address per block and register out of the loop,
    // which holds over a hundred registers and spills.
    int praw_l = 4 * g * L::kLdp + q;  // Praw row 4g, column q
    int frag_l = 4 * l;                // fragment slot of lane l
    int back_l = 4 * q + g;            // fragment slot (q + 16 r, kc = g) for r = 0
    asm volatile("" : "+v"(praw_l), "+v"(frag_l), "+v"(back_l));
    // 1. block row K to Praw
#pragma unroll
    for (int t = 0; t < L::kPPW; ++t) {
      const int p = pair_of<WAVES>(w, t);
#pragma unroll
      for (int k = 0; k < L::kTPP; ++k) {
        int I, J;
        slot_block<NB>(p, k, I, J);
        if (I == K && J < nk) {
#pragma unroll
          for (int r = 0; r < 4; ++r) s.Praw[praw_l + r * L::kLdp + 16 * J] = T[t][k][r];
        }
      }
    }
    wg_barrier<WAVES>();
    // 2. diagonal block: factor and invert
    if (w == 0) {
      const int f = diag_factor_invert<L::kLdp>(s, 16 * K, s.Vall + 256 * K);
      if (l == 0) s.Int[1] = f ? 16 * K + f : 0;
    }
    wg_barrier<WAVES>();
    const int fail = s.Int[1];
    if (fail) return fail;
    // 3. panel solve U_KJ = V^T A_KJ, stored as operand fragments:
    //    PF[J][l][kc] = U_KJ[4kc + l/16][l%16]
    {
      const f4 vf = *reinterpret_cast<const f4*>(s.Vfrag + l * 4);
      for (int J = K + 1 + w; J < nk; J += WAVES) {
        f4 x = {0.0f, 0.0f, 0.0f, 0.0f};
        const float* col = s.Praw + 16 * J + q;
        x = mfma4(vf[0], col[(0 + g) * L::kLdp], x);
        x = mfma4(vf[1], col[(4 + g) * L::kLdp], x);
        x = mfma4(vf[2], col[(8 + g) * L::kLdp], x);
        x = mfma4(vf[3], col[(12 + g) * L::kLdp], x);
        float* pf = s.PF + J * 256;
#pragma unroll
        for (int r = 0; r < 4; ++r) pf[back_l + 64 * r] = x[r];
      }
    }
    wg_barrier<WAVES>();
    // 4. the owner keeps row K; every wave updates its trailing blocks
#pragma unroll
    for (int t = 0; t < L::kPPW; ++t) {
      const int p = pair_of<WAVES>(w, t);
#pragma unroll
      for (int k = 0; k < L::kTPP; ++k) {
        int I, J;
        slot_block<NB>(p, k, I, J);
        if (J < nk) {
          if (I == K) {
            if (J > K) {  // (the diagonal block's inverse stays in LDS, s.Vall)
              const float* pf = s.PF + J * 256;
#pragma unroll
              for (int r = 0; r < 4; ++r) T[t][k][r] = pf[back_l + 64 * r];
            }
          } else if (I > K) {
            const f4 fi = -*reinterpret_cast<const f4*>(s.PF + I * 256 + frag_l);
            const f4 fj = *reinterpret_cast<const f4*>(s.PF + J * 256 + frag_l);
            f4 acc = T[t][k];
            acc = mfma4(fi[0], fj[0], acc);
            acc = mfma4(fi[1], fj[1], acc);
            acc = mfma4(fi[2], fj[2], acc);
            acc = mfma4(fi[3], fj[3], acc);
            T[t][k] = acc;
          }
        }
        // two blocks' operands at a time (their loads and MFMA chains
        // interleave): issued all at once, the fragment loads of every
        // block would be live together
        if (k & 1) __builtin_amdgcn_sched_barrier(0);
      }
    }
  }
  return 0;
}

// U^T U z = c in place in s.C (float32, padded with zeros to 16 nk), from the
// factored blocks: U off the diagonal in registers, V_K = U_KK^-1 in LDS.
template <int NB, int WAVES>
__device__ __forceinline__ void solve_tiles(const Tiles<NB, WAVES>& T, int nk, const Shared& s) {
  using L = Lay<NB, WAVES>;
  const int w = wave_id(), l = threadIdx.x & 63, g = l >> 4, q = l & 15;
  float* C = s.C;
  // forward, U^T y = c: at step I the owner of block row I forms
  // y_I = V_I^T c_I and pushes c_J -= U_IJ^T y_I down its row
  for (int I = 0; I < nk; ++I) {
    int col_l = q;  // opaque per step, as in factor_tiles
    asm volatile("" : "+v"(col_l));
    bool mine = false;
    int tsel = 0;  // the pair slot holding block row I (wave-uniform)
#pragma unroll
    for (int t = 0; t < L::kPPW; ++t) {
      const int p = pair_of<WAVES>(w, t);
      if (p == I || NB - 1 - p == I) {
        mine = true;
        tsel = t;
      }
    }
    if (mine) {
      const int ps = pair_of<WAVES>(w, tsel);
      // y_I[q] = sum_p V[p][q] c[p] (rows 4g..4g+3 here, then across groups)
      const f4 cv = *reinterpret_cast<const f4*>(C + 16 * I + 4 * g);
      const float* V = s.Vall + 256 * I + (4 * g) * 16 + col_l;
      float part = V[0] * cv[0] + V[16] * cv[1] + V[32] * cv[2] + V[48] * cv[3];
      part = group_sum4(part);  // y_I[q], the same in all four groups
      // push c_J -= U_IJ^T y_I over the pair's slots: every slot reads its
      // c_J (J < NB: inside C) before any product, so the reads are in
      // flight together; only row I's blocks right of the diagonal are
      // computed and written back (distinct J, so no slot reads another's
      // write, and none reads c_I)
      float cj[L::kTPP];
#pragma unroll
      for (int k = 0; k < L::kTPP; ++k) {
        int Ib, J;
        slot_block<NB>(ps, k, Ib, J);
        cj[k] = C[16 * J + col_l];
      }
      if (l < 16) C[16 * I + q] = part;
      // y_I[4g + r] from lane 4g + r (group 0) instead of through LDS
      f4 yv;
#pragma unroll
      for (int r = 0; r < 4; ++r) yv[r] = __shfl(part, 4 * g + r, 64);
#pragma unroll
      for (int k = 0; k < L::kTPP; ++k) asm volatile("" : "+v"(cj[k]));
#pragma unroll
      for (int k = 0; k < L::kTPP; ++k) {
        int Ib, J;
        slot_block<NB>(ps, k, Ib, J);
        if (Ib == I && J > I && J < nk) {
          f4 u = T[0][k];
#pragma unroll
          for (int t = 1; t < L::kPPW; ++t) u = tsel == t ? T[t][k] : u;
          float pt = u[0] * yv[0] + u[1] * yv[1] + u[2] * yv[2] + u[3] * yv[3];
          pt = group_sum4(pt);
          if (l < 16) C[16 * J + col_l] = cj[k] - pt;
        }
      }
    }
    wg_barrier<WAVES>();
  }
  // backward, U x = y: at step K the owner of block row K gathers
  // s = sum_J U_KJ x_J and forms x_K = V_K (y_K - s)
  for (int K = nk - 1; K >= 0; --K) {
    int col_l = q;
    asm volatile("" : "+v"(col_l));
    bool mine = false;
    int tsel = 0;
#pragma unroll
    for (int t = 0; t < L::kPPW; ++t) {
      const int p = pair_of<WAVES>(w, t);
      if (p == K || NB - 1 - p == K) {
        mine = true;
        tsel = t;
      }
    }
    if (mine) {
      // every slot of the pair holding row K reads its x_J (J < NB: inside
      // C), and y_K and V_K are read, before any product, so the LDS reads
      // are in flight together; only row K's blocks right of the diagonal
      // are multiplied
      const int ps = pair_of<WAVES>(w, tsel);
      const float yq = C[16 * K + q];
      const float* V = s.Vall + 256 * K + (4 * g) * 16 + col_l;
      float vr[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) vr[r] = V[16 * r];
      float xj[L::kTPP];
#pragma unroll
      for (int k = 0; k < L::kTPP; ++k) {
        int Ib, J;
        slot_block<NB>(ps, k, Ib, J);
        xj[k] = C[16 * J + col_l];
      }
#pragma unroll
      for (int k = 0; k < L::kTPP; ++k) asm volatile("" : "+v"(xj[k]));
      float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int k = 0; k < L::kTPP; ++k) {
        int Ib, J;
        slot_block<NB>(ps, k, Ib, J);
        if (Ib == K && J > K && J < nk) {
          f4 u = T[0][k];
#pragma unroll
          for (int t = 1; t < L::kPPW; ++t) u = tsel == t ? T[t][k] : u;
#pragma unroll
          for (int r = 0; r < 4; ++r) acc[r] = fmaf(u[r], xj[k], acc[r]);
        }
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[r] = row_sum16(acc[r]);  // s[4g + r], the same in the row's 16 lanes
      // s[q] from lane 16 (q / 4), entry q % 4, instead of through LDS
      float sv[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) sv[r] = __shfl(acc[r], 16 * (q >> 2), 64);
      const int qr = q & 3;
      const float sq = qr == 0 ? sv[0] : qr == 1 ? sv[1] : qr == 2 ? sv[2] : sv[3];
      const float tq = yq - sq;
      float xr[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) xr[r] = row_sum16(vr[r] * tq);
      wave_lds_sync();
      if (q < 4) C[16 * K + 4 * g + q] = q == 0 ? xr[0] : q == 1 ? xr[1] : q == 2 ? xr[2] : xr[3];
    }
    wg_barrier<WAVES>();
  }
}

// Four workgroup-wide maxima in one exchange.
template <int WAVES>
__device__ __forceinline__ void block_max4(double (&v)[4], const Shared& s) {
  const int w = wave_id();
#pragma unroll
  for (int i = 0; i < 4; ++i) v[i] = wave_max(v[i]);
  if constexpr (WAVES > 1) {
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
      for (int i = 0; i < 4; ++i) s.Red[4 * w + i] = v[i];
    }
    lds_barrier();
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      double m = s.Red[i];
#pragma unroll
      for (int ww = 1; ww < WAVES; ++ww) m = fmax(m, s.Red[4 * ww + i]);
      v[i] = m;
    }
    lds_barrier();
  }
}

// r = b - S x in FP64 (S: the symmetric matrix of A's upper triangle, x in
// s.X), and with want_abs the row sums of |S| (s.RowAbs).  Each wave reads
// its own block rows of the upper triangle once: block (I, J) adds A_IJ x_J
// to rows of I (a row reduction at the end of the block row) and, off the
// diagonal, A_IJ^T x_I to rows of J (per-wave column sums in LDS).
constexpr int kResGroup = 2;  // blocks per residual load group

template <int NB, int WAVES>
__device__ __forceinline__ void residual_tiles(const float* __restrict__ A, const float* __restrict__ B, int n, int nk,
                               const Shared& s, bool want_abs) {
  using L = Lay<NB, WAVES>;
  const int w = wave_id(), l = threadIdx.x & 63, g = l >> 4, q = l & 15;
  const __amdgpu_buffer_rsrc_t rs = matrix_rsrc(A, n);
  double* colp = s.ColPart + w * L::kN;
  double* cola = s.ColAbs + w * L::kN;
  for (int i = l; i < 16 * nk; i += 64) {
    colp[i] = 0.0;
    cola[i] = 0.0;
  }
  wave_lds_sync();
#pragma unroll
  for (int t = 0; t < L::kPPW; ++t) {
    const int p = pair_of<WAVES>(w, t);
#pragma unroll
    for (int half = 0; half < 2; ++half) {
      const int I = half == 0 ? p : NB - 1 - p;
      if (I < nk) {
        // FP64 products; the |A| row sums (for the norm in the backward
        // error's denominator) in float32, which is plenty for a norm
        double racc[4] = {0.0, 0.0, 0.0, 0.0};
        float rabs[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        // blocks J = I.. of the row in groups of kResGroup, the next group's
        // loads issued before the current group is reduced (kResGroup * 4
        // loads in flight per lane; blocks past the matrix read 0)
        float cur[kResGroup][4], nxt[kResGroup][4];
        auto load_group = [&](int J0, float (&dst)[kResGroup][4]) {
#pragma unroll
          for (int u = 0; u < kResGroup; ++u)
#pragma unroll
            for (int r = 0; r < 4; ++r) dst[u][r] = upper_at(rs, n, 16 * I + 4 * g + r, 16 * (J0 + u) + q);
        };
        load_group(I, cur);
        for (int J0 = I; J0 < nk; J0 += kResGroup) {
          load_group(J0 + kResGroup, nxt);
#pragma unroll
          for (int u = 0; u < kResGroup; ++u) {
            const int J = J0 + u;
            if (J < nk) {
              const int j = 16 * J + q;
              const double xj = s.X[j];
              double cs = 0.0;
              float ca = 0.0f;
#pragma unroll
              for (int r = 0; r < 4; ++r) {
                const int i = 16 * I + 4 * g + r;
                const double ad = static_cast<double>(cur[u][r]);
                racc[r] = fma(ad, xj, racc[r]);
                rabs[r] += fabsf(cur[u][r]);
                if (j > i) {  // strictly upper: the symmetric entry of row j
                  cs = fma(ad, s.X[i], cs);
                  ca += fabsf(cur[u][r]);
                }
              }
              cs = group_sum4(cs);
              if (want_abs) ca = group_sum4(ca);
              if (l < 16) {
                colp[j] += cs;
                if (want_abs) cola[j] += static_cast<double>(ca);
              }
            }
          }
#pragma unroll
          for (int u = 0; u < kResGroup; ++u)
#pragma unroll
            for (int r = 0; r < 4; ++r) cur[u][r] = nxt[u][r];
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          racc[r] = row_sum16(racc[r]);
          if (want_abs) rabs[r] = row_sum16(rabs[r]);
        }
        if (q < 4) {
          const double rv = q == 0 ? racc[0] : q == 1 ? racc[1] : q == 2 ? racc[2] : racc[3];
          s.R[16 * I + 4 * g + q] = rv;
          if (want_abs)
            s.RowAbs[16 * I + 4 * g + q] =
                static_cast<double>(q == 0 ? rabs[0] : q == 1 ? rabs[1] : q == 2 ? rabs[2] : rabs[3]);
        }
      }
    }
  }
  wg_barrier<WAVES>();
  for (int i = threadIdx.x; i < n; i += L::kThreads) {
    double sx = s.R[i], sa = want_abs ? s.RowAbs[i] : 0.0;
#pragma unroll
    for (int ww = 0; ww < WAVES; ++ww) {
      sx += s.ColPart[ww * L::kN + i];
      if (want_abs) sa += s.ColAbs[ww * L::kN + i];
    }
    s.R[i] = static_cast<double>(B[i]) - sx;
    if (want_abs) s.RowAbs[i] = sa;
  }
  wg_barrier<WAVES>();
}

// First index of order[lo, hi) whose order n[order[i]] <= v (the order is
// sorted by n descending), by binary search.
__device__ __forceinline__ int64_t first_at_most(const int32_t* __restrict__ ns, const int32_t* __restrict__ order, int64_t lo,
                                 int64_t hi, int v) {
  while (lo < hi) {
    const int64_t mid = lo + ((hi - lo) >> 1);
    if (ns[order[mid]] <= v)
      hi = mid;
    else
      lo = mid + 1;
  }
  return lo;
}

template <int NB, int WAVES>
constexpr int waves_per_eu() {
  return NB == 16 ? 2 : 4;
}

// KR classes: this launch solves the systems with n_lo < n <= n_hi.
template <int NB, int WAVES, bool kRefine, typename TX>
__global__ __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(waves_per_eu<NB, WAVES>())))
void chol_mfma_kernel(int64_t count, int n_lo, int n_hi, const int32_t* __restrict__ ns,
                      const int64_t* __restrict__ a_off, const float* __restrict__ a,
                      const int64_t* __restrict__ b_off, const float* __restrict__ b, TX* __restrict__ x,
                      const int32_t* __restrict__ order, int persistent,
                      const unsigned int* __restrict__ unsorted, unsigned int* __restrict__ queue, int max_refine,
                      double tol, int32_t* __restrict__ info, int32_t* __restrict__ iters,
                      double* __restrict__ eta) {
  using L = Lay<NB, WAVES>;
  extern __shared__ __align__(16) unsigned char lxa_smem[];
  const Shared s = carve<NB, WAVES>(lxa_smem);
  const int tid = threadIdx.x;
  int64_t lo = 0, hi = count;
  if (persistent && unsorted) {
    // the caller declared the order sorted (largest first) and the prepass
    // confirmed it: this class's systems are one contiguous range
    if (tid == 0) {
      int64_t a0 = 0, a1 = count;
      if (*unsorted == 0u) {
        a0 = first_at_most(ns, order, 0, count, n_hi);
        a1 = first_at_most(ns, order, a0, count, n_lo);
      }
      s.Int[2] = static_cast<int>(a0);
      s.Int[3] = static_cast<int>(a1);
    }
    wg_barrier<WAVES>();
    lo = s.Int[2];
    hi = s.Int[3];
  }
  int64_t next = persistent ? 0 : blockIdx.x;
  for (;;) {
    if (persistent) {
      if (tid == 0) s.Int[0] = static_cast<int>(lo + atomicAdd(queue, 1u));
      wg_barrier<WAVES>();
      next = s.Int[0];
      wg_barrier<WAVES>();
    }
    if (next >= hi) break;
    const int64_t sys = order[next];
    if (!persistent) next = hi;
    const int n = ns[sys];
    if (n <= n_lo || n > n_hi) continue;  // another class's system, or invalid (reported elsewhere); uniform
    const float* A = a + a_off[sys];
    const float* Bv = b + b_off[sys];
    TX* Xo = x + b_off[sys];
    const int nk = round16(n) >> 4;

    Tiles<NB, WAVES> T;
    load_tiles<NB, WAVES>(T, A, n, nk);
    const int fail = factor_tiles<NB, WAVES>(T, nk, s);
    if (fail) {
      for (int i = tid; i < n; i += L::kThreads) Xo[i] = static_cast<TX>(quiet_nan<double>());
      if (tid == 0) {
        info[sys] = fail;
        if (iters) iters[sys] = 0;
        if (eta) eta[sys] = -1.0;
      }
      wg_barrier<WAVES>();
      continue;
    }
    // Pass -1 solves with b; later passes form the FP64 residual, test the
    // backward error and solve for the correction.
    int steps = 0;
    double last_eta = -1.0, bnorm = 0.0, anorm = 0.0;
    const int passes = kRefine ? max_refine : 0;
    for (int it = -1; it < passes; ++it) {
      if constexpr (kRefine) {
        if (it >= 0) {
          residual_tiles<NB, WAVES>(A, Bv, n, nk, s, it == 0);
          double m4[4] = {0.0, 0.0, 0.0, 0.0};
          for (int i = tid; i < n; i += L::kThreads) {
            m4[0] = fmax(m4[0], fabs(s.R[i]));
            m4[1] = fmax(m4[1], fabs(s.X[i]));
            if (it == 0) {
              m4[2] = fmax(m4[2], s.RowAbs[i]);
              m4[3] = fmax(m4[3], fabs(static_cast<double>(Bv[i])));
            }
          }
          block_max4<WAVES>(m4, s);
          if (it == 0) {
            anorm = m4[2];
            bnorm = m4[3];
          }
          const double denom = anorm * m4[1] + bnorm;
          last_eta = denom > 0.0 ? m4[0] / denom : (m4[0] > 0.0 ? INFINITY : 0.0);
          ++steps;
          if (!(last_eta > tol)) break;  // also stops on NaN
        }
      }
      for (int i = tid; i < 16 * nk; i += L::kThreads)
        s.C[i] = i < n ? (it < 0 ? Bv[i] : static_cast<float>(s.R[i])) : 0.0f;
      wg_barrier<WAVES>();
      solve_tiles<NB, WAVES>(T, nk, s);
      // x beyond n stays zero: the residual multiplies it by A's zero padding
      for (int i = tid; i < 16 * nk; i += L::kThreads)
        s.X[i] = i < n ? (it < 0 ? 0.0 : s.X[i]) + static_cast<double>(s.C[i]) : 0.0;
      wg_barrier<WAVES>();
    }
    for (int i = tid; i < n; i += L::kThreads) Xo[i] = static_cast<TX>(s.X[i]);
    if (tid == 0) {
      info[sys] = 0;
      if (iters) iters[sys] = steps;
      if (eta) eta[sys] = last_eta;
    }
    wg_barrier<WAVES>();
  }
}

// Marks systems whose order lies outside [1, max_n]; with `unsorted` non-null
// also checks that the order is sorted by n descending and sets *unsorted
// when it is not (the class launches then sweep the whole order).
__global__ void uneven_prepass_kernel(int64_t count, int max_n, const int32_t* __restrict__ ns,
                                      const int32_t* __restrict__ order, int32_t* __restrict__ info,
                                      int32_t* __restrict__ iters, double* __restrict__ eta,
                                      unsigned int* __restrict__ unsorted) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const int64_t sys = order[i];
  const int n = ns[sys];
  if (n < 1 || n > max_n) {
    info[sys] = -1;
    if (iters) iters[sys] = 0;
    if (eta) eta[sys] = -1.0;
  }
  if (unsorted && i + 1 < count && ns[order[i + 1]] > n) atomicOr(unsorted, 1u);
}

__global__ void zero_u32x4_kernel(unsigned int* p) {
  if (threadIdx.x < 4) p[threadIdx.x] = 0u;
}

template <typename K>
int allow_lds(K kernel, size_t bytes) {
  if (bytes <= 64 * 1024) return LXA_OK;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes));
  return e == hipSuccess ? LXA_OK : hip_error(e, "hipFuncSetAttribute(MaxDynamicSharedMemorySize)");
}

template <int NB, int WAVES, bool kRefine, typename TX>
int launch_mfma_class(hipStream_t s, int64_t count, int n_lo, int n_hi, const int32_t* n, const int64_t* a_offset,
                      const float* a, const int64_t* b_offset, const float* b, void* x, const int32_t* order,
                      bool persistent, const unsigned int* unsorted, unsigned int* queue, int max_refine, double tol,
                      int32_t* info, int32_t* iters, double* eta) {
  using L = Lay<NB, WAVES>;
  auto kernel = chol_mfma_kernel<NB, WAVES, kRefine, TX>;
  const size_t shmem = L::kBytes;
  if (int rc = allow_lds(kernel, shmem)) return rc;
  int64_t grid = count;
  if (persistent) {
    int per_cu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(kernel),
                                                                L::kThreads, shmem);
    if (e != hipSuccess) return hip_error(e, "hipOccupancyMaxActiveBlocksPerMultiprocessor");
    if (per_cu < 1) per_cu = 1;
    grid = static_cast<int64_t>(num_cus()) * per_cu;
    if (grid > count) grid = count;
  }
  hipLaunchKernelGGL(kernel, dim3(static_cast<unsigned>(grid)), dim3(L::kThreads), shmem, s, count, n_lo, n_hi, n,
                     a_offset, a, b_offset, b, static_cast<TX*>(x), order, persistent ? 1 : 0, unsorted, queue,
                     max_refine, tol, info, iters, eta);
  LXA_LAUNCH_CHECK("chol_mfma_kernel launch");
  return LXA_OK;
}

template <int NB, int WAVES>
int launch_mfma_any(hipStream_t s, int64_t count, int n_lo, int n_hi, const int32_t* n, const int64_t* a_offset,
                    const float* a, const int64_t* b_offset, const float* b, void* x, int x_dtype,
                    const int32_t* order, bool persistent, const unsigned int* sorted, unsigned int* queue,
                    int max_refine, double tol, int32_t* info, int32_t* iters, double* eta) {
  if (max_refine > 0)
    return launch_mfma_class<NB, WAVES, true, double>(s, count, n_lo, n_hi, n, a_offset, a, b_offset, b, x, order,
                                                      persistent, sorted, queue, max_refine, tol, info, iters, eta);
  if (x_dtype == LXA_F64)
    return launch_mfma_class<NB, WAVES, false, double>(s, count, n_lo, n_hi, n, a_offset, a, b_offset, b, x, order,
                                                       persistent, sorted, queue, 0, tol, info, iters, eta);
  return launch_mfma_class<NB, WAVES, false, float>(s, count, n_lo, n_hi, n, a_offset, a, b_offset, b, x, order,
                                                    persistent, sorted, queue, 0, tol, info, iters, eta);
}

}  // namespace

// Tile classes: orders up to max4 (<= 64) on 4 x 4 blocks of one wave, up
// to max8 (<= 128) on 8 x 8 blocks of four waves, the rest on 16 x 16 blocks
// of four waves (the boundaries come from the autotune table, or from the
// tuner's own calls).  `queue` holds four counters: one per class, then the
// prepass's unsorted flag.
int chol_mfma_launch(hipStream_t s, int64_t count, int max_n, const int32_t* n, const int64_t* a_offset,
                     const float* a, const int64_t* b_offset, const float* b, void* x, int x_dtype,
                     const int32_t* order, bool persistent, bool sorted, int max_refine, double tol, int32_t* info,
                     int32_t* iters, double* eta, unsigned int* queue, int max4, int max8) {
  if (max4 < 0 || max4 > 64 || max8 < max4 || max8 > 128)
    return set_error(LXA_EINVAL, "uneven Cholesky: class boundaries %d, %d outside 0 <= max4 <= 64, max4 <= max8 <= 128",
                     max4, max8);
  if (persistent) {
    hipLaunchKernelGGL(zero_u32x4_kernel, dim3(1), dim3(64), 0, s, queue);
    LXA_LAUNCH_CHECK("zero_u32x4_kernel");
  }
  unsigned int* unsorted = (persistent && sorted) ? queue + 3 : nullptr;
  const unsigned blocks = static_cast<unsigned>((count + 255) / 256);
  hipLaunchKernelGGL(uneven_prepass_kernel, dim3(blocks), dim3(256), 0, s, count, max_n, n, order, info, iters, eta,
                     unsorted);
  LXA_LAUNCH_CHECK("uneven_prepass_kernel");
  if (!persistent) {
    // one workgroup per listed system, on the class that covers max_n
    if (max_n > max8)
      return launch_mfma_any<16, 4>(s, count, 0, max_n, n, a_offset, a, b_offset, b, x, x_dtype, order, false,
                                    nullptr, queue, max_refine, tol, info, iters, eta);
    if (max_n > max4)
      return launch_mfma_any<8, 4>(s, count, 0, max_n, n, a_offset, a, b_offset, b, x, x_dtype, order, false, nullptr,
                                   queue, max_refine, tol, info, iters, eta);
    return launch_mfma_any<4, 1>(s, count, 0, max_n, n, a_offset, a, b_offset, b, x, x_dtype, order, false, nullptr,
                                 queue, max_refine, tol, info, iters, eta);
  }
  int rc = LXA_OK;
  if (max_n > max8) {
    rc = launch_mfma_any<16, 4>(s, count, max8, max_n, n, a_offset, a, b_offset, b, x, x_dtype, order, true,
                                unsorted, queue + 0, max_refine, tol, info, iters, eta);
    if (rc) return rc;
  }
  if (max_n > max4 && max8 > max4) {
    rc = launch_mfma_any<8, 4>(s, count, max4, max_n < max8 ? max_n : max8, n, a_offset, a, b_offset, b, x, x_dtype,
                               order, true, unsorted, queue + 1, max_refine, tol, info, iters, eta);
    if (rc) return rc;
  }
  if (max4 > 0)
    rc = launch_mfma_any<4, 1>(s, count, 0, max_n < max4 ? max_n : max4, n, a_offset, a, b_offset, b, x, x_dtype,
                               order, true, unsorted, queue + 2, max_refine, tol, info, iters, eta);
  return rc;
}

}  // namespace lxa
