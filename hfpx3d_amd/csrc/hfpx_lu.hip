// hfpx_lu.hip -- dense LU with partial pivoting and triangular solves on
// gfx950, FP64 and FP32.
//
// Replaces il::LU<il::Array2D<double>> (il/linear_algebra/dense/factorization/
// LU.h:79-138), i.e. LAPACKE_dgetrf / LAPACKE_dgetrs: same result semantics
// (P A = L U in place, unit lower L, 1-based sequential row interchanges in
// ipiv, info = first exactly-zero pivot column), FP32 variants for the
// mixed-precision solve.
//
// Algorithm: right-looking blocked LU, two levels.  An outer panel of NB
// columns is factorised as a sequence of inner panels of IB columns; each
// inner panel is factorised by one cooperative kernel whose workgroups keep
// their rows of the panel in registers and meet once per column at a grid
// barrier (pivot search -> interchange -> rank-1 update).  The rest of the
// outer panel, and then the trailing matrix, get the row interchanges, a
// unit-lower triangular solve for the U rows, and the Schur-complement update
// C -= L21 U12 on MFMA (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32).
#include <hip/hip_cooperative_groups.h>

#include <algorithm>

#include "hfpx_internal.h"

namespace cg = cooperative_groups;

namespace hfpx {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

template <typename T>
struct MfmaVec;
template <>
struct MfmaVec<float> {
  typedef floatx4 type;
  static __device__ __forceinline__ floatx4 mfma(float a, float b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // result element v of lane l is D[4 (l / 16) + v][l % 16]
  static __device__ __forceinline__ int drow(int lane, int v) { return 4 * (lane >> 4) + v; }
};
template <>
struct MfmaVec<double> {
  typedef doublex4 type;
  static __device__ __forceinline__ doublex4 mfma(double a, double b, doublex4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // result element v of lane l is D[4 v + l / 16][l % 16] (measured on gfx950)
  static __device__ __forceinline__ int drow(int lane, int v) { return 4 * v + (lane >> 4); }
};

// ---------------------------------------------------------------------
// C (M x N) -= A (M x K) * B (K x N), all column-major.
//
// Workgroup tile 128 x 128 (4 waves of 64 x 64), K in chunks of KC staged
// through LDS.  The MFMA computes the transposed product (U^T L^T) so that the
// 16 lanes of a group map to 16 consecutive rows of one column of C: the
// epilogue then writes contiguous column segments.
// MFMA 16x16x4 operand layout: lane l supplies A[l%16][l/16] and B[l/16][l%16];
// the row of result element v differs between the f32 and f64 forms (drow).
// ---------------------------------------------------------------------
constexpr int GT = 128;  // tile rows (M)
constexpr int KC = 16;   // K chunk

// GTN = tile columns (N): 128, or 64 for skinny products (many right-hand
// sides), where a 128-wide tile would leave half of every MFMA idle.
template <typename T, int GTN>
__global__ __launch_bounds__(256) void k_gemm_sub(T* __restrict__ C, int64_t ldc,
                                                  const T* __restrict__ A, int64_t lda,
                                                  const T* __restrict__ B, int64_t ldb, int64_t M,
                                                  int64_t N, int64_t K) {
  typedef typename MfmaVec<T>::type vec4;
  constexpr int WN = GTN / 2;      // wave tile columns
  constexpr int TC = WN / 16;      // 16x16 MFMA tiles across a wave's columns
  constexpr int BQ = KC * GTN / 256;  // B elements loaded per thread per chunk
  __shared__ T sA[2][KC][GT];      // sA[k][i] = A(row0 + i, k0 + k)
  __shared__ T sB[2][KC][GTN];     // sB[k][j] = B(k0 + k, col0 + j)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t row0 = (int64_t)blockIdx.x * GT, col0 = (int64_t)blockIdx.y * GTN;
  const int wr = (wave & 1) * 64, wc = (wave >> 1) * WN;  // wave's 64 x WN sub-tile

  vec4 acc[TC][4];
#pragma unroll
  for (int a = 0; a < TC; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = vec4{0, 0, 0, 0};

  T ra[8], rb[BQ];
  auto load_global = [&](int64_t k0) {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int e = tid + 256 * q;  // 0 .. KC * GT - 1
      const int ai = e & (GT - 1), ak = e >> 7;
      const int64_t gr = row0 + ai, gk = k0 + ak;
      ra[q] = (gr < M && gk < K) ? A[gr + gk * lda] : T(0);
    }
#pragma unroll
    for (int q = 0; q < BQ; ++q) {
      const int e = tid + 256 * q;  // 0 .. KC * GTN - 1
      const int bk = e & (KC - 1), bj = e >> 4;
      const int64_t gc = col0 + bj, gk2 = k0 + bk;
      rb[q] = (gc < N && gk2 < K) ? B[gk2 + gc * ldb] : T(0);
    }
  };
  auto store_lds = [&](int buf) {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int e = tid + 256 * q;
      sA[buf][e >> 7][e & (GT - 1)] = ra[q];
    }
#pragma unroll
    for (int q = 0; q < BQ; ++q) {
      const int e = tid + 256 * q;
      sB[buf][e & (KC - 1)][e >> 4] = rb[q];
    }
  };

  const int nchunks = (int)((K + KC - 1) / KC);
  load_global(0);
  store_lds(0);
  __syncthreads();
  for (int c = 0; c < nchunks; ++c) {
    const int buf = c & 1;
    if (c + 1 < nchunks) load_global((int64_t)(c + 1) * KC);
#pragma unroll
    for (int k4 = 0; k4 < KC; k4 += 4) {
      const int kk = k4 + (lane >> 4);
      T fu[TC], fl[4];
#pragma unroll
      for (int t = 0; t < TC; ++t) fu[t] = sB[buf][kk][wc + 16 * t + (lane & 15)];  // U^T rows = C columns
#pragma unroll
      for (int t = 0; t < 4; ++t) fl[t] = sA[buf][kk][wr + 16 * t + (lane & 15)];   // L^T cols = C rows
#pragma unroll
      for (int tc = 0; tc < TC; ++tc)
#pragma unroll
        for (int tr = 0; tr < 4; ++tr) acc[tc][tr] = MfmaVec<T>::mfma(fu[tc], fl[tr], acc[tc][tr]);
    }
    if (c + 1 < nchunks) store_lds(buf ^ 1);
    __syncthreads();
  }
  // acc[tc][tr][v] = (A B)(row = wr + 16 tr + lane%16, col = wc + 16 tc + drow(lane, v))
#pragma unroll
  for (int tc = 0; tc < TC; ++tc)
#pragma unroll
    for (int tr = 0; tr < 4; ++tr)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t r = row0 + wr + 16 * tr + (lane & 15);
        const int64_t cc = col0 + wc + 16 * tc + MfmaVec<T>::drow(lane, v);
        if (r < M && cc < N) C[r + cc * ldc] -= acc[tc][tr][v];
      }
}

// ---------------------------------------------------------------------
// Cooperative panel factorisation of A[j0:n, j0:j0+jb] (jb <= IB).
// Workgroup w owns rows [j0 + w*chunk, j0 + (w+1)*chunk); thread t holds
// rows base + t + WG*r, r < RPT, of all jb columns in registers.
// ---------------------------------------------------------------------
constexpr int PWG = 512;  // threads per panel workgroup

template <typename T>
struct PanelScratch {
  T val[2][1024];          // per-workgroup candidate |pivot| (double-buffered)
  int64_t idx[2][1024];    // per-workgroup candidate row (relative)
};

template <typename T>
struct PanelCtx {
  int nwg, w, t;
  int jb;
  int64_t j0;
  PanelScratch<T>* ps;
  T* rows;
  int32_t* ipiv;
  int32_t* info;
  T* red_v;
  int64_t* red_i;
  int* red_w;
  T* s_piv;
};

// One column step of the panel (column J known at compile time, so the
// register-resident rows are indexed statically).
template <typename T, int IB, int RPT, int J>
__device__ __forceinline__ void panel_step(T (&a)[RPT][IB], const int64_t (&ri)[RPT],
                                           const PanelCtx<T>& P, cg::grid_group& grid) {
  if constexpr (J < IB) {
    if (J < P.jb) {
      const int par = J & 1;
      const int t = P.t;
      // 1. local pivot candidate among rows >= J
      T best = T(-1);
      int64_t bidx = -1;
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
        if (ri[r] >= J) {
          const T v = fabs(a[r][J]);
          if (v > best || (v == best && ri[r] < bidx)) {
            best = v;
            bidx = ri[r];
          }
        }
      }
      for (int off = 32; off > 0; off >>= 1) {
        const T ov = __shfl_xor(best, off);
        const int64_t oi = __shfl_xor(bidx, off);
        if (ov > best || (ov == best && oi >= 0 && (bidx < 0 || oi < bidx))) {
          best = ov;
          bidx = oi;
        }
      }
      if ((t & 63) == 0) {
        P.red_v[t >> 6] = best;
        P.red_i[t >> 6] = bidx;
      }
      __syncthreads();
      if (t == 0) {
        for (int q = 1; q < PWG / 64; ++q) {
          const T ov = P.red_v[q];
          const int64_t oi = P.red_i[q];
          if (ov > best || (ov == best && oi >= 0 && (bidx < 0 || oi < bidx))) {
            best = ov;
            bidx = oi;
          }
        }
        P.ps->val[par][P.w] = best;
        P.ps->idx[par][P.w] = bidx;
        P.red_i[0] = bidx;
      }
      // publish the candidate row and (owner of row J) row J itself
      T* cand = P.rows + ((size_t)par * (P.nwg + 1) + P.w) * IB;
      T* rowj = P.rows + ((size_t)par * (P.nwg + 1) + P.nwg) * IB;
      __syncthreads();
      const int64_t wg_best = P.red_i[0];
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
        if (ri[r] >= 0 && ri[r] == wg_best) {
#pragma unroll
          for (int c = 0; c < IB; ++c) cand[c] = a[r][c];
        }
        if (ri[r] == J) {
#pragma unroll
          for (int c = 0; c < IB; ++c) rowj[c] = a[r][c];
        }
      }
      __threadfence();
      grid.sync();
      // 2. global pivot: every thread reads candidates q = t, t + PWG, ...,
      //    then a workgroup arg-max (every thread read red_i[0] before the
      //    grid barrier, so the LDS slots can be reused)
      {
        T gb = T(-1);
        int64_t gi = -1;
        int gw = -1;
        for (int q = t; q < P.nwg; q += PWG) {
          const T ov = P.ps->val[par][q];
          const int64_t oi = P.ps->idx[par][q];
          if (oi >= 0 && (ov > gb || (ov == gb && oi < gi))) {
            gb = ov;
            gi = oi;
            gw = q;
          }
        }
        for (int off = 32; off > 0; off >>= 1) {
          const T ov = __shfl_xor(gb, off);
          const int64_t oi = __shfl_xor(gi, off);
          const int ow = __shfl_xor(gw, off);
          if (oi >= 0 && (ov > gb || (ov == gb && (gi < 0 || oi < gi)))) {
            gb = ov;
            gi = oi;
            gw = ow;
          }
        }
        if ((t & 63) == 0) {
          P.red_v[t >> 6] = gb;
          P.red_i[t >> 6] = gi;
          P.red_w[t >> 6] = gw;
        }
        __syncthreads();
        if (t == 0) {
          for (int q = 1; q < PWG / 64; ++q) {
            const T ov = P.red_v[q];
            const int64_t oi = P.red_i[q];
            if (oi >= 0 && (ov > gb || (ov == gb && (gi < 0 || oi < gi)))) {
              gb = ov;
              gi = oi;
              gw = P.red_w[q];
            }
          }
          P.red_i[0] = gi;
          P.red_i[1] = gw;
        }
        __syncthreads();
      }
      const int64_t p = P.red_i[0];
      const int pw = (int)P.red_i[1];
      const T* prow = P.rows + ((size_t)par * (P.nwg + 1) + pw) * IB;
      if (t < IB) P.s_piv[t] = prow[t];
      __syncthreads();
      const T pv = P.s_piv[J];
      const bool zero = (pv == T(0));
      if (P.w == 0 && t == 0) {
        P.ipiv[P.j0 + J] = (int32_t)(P.j0 + (zero ? J : p) + 1);
        if (zero && *P.info == 0) *P.info = (int32_t)(P.j0 + J + 1);
      }
      const T inv = zero ? T(0) : T(1) / pv;
      // 3. interchange + rank-1 update of the rows this thread holds
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
        if (ri[r] < J) continue;
        if (!zero && ri[r] == J) {
#pragma unroll
          for (int c = 0; c < IB; ++c) a[r][c] = P.s_piv[c];
          continue;
        }
        if (!zero && ri[r] == p) {
#pragma unroll
          for (int c = 0; c < IB; ++c) a[r][c] = rowj[c];
        }
        const T l = a[r][J] * inv;
        a[r][J] = l;
#pragma unroll
        for (int c = J + 1; c < IB; ++c) a[r][c] -= l * P.s_piv[c];
      }
    }
    panel_step<T, IB, RPT, J + 1>(a, ri, P, grid);
  }
}

template <typename T, int IB, int RPT>
__global__ __launch_bounds__(PWG) void k_panel(T* __restrict__ A, int64_t lda, int64_t n,
                                               int64_t j0, int jb, int32_t* __restrict__ ipiv,
                                               int32_t* __restrict__ info,
                                               PanelScratch<T>* __restrict__ ps,
                                               T* __restrict__ rows /* [2][nwg+1][IB] */) {
  cg::grid_group grid = cg::this_grid();
  const int nwg = gridDim.x, w = blockIdx.x, t = threadIdx.x;
  const int64_t m = n - j0;
  const int64_t chunk = (m + nwg - 1) / nwg;
  const int64_t base = (int64_t)w * chunk;
  __shared__ T red_v[PWG / 64];
  __shared__ int64_t red_i[PWG / 64];
  __shared__ int red_w[PWG / 64];
  __shared__ T s_piv[IB];

  T a[RPT][IB];
  int64_t ri[RPT];
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    ri[r] = base + t + PWG * r;
    const bool ok = (t + PWG * r) < chunk && ri[r] < m;
    if (!ok) ri[r] = -1;
#pragma unroll
    for (int c = 0; c < IB; ++c)
      a[r][c] = (ok && c < jb) ? A[(j0 + ri[r]) + (j0 + c) * lda] : T(0);
  }
  PanelCtx<T> P{nwg, w, t, jb, j0, ps, rows, ipiv, info, red_v, red_i, red_w, s_piv};
  panel_step<T, IB, RPT, 0>(a, ri, P, grid);
#pragma unroll
  for (int r = 0; r < RPT; ++r)
    if (ri[r] >= 0)
#pragma unroll
      for (int c = 0; c < IB; ++c)
        if (c < jb) A[(j0 + ri[r]) + (j0 + c) * lda] = a[r][c];
}

// ---------------------------------------------------------------------
// Row interchanges ipiv[k0:k1) (1-based, sequential) on columns [c0, c1).
// One thread per column.
// ---------------------------------------------------------------------
template <typename T>
__global__ void k_laswp(T* __restrict__ A, int64_t lda, const int32_t* __restrict__ ipiv,
                        int64_t k0, int64_t k1, int64_t c0, int64_t c1) {
  const int64_t c = c0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= c1) return;
  T* col = A + c * lda;
  for (int64_t j = k0; j < k1; ++j) {
    const int64_t p = ipiv[j] - 1;
    if (p != j) {
      const T tmp = col[j];
      col[j] = col[p];
      col[p] = tmp;
    }
  }
}

// ---------------------------------------------------------------------
// Leaf triangular solves for kb <= LB rows: one thread per column of B,
// the column held in registers, the kb x kb triangle in LDS (read as
// wave-wide broadcasts).  B := L^{-1} B (unit lower) or U^{-1} B (upper).
// ---------------------------------------------------------------------
constexpr int LB = 64;

template <typename T, bool LOWER>
__global__ __launch_bounds__(256) void k_trsm_leaf(const T* __restrict__ Lm, int64_t ldl, T* __restrict__ B,
                                                   int64_t ldb, int kb, int64_t ncols) {
  __shared__ T sL[LB][LB + 1];
  for (int e = threadIdx.x; e < LB * LB; e += 256) {
    const int i = e % LB, j = e / LB;
    sL[i][j] = (i < kb && j < kb) ? Lm[i + (int64_t)j * ldl] : T(0);
  }
  __syncthreads();
  const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= ncols) return;
  T* col = B + c * ldb;
  T x[LB];
#pragma unroll
  for (int i = 0; i < LB; ++i) x[i] = (i < kb) ? col[i] : T(0);
  if (LOWER) {
#pragma unroll
    for (int i = 1; i < LB; ++i) {
      // the guard keeps the scheduler from hoisting every LDS read of the
      // unrolled loop nest at once (which spills)
      if (i < kb) {
        T acc = x[i];
#pragma unroll
        for (int j = 0; j < i; ++j) acc -= sL[i][j] * x[j];
        x[i] = acc;
      }
    }
  } else {
#pragma unroll
    for (int i = LB - 1; i >= 0; --i) {
      if (i < kb) {
        T acc = x[i];
#pragma unroll
        for (int j = i + 1; j < LB; ++j) acc -= sL[i][j] * x[j];
        x[i] = acc / sL[i][i];
      }
    }
  }
#pragma unroll
  for (int i = 0; i < LB; ++i)
    if (i < kb) col[i] = x[i];
}

// ---------------------------------------------------------------------
// host drivers
// ---------------------------------------------------------------------
template <typename T>
static int gemm_sub(hfpx_ctx* ctx, T* C, int64_t ldc, const T* A, int64_t lda, const T* B,
                    int64_t ldb, int64_t M, int64_t N, int64_t K) {
  if (M <= 0 || N <= 0 || K <= 0) return HFPX_OK;
  const bool skinny = N <= 64;
  const int gtn = skinny ? 64 : 128;
  dim3 grid((unsigned)((M + GT - 1) / GT), (unsigned)((N + gtn - 1) / gtn));
  if (grid.y > 65535u) return set_error(ctx, HFPX_ERR_INVALID, "gemm: N too large");
  const int tok = prof_begin(ctx);
  if (skinny)
    hipLaunchKernelGGL((k_gemm_sub<T, 64>), grid, dim3(256), 0, ctx->stream, C, ldc, A, lda, B, ldb, M, N, K);
  else
    hipLaunchKernelGGL((k_gemm_sub<T, 128>), grid, dim3(256), 0, ctx->stream, C, ldc, A, lda, B, ldb, M, N, K);
  // algorithmic: 2MNK flops; bytes: C read + written, A and B read once
  prof_end(ctx, tok, sizeof(T) == 8 ? HFPX_KC_GEMM_F64 : HFPX_KC_GEMM_F32, 2.0 * M * N * K,
           (double)sizeof(T) * (2.0 * M * N + (double)M * K + (double)K * N));
  HFPX_TRY_HIP(ctx, hipGetLastError());
  return HFPX_OK;
}

template <typename T>
static int laswp(hfpx_ctx* ctx, T* A, int64_t lda, const int32_t* ipiv, int64_t k0, int64_t k1,
                 int64_t c0, int64_t c1) {
  if (c1 <= c0 || k1 <= k0) return HFPX_OK;
  const int bs = 256;
  const int tok = prof_begin(ctx);
  hipLaunchKernelGGL(k_laswp<T>, dim3((unsigned)((c1 - c0 + bs - 1) / bs)), dim3(bs), 0, ctx->stream,
                     A, lda, ipiv, k0, k1, c0, c1);
  prof_end(ctx, tok, HFPX_KC_LASWP, 0.0, 0.0);
  HFPX_TRY_HIP(ctx, hipGetLastError());
  return HFPX_OK;
}

template <typename T, bool LOWER>
static int trsm_leaf(hfpx_ctx* ctx, const T* L, int64_t ldl, T* B, int64_t ldb, int kb, int64_t ncols) {
  const int tok = prof_begin(ctx);
  hipLaunchKernelGGL((k_trsm_leaf<T, LOWER>), dim3((unsigned)((ncols + 255) / 256)), dim3(256), 0, ctx->stream,
                     L, ldl, B, ldb, kb, ncols);
  prof_end(ctx, tok, HFPX_KC_TRSM, (double)kb * kb * ncols, 0.0);
  HFPX_TRY_HIP(ctx, hipGetLastError());
  return HFPX_OK;
}

// B (kb x ncols) := L^{-1} B, L unit lower kb x kb: leaves of LB rows with
// MFMA updates from the rows already solved (left-looking).
template <typename T>
static int trsm_lunit(hfpx_ctx* ctx, const T* L, int64_t ldl, T* B, int64_t ldb, int kb, int64_t ncols) {
  if (ncols <= 0 || kb <= 0) return HFPX_OK;
  int rc;
  for (int i0 = 0; i0 < kb; i0 += LB) {
    const int ib = std::min(LB, kb - i0);
    if (i0 > 0 && (rc = gemm_sub<T>(ctx, B + i0, ldb, L + i0, ldl, B, ldb, ib, ncols, i0))) return rc;
    if ((rc = trsm_leaf<T, true>(ctx, L + i0 + (int64_t)i0 * ldl, ldl, B + i0, ldb, ib, ncols))) return rc;
  }
  return HFPX_OK;
}

// B (kb x ncols) := U^{-1} B, U upper (non-unit) kb x kb, bottom-up leaves.
template <typename T>
static int trsm_upper(hfpx_ctx* ctx, const T* U, int64_t ldu, T* B, int64_t ldb, int kb, int64_t ncols) {
  if (ncols <= 0 || kb <= 0) return HFPX_OK;
  int rc;
  const int nl = (kb + LB - 1) / LB;
  for (int b = nl - 1; b >= 0; --b) {
    const int i0 = b * LB, ib = std::min(LB, kb - i0);
    const int done = kb - (i0 + ib);
    if (done > 0 && (rc = gemm_sub<T>(ctx, B + i0, ldb, U + i0 + (int64_t)(i0 + ib) * ldu, ldu, B + i0 + ib, ldb,
                                      ib, ncols, done)))
      return rc;
    if ((rc = trsm_leaf<T, false>(ctx, U + i0 + (int64_t)i0 * ldu, ldu, B + i0, ldb, ib, ncols))) return rc;
  }
  return HFPX_OK;
}

template <typename T, int IB, int RPT>
static int panel_launch(hfpx_ctx* ctx, T* A, int64_t lda, int64_t n, int64_t j0, int jb,
                        int32_t* ipiv, int32_t* info, PanelScratch<T>* ps, T* rows, int nwg) {
  void* args[] = {&A, &lda, &n, &j0, &jb, &ipiv, &info, &ps, &rows};
  const int tok = prof_begin(ctx);
  const hipError_t e = hipLaunchCooperativeKernel((const void*)k_panel<T, IB, RPT>, dim3(nwg), dim3(PWG),
                                                  args, 0, ctx->stream);
  prof_end(ctx, tok, HFPX_KC_PANEL, (double)(n - j0) * jb * jb, 0.0);
  HFPX_TRY_HIP(ctx, e);
  return HFPX_OK;
}

template <typename T, int IB>
static int panel(hfpx_ctx* ctx, T* A, int64_t lda, int64_t n, int64_t j0, int jb, int32_t* ipiv,
                 int32_t* info, PanelScratch<T>* ps, T* rows, int max_wg) {
  const int64_t m = n - j0;
  // workgroups: enough that each thread holds at most RPT rows, at most max_wg
  int64_t nwg = (m + PWG - 1) / PWG;
  if (nwg > max_wg) nwg = max_wg;
  if (nwg < 1) nwg = 1;
  const int64_t chunk = (m + nwg - 1) / nwg;
  const int64_t rpt = (chunk + PWG - 1) / PWG;
  if (rpt <= 1) return panel_launch<T, IB, 1>(ctx, A, lda, n, j0, jb, ipiv, info, ps, rows, (int)nwg);
  if (rpt <= 2) return panel_launch<T, IB, 2>(ctx, A, lda, n, j0, jb, ipiv, info, ps, rows, (int)nwg);
  if (rpt <= 4) return panel_launch<T, IB, 4>(ctx, A, lda, n, j0, jb, ipiv, info, ps, rows, (int)nwg);
  return set_error(ctx, HFPX_ERR_INVALID, "getrf: n = %lld too large for the panel kernel", (long long)n);
}

template <typename T>
struct LuCfg;
template <>
struct LuCfg<double> {
  static constexpr int IB = 32;
  static constexpr int NB = 256;
};
template <>
struct LuCfg<float> {
  static constexpr int IB = 32;
  static constexpr int NB = 256;
};

template <typename T>
static int getrf_impl(hfpx_ctx* ctx, T* A, int64_t lda, int64_t n, int32_t* ipiv, int64_t* info_out) {
  constexpr int IB = LuCfg<T>::IB, NB = LuCfg<T>::NB;
  // co-resident panel workgroups
  int occ1 = 0, occ2 = 0, occ4 = 0;
  HFPX_TRY_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ1, k_panel<T, IB, 1>, PWG, 0));
  HFPX_TRY_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ2, k_panel<T, IB, 2>, PWG, 0));
  HFPX_TRY_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ4, k_panel<T, IB, 4>, PWG, 0));
  const int occ = std::min(occ1, std::min(occ2, occ4));
  if (occ < 1) return set_error(ctx, HFPX_ERR_HIP, "getrf: panel kernel cannot be resident");
  int max_wg = std::min(1024, occ * std::max(1, ctx->num_cus));
  const size_t rows_bytes = sizeof(T) * 2 * (size_t)(max_wg + 1) * IB;
  const size_t need = sizeof(PanelScratch<T>) + rows_bytes + 256;
  int rc = ensure_work(ctx, need + 64);
  if (rc) return rc;
  char* w = (char*)ctx->work;
  int32_t* dinfo = (int32_t*)w;
  PanelScratch<T>* ps = (PanelScratch<T>*)(w + 256);
  T* rows = (T*)(w + 256 + sizeof(PanelScratch<T>));
  HFPX_TRY_HIP(ctx, hipMemsetAsync(dinfo, 0, sizeof(int32_t), ctx->stream));

  for (int64_t k = 0; k < n; k += NB) {
    const int64_t kb = std::min<int64_t>(NB, n - k);
    for (int64_t j = k; j < k + kb; j += IB) {
      const int jb = (int)std::min<int64_t>(IB, k + kb - j);
      rc = panel<T, IB>(ctx, A, lda, n, j, jb, ipiv, dinfo, ps, rows, max_wg);
      if (rc) return rc;
      // interchanges on the outer panel's other columns
      if ((rc = laswp<T>(ctx, A, lda, ipiv, j, j + jb, k, j))) return rc;
      if ((rc = laswp<T>(ctx, A, lda, ipiv, j, j + jb, j + jb, k + kb))) return rc;
      const int64_t rest = k + kb - (j + jb);
      if (rest > 0) {
        T* a12 = A + j + (j + jb) * lda;
        if ((rc = trsm_lunit<T>(ctx, A + j + j * lda, lda, a12, lda, jb, rest))) return rc;
        if ((rc = gemm_sub<T>(ctx, A + (j + jb) + (j + jb) * lda, lda, A + (j + jb) + j * lda, lda, a12,
                              lda, n - (j + jb), rest, jb)))
          return rc;
      }
    }
    // interchanges on the columns left and right of the outer panel
    if ((rc = laswp<T>(ctx, A, lda, ipiv, k, k + kb, 0, k))) return rc;
    if ((rc = laswp<T>(ctx, A, lda, ipiv, k, k + kb, k + kb, n))) return rc;
    const int64_t rest = n - (k + kb);
    if (rest > 0) {
      T* a12 = A + k + (k + kb) * lda;
      if ((rc = trsm_lunit<T>(ctx, A + k + k * lda, lda, a12, lda, (int)kb, rest))) return rc;
      if ((rc = gemm_sub<T>(ctx, A + (k + kb) + (k + kb) * lda, lda, A + (k + kb) + k * lda, lda, a12,
                            lda, rest, rest, kb)))
        return rc;
    }
  }
  int32_t hinfo = 0;
  HFPX_TRY_HIP(ctx, hipMemcpyAsync(&hinfo, dinfo, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
  HFPX_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (info_out) *info_out = hinfo;
  return hinfo ? HFPX_ERR_SINGULAR : HFPX_OK;
}

// B (n x nrhs) := A^{-1} B using the factors (dgetrs 'N').
template <typename T>
static int getrs_impl(hfpx_ctx* ctx, const T* LU, int64_t lda, const int32_t* ipiv, int64_t n,
                      T* B, int64_t ldb, int64_t nrhs) {
  constexpr int NB = 256;
  int rc;
  // row interchanges on B (one thread per right-hand side)
  if ((rc = laswp<T>(ctx, B, ldb, ipiv, 0, n, 0, nrhs))) return rc;
  // forward substitution with unit L, blocked
  for (int64_t k = 0; k < n; k += NB) {
    const int64_t kb = std::min<int64_t>(NB, n - k);
    if ((rc = trsm_lunit<T>(ctx, LU + k + k * lda, lda, B + k, ldb, (int)kb, nrhs))) return rc;
    if (k + kb < n)
      if ((rc = gemm_sub<T>(ctx, B + k + kb, ldb, LU + (k + kb) + k * lda, lda, B + k, ldb, n - k - kb, nrhs,
                            kb)))
        return rc;
  }
  // back substitution with U, blocked (bottom-up)
  const int64_t nblk = (n + NB - 1) / NB;
  for (int64_t b = nblk - 1; b >= 0; --b) {
    const int64_t k = b * NB, kb = std::min<int64_t>(NB, n - k);
    if ((rc = trsm_upper<T>(ctx, LU + k + k * lda, lda, B + k, ldb, (int)kb, nrhs))) return rc;
    if (k > 0)
      if ((rc = gemm_sub<T>(ctx, B, ldb, LU + k * lda, lda, B + k, ldb, k, nrhs, kb))) return rc;
  }
  return HFPX_OK;
}

template int getrf_impl<double>(hfpx_ctx*, double*, int64_t, int64_t, int32_t*, int64_t*);
template int getrf_impl<float>(hfpx_ctx*, float*, int64_t, int64_t, int32_t*, int64_t*);

int lu_getrf(hfpx_ctx* ctx, int dtype, void* A, int64_t lda, int64_t n, int32_t* ipiv, int64_t* info) {
  if (dtype == HFPX_F64) return getrf_impl<double>(ctx, (double*)A, lda, n, ipiv, info);
  return getrf_impl<float>(ctx, (float*)A, lda, n, ipiv, info);
}

int lu_getrs(hfpx_ctx* ctx, int dtype, const void* LU, int64_t lda, const int32_t* ipiv, int64_t n,
             void* B, int64_t ldb, int64_t nrhs) {
  if (dtype == HFPX_F64) return getrs_impl<double>(ctx, (const double*)LU, lda, ipiv, n, (double*)B, ldb, nrhs);
  return getrs_impl<float>(ctx, (const float*)LU, lda, ipiv, n, (float*)B, ldb, nrhs);
}

int gemm_sub_any(hfpx_ctx* ctx, int dtype, void* C, int64_t ldc, const void* A, int64_t lda,
                 const void* B, int64_t ldb, int64_t M, int64_t N, int64_t K) {
  if (dtype == HFPX_F64)
    return gemm_sub<double>(ctx, (double*)C, ldc, (const double*)A, lda, (const double*)B, ldb, M, N, K);
  return gemm_sub<float>(ctx, (float*)C, ldc, (const float*)A, lda, (const float*)B, ldb, M, N, K);
}

}  // namespace hfpx
