// hfpx_gemm.hip -- C -= A B on the MFMA units of gfx950, FP64 and FP32:
// the Schur-complement updates of the LU and the blocked triangular solves,
// and the FP64 residuals of the mixed-precision solve.
//
// No direct reference counterpart: inside LAPACKE_dgetrf/dgetrs
// (il/linear_algebra/dense/factorization/LU.h:79-138) this is dgemm.
#include <algorithm>

#include "hfpx_internal.h"

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
// One 128 x GTN tile of C at (row0, col0).
template <typename T, int GTN>
__device__ __forceinline__ void gemm_tile(T* __restrict__ C, int64_t ldc, const T* __restrict__ A, int64_t lda,
                                          const T* __restrict__ B, int64_t ldb, int64_t M, int64_t N, int64_t K,
                                          int64_t row0, int64_t col0) {
  typedef typename MfmaVec<T>::type vec4;
  constexpr int WN = GTN / 2;      // wave tile columns
  constexpr int TC = WN / 16;      // 16x16 MFMA tiles across a wave's columns
  constexpr int BQ = KC * GTN / 256;  // B elements loaded per thread per chunk
  __shared__ T sA[2][KC][GT];      // sA[k][i] = A(row0 + i, k0 + k)
  __shared__ T sB[2][KC][GTN];     // sB[k][j] = B(k0 + k, col0 + j)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
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

template <typename T, int GTN>
__global__ __launch_bounds__(256) void k_gemm_sub(T* __restrict__ C, int64_t ldc,
                                                  const T* __restrict__ A, int64_t lda,
                                                  const T* __restrict__ B, int64_t ldb, int64_t M,
                                                  int64_t N, int64_t K, int64_t kchunk, int64_t cz) {
  // split K: slice z = blockIdx.z covers k in [z kchunk, (z+1) kchunk) and
  // writes C + z cz (cz = 0 and one slice: the plain product)
  const int64_t z = blockIdx.z, kz = z * kchunk;
  gemm_tile<T, GTN>(C + z * cz, ldc, A + kz * lda, lda, B + kz, ldb, M, N, (K - kz) < kchunk ? (K - kz) : kchunk,
                    (int64_t)blockIdx.x * GT, (int64_t)blockIdx.y * GTN);
}

// ---------------------------------------------------------------------
// FP32 C (M x N) -= A (M x K) * B (K x N) on v_mfma_f32_32x32x2f32, for the
// large trailing updates of the LU.
//
// Workgroup = 4 waves in a 2 x 2 arrangement, wave tile (32 WR) x (32 WC),
// workgroup tile GM = 64 WR rows by GN = 64 WC columns; K in chunks of KC
// through double-buffered LDS (one barrier per chunk).  As above, the MFMA
// forms the transposed product: operand a (32 x 2) is a block of B^T, operand
// b (2 x 32) a block of A^T, so that lane l % 32 indexes 32 consecutive rows
// of C and every result element of a lane group lands in one 128-byte column
// segment.  32x32x2 operand layout: lane l supplies a[l%32][l/32] and
// b[l/32][l%32]; result element v of lane l is D[8(v/4) + 4(l/32) + v%4][l%32].
// LDS rows are padded (A by 32 floats, B by 33) so that the transposing B
// stores and the two half-waves' operand reads fall in distinct banks.
// ---------------------------------------------------------------------
typedef float floatx16 __attribute__((ext_vector_type(16)));

// Options of k_sgemm32 (bit flags):
//  SG_INT   interior launch: every tile lies wholly inside M x N (the host
//           launches the edge bands separately), so loads and stores carry
//           no row/column guards; only a K chunk that crosses K checks k;
//  SG_PIPE  the LDS operands of k-step s+1 are read while the MFMAs of step s
//           run (register double buffer);
//  SG_SWZ   1-D grid remapped so that each of the 8 XCDs (workgroups are
//           dealt to XCDs round-robin) owns one contiguous eighth of the
//           tiles, walked in bands of 8 row tiles, so the A and B panels that
//           concurrent workgroups share stay in that XCD's L2;
//  SG_CPRE  C preloaded into the accumulators (no read-modify-write at the end).
constexpr int SG_INT = 1, SG_PIPE = 2, SG_SWZ = 4, SG_CPRE = 8;

// LDS footprint of one 32x32x2 tile (floats per buffer, two buffers).
template <int WR, int WC, int KC>
struct Sg32 {
  static constexpr int GM = 64 * WR, GN = 64 * WC;
  static constexpr int PA = GM + 32, PB = GN + 33;
  static constexpr int LA = KC * PA, LB = KC * PB;
};

// One GM x GN tile of C at (row0, col0); sA and sB are the caller's LDS
// buffers (2 LA and 2 LB floats).  With SG_INT the tile must lie wholly
// inside M x N.
template <int WR, int WC, int KC, int OPT>
__device__ __forceinline__ void sgemm32_tile(float* __restrict__ sA, float* __restrict__ sB,
                                             float* __restrict__ C, int64_t ldc, const float* __restrict__ A,
                                             int64_t lda, const float* __restrict__ B, int64_t ldb, int64_t M,
                                             int64_t N, int64_t K, int64_t row0, int64_t col0) {
  typedef Sg32<WR, WC, KC> S;
  constexpr int GM = S::GM, GN = S::GN, PA = S::PA, PB = S::PB, LA = S::LA, LB = S::LB;
  constexpr int QA = GM * KC / 256, QB = GN * KC / 256;
  static_assert(QA >= 1 && QB >= 1, "tile too small for the loaders");
  static_assert(256 % GM == 0 && 256 % KC == 0, "loader strides");
  // sA[buf LA + k PA + i] = A(row0 + i, k0 + k); sB[buf LB + k PB + j] = B(k0 + k, col0 + j)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = (wave & 1) * 32 * WR, wc = (wave >> 1) * 32 * WC;

  const int l32 = lane & 31, hk = lane >> 5;
  // acc[tc][tr][v] holds element (row0 + wr + 32 tr + l32, col0 + wc + 32 tc
  // + 8 (v/4) + 4 hk + v%4) of A B, or with SG_CPRE of C - A B: the
  // accumulators start as the tile of C (loaded before the first operands,
  // so that its latency overlaps theirs), B is staged negated and the
  // epilogue is a plain store.
  // Tile addresses are 32-bit offsets from the tile's corner (the host
  // keeps GN ldc below 2^31).
  float* __restrict__ Ct = C + row0 + col0 * ldc;
  const int ld = (int)ldc;
  const int off0 = wr + l32 + (wc + 4 * hk) * ld;
  const int mrem = (int)((M - row0) < GM ? (M - row0) : GM), nrem = (int)((N - col0) < GN ? (N - col0) : GN);
  floatx16 acc[WC][WR];
#pragma unroll
  for (int tc = 0; tc < WC; ++tc)
#pragma unroll
    for (int tr = 0; tr < WR; ++tr) {
      const int i = wr + 32 * tr + l32;
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int jc = wc + 32 * tc + 8 * (v >> 2) + 4 * hk + (v & 3);
        const int off = off0 + 32 * tr + (32 * tc + 8 * (v >> 2) + (v & 3)) * ld;
        if constexpr ((OPT & SG_CPRE) != 0)
          acc[tc][tr][v] = ((OPT & SG_INT) != 0 || (i < mrem && jc < nrem)) ? Ct[off] : 0.f;
        else
          acc[tc][tr][v] = 0.f;
      }
    }

  float ra[QA], rb[QB];
  auto load_global = [&](int64_t k0) {
    if constexpr ((OPT & SG_INT) != 0) {
      // e = tid + 256 q: A row tid % GM, k = tid / GM + q (256 / GM);
      // B k = tid % KC, column tid / KC + q (256 / KC)
      const float* pa = A + (row0 + tid % GM) + (k0 + tid / GM) * lda;
      const float* pb = B + (k0 + tid % KC) + (col0 + tid / KC) * ldb;
      if (k0 + KC <= K) {
#pragma unroll
        for (int q = 0; q < QA; ++q) ra[q] = pa[(int64_t)q * (256 / GM) * lda];
#pragma unroll
        for (int q = 0; q < QB; ++q) rb[q] = pb[(int64_t)q * (256 / KC) * ldb];
      } else {
#pragma unroll
        for (int q = 0; q < QA; ++q)
          ra[q] = (k0 + tid / GM + q * (256 / GM) < K) ? pa[(int64_t)q * (256 / GM) * lda] : 0.f;
        const bool kin = k0 + tid % KC < K;
#pragma unroll
        for (int q = 0; q < QB; ++q) rb[q] = kin ? pb[(int64_t)q * (256 / KC) * ldb] : 0.f;
      }
      return;
    }
#pragma unroll
    for (int q = 0; q < QA; ++q) {
      const int e = tid + 256 * q;  // consecutive lanes: consecutive rows of one column
      const int i = e % GM, k = e / GM;
      const int64_t gr = row0 + i, gk = k0 + k;
      ra[q] = (gr < M && gk < K) ? A[gr + gk * lda] : 0.f;
    }
#pragma unroll
    for (int q = 0; q < QB; ++q) {
      const int e = tid + 256 * q;  // consecutive lanes: consecutive k of one column
      const int k = e % KC, j = e / KC;
      const int64_t gc = col0 + j, gk = k0 + k;
      rb[q] = (gc < N && gk < K) ? B[gk + gc * ldb] : 0.f;
    }
  };
  auto store_lds = [&](int buf) {
#pragma unroll
    for (int q = 0; q < QA; ++q) {
      const int e = tid + 256 * q;
      sA[buf * LA + (e / GM) * PA + (e % GM)] = ra[q];
    }
#pragma unroll
    for (int q = 0; q < QB; ++q) {
      const int e = tid + 256 * q;
      sB[buf * LB + (e % KC) * PB + (e / KC)] = (OPT & SG_CPRE) != 0 ? -rb[q] : rb[q];
    }
  };

  const int nchunks = (int)((K + KC - 1) / KC);
  load_global(0);
  store_lds(0);
  __syncthreads();
  for (int c = 0; c < nchunks; ++c) {
    const int buf = c & 1;
    if (c + 1 < nchunks) load_global((int64_t)(c + 1) * KC);
    const float* pa = sA + buf * LA + wr + l32;
    const float* pb = sB + buf * LB + wc + l32;
    if constexpr ((OPT & SG_PIPE) != 0) {
      float fa[2][WC], fb[2][WR];
#pragma unroll
      for (int t = 0; t < WC; ++t) fa[0][t] = pb[hk * PB + 32 * t];
#pragma unroll
      for (int t = 0; t < WR; ++t) fb[0][t] = pa[hk * PA + 32 * t];
#pragma unroll
      for (int s = 0; s < KC / 2; ++s) {
        const int cur = s & 1;
        if (s + 1 < KC / 2) {
          const int k = 2 * (s + 1) + hk;
#pragma unroll
          for (int t = 0; t < WC; ++t) fa[cur ^ 1][t] = pb[k * PB + 32 * t];
#pragma unroll
          for (int t = 0; t < WR; ++t) fb[cur ^ 1][t] = pa[k * PA + 32 * t];
        }
#pragma unroll
        for (int tc = 0; tc < WC; ++tc)
#pragma unroll
          for (int tr = 0; tr < WR; ++tr)
            acc[tc][tr] = __builtin_amdgcn_mfma_f32_32x32x2f32(fa[cur][tc], fb[cur][tr], acc[tc][tr], 0, 0, 0);
      }
    } else {
#pragma unroll
      for (int s = 0; s < KC / 2; ++s) {
        const int k = 2 * s + hk;
        float fa[WC], fb[WR];
#pragma unroll
        for (int t = 0; t < WC; ++t) fa[t] = pb[k * PB + 32 * t];  // C columns
#pragma unroll
        for (int t = 0; t < WR; ++t) fb[t] = pa[k * PA + 32 * t];  // C rows
#pragma unroll
        for (int tc = 0; tc < WC; ++tc)
#pragma unroll
          for (int tr = 0; tr < WR; ++tr)
            acc[tc][tr] = __builtin_amdgcn_mfma_f32_32x32x2f32(fa[tc], fb[tr], acc[tc][tr], 0, 0, 0);
      }
    }
    if (c + 1 < nchunks) store_lds(buf ^ 1);
    __syncthreads();
  }
#pragma unroll
  for (int tc = 0; tc < WC; ++tc)
#pragma unroll
    for (int tr = 0; tr < WR; ++tr) {
      const int i = wr + 32 * tr + l32;
      if ((OPT & SG_INT) == 0 && i >= mrem) continue;
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int jc = wc + 32 * tc + 8 * (v >> 2) + 4 * hk + (v & 3);
        const int off = off0 + 32 * tr + (32 * tc + 8 * (v >> 2) + (v & 3)) * ld;
        if ((OPT & SG_INT) != 0 || jc < nrem) {
          if constexpr ((OPT & SG_CPRE) != 0)
            Ct[off] = acc[tc][tr][v];
          else
            Ct[off] -= acc[tc][tr][v];
        }
      }
    }
}

template <int WR, int WC, int KC, int OPT>
__global__ __launch_bounds__(256) void k_sgemm32(float* __restrict__ C, int64_t ldc,
                                                 const float* __restrict__ A, int64_t lda,
                                                 const float* __restrict__ B, int64_t ldb, int64_t M,
                                                 int64_t N, int64_t K) {
  typedef Sg32<WR, WC, KC> S;
  __shared__ float sA[2 * S::LA];
  __shared__ float sB[2 * S::LB];
  int64_t tile_r, tile_c;
  if constexpr ((OPT & SG_SWZ) != 0) {
    const int64_t tm = (M + S::GM - 1) / S::GM, tn = (N + S::GN - 1) / S::GN;
    sg_band_tile(sg_xcd_order(blockIdx.x, tm * tn), tm, tn, tile_r, tile_c);
  } else {
    tile_r = blockIdx.x;
    tile_c = blockIdx.y;
  }
  sgemm32_tile<WR, WC, KC, OPT>(sA, sB, C, ldc, A, lda, B, ldb, M, N, K, tile_r * S::GM, tile_c * S::GN);
}

// The same for the larger tiles (WR WC > 4), whose accumulators would
// otherwise be placed in AGPRs beside a full VGPR set (one wave per SIMD):
// bounded to two waves per SIMD, the compiler keeps everything in VGPRs.
template <int WR, int WC, int KC, int OPT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_sgemm32_w2(float* __restrict__ C, int64_t ldc,
                                                 const float* __restrict__ A, int64_t lda,
                                                 const float* __restrict__ B, int64_t ldb, int64_t M,
                                                 int64_t N, int64_t K) {
  typedef Sg32<WR, WC, KC> S;
  __shared__ float sA[2 * S::LA];
  __shared__ float sB[2 * S::LB];
  int64_t tile_r, tile_c;
  if constexpr ((OPT & SG_SWZ) != 0) {
    const int64_t tm = (M + S::GM - 1) / S::GM, tn = (N + S::GN - 1) / S::GN;
    sg_band_tile(sg_xcd_order(blockIdx.x, tm * tn), tm, tn, tile_r, tile_c);
  } else {
    tile_r = blockIdx.x;
    tile_c = blockIdx.y;
  }
  sgemm32_tile<WR, WC, KC, OPT>(sA, sB, C, ldc, A, lda, B, ldb, M, N, K, tile_r * S::GM, tile_c * S::GN);
}

// ---------------------------------------------------------------------
// FP64 C (M x N) -= A (M x K) * B (K x N) for the large Schur-complement
// updates of the FP64 LU, on v_mfma_f64_16x16x4f64.
//
// Workgroup tile 128 x 128 (4 waves of 64 x 64, sixteen 16 x 16 MFMA tiles
// each), K in chunks of 16 through double-buffered LDS, one barrier per
// chunk.  The kernel is bounded to two waves per SIMD (accumulators in
// AGPRs, at most 256 registers per lane), so two workgroups share a CU and
// one's loads, LDS stores and barrier overlap the other's MFMAs: with one
// wave per SIMD every barrier idles the MFMA pipe.  As in gemm_tile the
// MFMA forms the transposed product, so 16 lanes write 16 consecutive rows
// of one column of C.  LDS rows are padded by 16 doubles (half of the
// banks), so the two half-waves' operand reads of rows k and k + 1 fall in
// distinct banks, and the B tile's columns are XOR-swizzled within groups
// of 16 by (k & 14): its transposing stores (16 k rows per half-wave) then
// hit distinct banks too, where unswizzled rows k and k + 2 share them
// (measured: a third of the cycles stalled on bank conflicts).  Options as
// k_sgemm32: SG_INT (interior tiles, no
// row / column guards), SG_SWZ (XCD-grouped walk in bands of `band` row
// tiles), SG_CPRE (C preloaded into the accumulators, B staged negated,
// plain store at the end).
// ---------------------------------------------------------------------
constexpr int DG_T = 128, DG_KC = 16, DG_P = DG_T + 16;

template <int OPT>
__device__ __forceinline__ void dgemm128_tile(double* __restrict__ sA, double* __restrict__ sB, double* __restrict__ C,
                                              int64_t ldc, const double* __restrict__ A, int64_t lda,
                                              const double* __restrict__ B, int64_t ldb, int64_t M, int64_t N,
                                              int64_t K, int64_t row0, int64_t col0) {
  // sA[buf][k][i] = A(row0 + i, k0 + k), sB[buf][k][j] = (+-) B(k0 + k, col0 + j)
  constexpr int L = DG_KC * DG_P;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = (wave & 1) * 64, wc = (wave >> 1) * 64;
  const int l16 = lane & 15, lq = lane >> 4;
  const int mrem = (int)((M - row0) < DG_T ? (M - row0) : DG_T);
  const int nrem = (int)((N - col0) < DG_T ? (N - col0) : DG_T);
  constexpr bool INT = (OPT & SG_INT) != 0, CPRE = (OPT & SG_CPRE) != 0;

  // acc[tc][tr][v]: row wr + 16 tr + l16, column wc + 16 tc + 4 v + lq.
  // C is addressed by 32-bit offsets from the tile's corner (the host keeps
  // 128 ldc below 2^31), so the 64 addresses cost no 64-bit registers.
  double* __restrict__ Ct = C + row0 + col0 * ldc;
  const int ld = (int)ldc;
  const int off0 = wr + l16 + (wc + lq) * ld;
  doublex4 acc[4][4];
#pragma unroll
  for (int tc = 0; tc < 4; ++tc)
#pragma unroll
    for (int tr = 0; tr < 4; ++tr)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = wr + 16 * tr + l16, j = wc + 16 * tc + 4 * v + lq;
        if constexpr (CPRE)
          acc[tc][tr][v] = (INT || (i < mrem && j < nrem)) ? Ct[off0 + 16 * tr + (16 * tc + 4 * v) * ld] : 0.0;
        else
          acc[tc][tr][v] = 0.0;
      }

  // loaders: A row tid % 128, k = tid / 128 + 2 q; B k = tid % 16, column
  // tid / 16 + 16 q (consecutive lanes: consecutive addresses)
  const int ai = tid & 127, ak = tid >> 7, bk = tid & 15, bj = tid >> 4;
  const double* pa = A + (row0 + ai) + (int64_t)ak * lda;
  const double* pb = B + bk + (col0 + bj) * ldb;
  double ra[8], rb[8];
  auto load_global = [&](int64_t k0) {
    const bool kfull = k0 + DG_KC <= K;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const bool ok = (INT || ai < mrem) && (kfull || k0 + ak + 2 * q < K);
      ra[q] = ok ? pa[(k0 + 2 * q) * lda] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const bool ok = (INT || bj + 16 * q < nrem) && (kfull || k0 + bk < K);
      rb[q] = ok ? pb[k0 + (int64_t)(16 * q) * ldb] : 0.0;
    }
  };
  auto store_lds = [&](int buf) {
#pragma unroll
    for (int q = 0; q < 8; ++q) sA[buf * L + (ak + 2 * q) * DG_P + ai] = ra[q];
    const int swz = bk & 14;  // column swizzle of row bk (bj < 16)
#pragma unroll
    for (int q = 0; q < 8; ++q) sB[buf * L + bk * DG_P + 16 * q + (bj ^ swz)] = CPRE ? -rb[q] : rb[q];
  };

  const int nchunks = (int)((K + DG_KC - 1) / DG_KC);
  load_global(0);
  store_lds(0);
  __syncthreads();
  for (int c = 0; c < nchunks; ++c) {
    const int buf = c & 1;
    if (c + 1 < nchunks) load_global((int64_t)(c + 1) * DG_KC);
    const double* qa = sA + buf * L + wr + l16;
    const double* qb = sB + buf * L + wc;
#pragma unroll
    for (int k4 = 0; k4 < DG_KC; k4 += 4) {
      const int kk = (k4 + lq) * DG_P;
      const int cb = l16 ^ ((k4 + lq) & 14);  // the store's swizzle of row k4 + lq
      double fu[4], fl[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) fu[t] = qb[kk + 16 * t + cb];  // B^T rows = C columns
#pragma unroll
      for (int t = 0; t < 4; ++t) fl[t] = qa[kk + 16 * t];  // A^T columns = C rows
#pragma unroll
      for (int tc = 0; tc < 4; ++tc)
#pragma unroll
        for (int tr = 0; tr < 4; ++tr)
          acc[tc][tr] = __builtin_amdgcn_mfma_f64_16x16x4f64(fu[tc], fl[tr], acc[tc][tr], 0, 0, 0);
    }
    if (c + 1 < nchunks) store_lds(buf ^ 1);
    __syncthreads();
  }
#pragma unroll
  for (int tc = 0; tc < 4; ++tc)
#pragma unroll
    for (int tr = 0; tr < 4; ++tr) {
      const int i = wr + 16 * tr + l16;
      if (!INT && i >= mrem) continue;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int j = wc + 16 * tc + 4 * v + lq;
        if (INT || j < nrem) {
          double* pc = Ct + off0 + 16 * tr + (16 * tc + 4 * v) * ld;
          if constexpr (CPRE)
            *pc = acc[tc][tr][v];
          else
            *pc -= acc[tc][tr][v];
        }
      }
    }
}

template <int OPT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_dgemm128(
    double* __restrict__ C, int64_t ldc, const double* __restrict__ A, int64_t lda, const double* __restrict__ B,
    int64_t ldb, int64_t M, int64_t N, int64_t K, int band) {
  __shared__ double sA[2 * DG_KC * DG_P];
  __shared__ double sB[2 * DG_KC * DG_P];
  const int64_t tm = (M + DG_T - 1) / DG_T, tn = (N + DG_T - 1) / DG_T;
  int64_t tile_r, tile_c;
  if constexpr ((OPT & SG_SWZ) != 0) {
    sg_band_tile(sg_xcd_order(blockIdx.x, tm * tn), tm, tn, tile_r, tile_c, band);
  } else {
    tile_r = blockIdx.x % tm;
    tile_c = blockIdx.x / tm;
  }
  dgemm128_tile<OPT>(sA, sB, C, ldc, A, lda, B, ldb, M, N, K, tile_r * DG_T, tile_c * DG_T);
}

// ---------------------------------------------------------------------
// FP64 P_z = -A(:, slice z) X(slice z, :) for N <= 64 and long K (the
// residuals b - A x of the mixed-precision solve): A is streamed once, so
// every lane loads its MFMA operands straight from HBM into registers (no
// LDS), while the slice of X is staged through LDS for all four waves.
// Wave w owns rows 32 w .. 32 w + 31 of the workgroup's 128; columns are
// NB blocks of 16 (N rounded up), so even a single right-hand side keeps the
// kernel bound by HBM rather than by v_mfma_f64_16x16x4f64.
// Operand layout: lane l supplies A[i = l%16][k = l/16] and X[k = l/16][j =
// l%16]; result element v of lane l is D[4v + l/16][l%16].
// ---------------------------------------------------------------------
// rabs (optional): the row sums of |A| over the slice, rabs[z M + row],
// from the A values the product streams anyway (the normwise backward
// error's |A|_inf without a pass of its own over A).
template <int NB>
__global__ __launch_bounds__(256) void k_dgemm_stream(double* __restrict__ P, const double* __restrict__ A,
                                                      int64_t lda, const double* __restrict__ X, int64_t ldx,
                                                      int64_t M, int64_t N, int64_t K, int64_t kslice,
                                                      double* __restrict__ rabs) {
  constexpr int KC2 = 16;
  constexpr int XQ = KC2 * NB * 16 / 256;  // X elements staged per thread per chunk
  __shared__ double sX[2][KC2][NB * 16 + 1];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, l16 = lane & 15, lq = lane >> 4;
  const int64_t row0 = (int64_t)blockIdx.x * 128 + 32 * wave;
  const int64_t z = blockIdx.y, kb = z * kslice;
  const int64_t ke = (K - kb) < kslice ? K : kb + kslice;
  doublex4 acc[2][NB];
#pragma unroll
  for (int rb = 0; rb < 2; ++rb)
#pragma unroll
    for (int cb = 0; cb < NB; ++cb) acc[rb][cb] = doublex4{0, 0, 0, 0};
  const int64_t r_a = row0 + l16, r_b = row0 + 16 + l16;
  const bool ok_a = r_a < M, ok_b = r_b < M;
  // A is fetched two chunks ahead into two static register sets (the loop
  // is unrolled by two so each set keeps its registers), the X slice one
  // chunk ahead; X is issued before A so that waiting for it never waits
  // for the younger A loads
  double a0[2][4], a1[2][4], xn[XQ];
  auto fetch_a = [&](double (&dst)[2][4], int64_t k0) {
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int64_t k = k0 + 4 * s + lq;
      const bool kin = k < ke;
      dst[0][s] = (ok_a && kin) ? A[r_a + k * lda] : 0.0;
      dst[1][s] = (ok_b && kin) ? A[r_b + k * lda] : 0.0;
    }
  };
  auto fetch_x = [&](int64_t k0) {
#pragma unroll
    for (int q = 0; q < XQ; ++q) {
      const int e = tid + 256 * q, kk = e % KC2, j = e / KC2;
      const int64_t k = k0 + kk;
      xn[q] = (k < ke && j < N) ? X[k + j * ldx] : 0.0;
    }
  };
  fetch_x(kb);
  fetch_a(a0, kb);
  if (kb + KC2 < ke) fetch_a(a1, kb + KC2);
  int buf = 0;
  double sa = 0.0, sb = 0.0;  // |A| row sums of rows r_a, r_b over this lane's k
  auto step = [&](double (&cur)[2][4], int64_t k0) {
    double av[2][4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      av[0][s] = cur[0][s];
      av[1][s] = cur[1][s];
    }
    if (rabs) {
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        sa += fabs(av[0][s]);
        sb += fabs(av[1][s]);
      }
    }
#pragma unroll
    for (int q = 0; q < XQ; ++q) {
      const int e = tid + 256 * q;
      sX[buf][e % KC2][e / KC2] = xn[q];
    }
    __syncthreads();  // one barrier per chunk: sX is double-buffered
    if (k0 + KC2 < ke) fetch_x(k0 + KC2);
    if (k0 + 2 * KC2 < ke) fetch_a(cur, k0 + 2 * KC2);
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int cb = 0; cb < NB; ++cb) {
        const double xv = sX[buf][4 * s + lq][16 * cb + l16];
#pragma unroll
        for (int rb = 0; rb < 2; ++rb)
          acc[rb][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[rb][s], xv, acc[rb][cb], 0, 0, 0);
      }
    buf ^= 1;
  };
  for (int64_t k0 = kb; k0 < ke; k0 += 2 * KC2) {
    step(a0, k0);
    if (k0 + KC2 < ke) step(a1, k0 + KC2);
  }
  if (rabs) {
    // the four k groups of a row are lanes l16 + 16 lq
    sa += __shfl_xor(sa, 16);
    sa += __shfl_xor(sa, 32);
    sb += __shfl_xor(sb, 16);
    sb += __shfl_xor(sb, 32);
    if (lq == 0) {
      if (ok_a) rabs[z * M + r_a] = sa;
      if (ok_b) rabs[z * M + r_b] = sb;
    }
  }
  double* Pz = P + z * M * N;
#pragma unroll
  for (int rb = 0; rb < 2; ++rb)
#pragma unroll
    for (int cb = 0; cb < NB; ++cb)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t r = row0 + 16 * rb + 4 * v + lq;
        const int64_t c = 16 * cb + l16;
        if (r < M && c < N) Pz[r + c * M] = -acc[rb][cb][v];
      }
}

// out[i] = sum over slices z of part[z M + i], in slice order.
__global__ void k_slice_rowsum(const double* __restrict__ part, int64_t M, int ks, double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  double s = 0.0;
  for (int z = 0; z < ks; ++z) s += part[z * M + i];
  out[i] = s;
}

// C (M x N, ldc) += sum over slices z = 0 .. ks-1 of P_z (M x N, ld M),
// in slice order (deterministic).
template <typename T>
__global__ __launch_bounds__(256) void k_splitk_sum(T* __restrict__ C, int64_t ldc, const T* __restrict__ P,
                                                    int64_t M, int64_t N, int ks) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= M * N) return;
  const int64_t i = e % M, j = e / M;
  T s = C[i + j * ldc];
  for (int z = 0; z < ks; ++z) s += P[e + (int64_t)z * M * N];
  C[i + j * ldc] = s;
}

constexpr int64_t SPLITK_MIN_K = 8192;


constexpr int64_t SPLITK_MAX = 32;

// Grow-only workspace of the split-K products (not ctx->work, which callers
// of gemm_sub may be holding).
static int ensure_gemm_work(hfpx_ctx* ctx, size_t bytes) {
  if (ctx->gemm_work_bytes >= bytes) return HFPX_OK;
  if (ctx->gemm_work) {
    HFPX_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    HFPX_TRY_HIP(ctx, hipFree(ctx->gemm_work));
    ctx->gemm_work = nullptr;
    ctx->gemm_work_bytes = 0;
  }
  HFPX_TRY_HIP(ctx, hipMalloc(&ctx->gemm_work, bytes));
  ctx->gemm_work_bytes = bytes;
  return HFPX_OK;
}

bool dgemm_stream_applies(int64_t N, int64_t K) { return N <= 64 && K >= SPLITK_MIN_K; }

// C -= A B for few columns (N <= 64) and long K, FP64: A streamed once in
// 128-row tiles x K slices, the slices' partial products summed in order;
// with rowabs, also rowabs[i] = sum_k |A(i, k)| (summed over the slices in
// order).
int dgemm_stream_sub(hfpx_ctx* ctx, double* C, int64_t ldc, const double* A, int64_t lda, const double* B,
                     int64_t ldb, int64_t M, int64_t N, int64_t K, double* rowabs) {
  const int64_t nrt = (M + 127) / 128;
  int64_t ks = std::min<int64_t>(SPLITK_MAX, std::max<int64_t>(1, (8 * ctx->num_cus + nrt - 1) / nrt));
  const int64_t kslice = ((K + ks - 1) / ks + 15) / 16 * 16;
  ks = (K + kslice - 1) / kslice;
  const size_t slice = (size_t)M * N;
  int rc = ensure_gemm_work(ctx, sizeof(double) * (slice * ks + (rowabs ? (size_t)M * ks : 0)));
  if (rc) return rc;
  double* P = (double*)ctx->gemm_work;
  double* RA = rowabs ? P + slice * ks : nullptr;
  const dim3 g((unsigned)nrt, (unsigned)ks);
  if (N <= 16)
    hipLaunchKernelGGL(k_dgemm_stream<1>, g, dim3(256), 0, ctx->stream, P, A, lda, B, ldb, M, N, K, kslice, RA);
  else if (N <= 32)
    hipLaunchKernelGGL(k_dgemm_stream<2>, g, dim3(256), 0, ctx->stream, P, A, lda, B, ldb, M, N, K, kslice, RA);
  else
    hipLaunchKernelGGL(k_dgemm_stream<4>, g, dim3(256), 0, ctx->stream, P, A, lda, B, ldb, M, N, K, kslice, RA);
  const int64_t tot = M * N;
  hipLaunchKernelGGL(k_splitk_sum<double>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, ctx->stream, C, ldc,
                     P, M, N, (int)ks);
  if (rowabs)
    hipLaunchKernelGGL(k_slice_rowsum, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, ctx->stream, RA, M,
                       (int)ks, rowabs);
  HFPX_TRY_HIP(ctx, hipGetLastError());
  return HFPX_OK;
}

// ---------------------------------------------------------------------
// host dispatch
// ---------------------------------------------------------------------
// FP32 trailing-update kernels selectable through HFPX_KNOB_GEMM_F32
// (0 = 16x16x4 128 x 128 tile; 1.. = 32x32x2 family).
template <int WR, int WC, int KC, int OPT>
static void launch_sgemm32_grid(hfpx_ctx* ctx, float* C, int64_t ldc, const float* A, int64_t lda,
                                const float* B, int64_t ldb, int64_t M, int64_t N, int64_t K) {
  if (M <= 0 || N <= 0) return;
  const int64_t tm = (M + 64 * WR - 1) / (64 * WR), tn = (N + 64 * WC - 1) / (64 * WC);
  const dim3 grid = (OPT & SG_SWZ) ? dim3((unsigned)(tm * tn)) : dim3((unsigned)tm, (unsigned)tn);
  if constexpr (WR * WC > 4)
    hipLaunchKernelGGL((k_sgemm32_w2<WR, WC, KC, OPT>), grid, dim3(256), 0, ctx->stream, C, ldc, A, lda, B, ldb,
                       M, N, K);
  else
    hipLaunchKernelGGL((k_sgemm32<WR, WC, KC, OPT>), grid, dim3(256), 0, ctx->stream, C, ldc, A, lda, B, ldb, M,
                       N, K);
}

// With SG_INT: the whole tiles by the interior kernel, then the bottom band
// (all columns) and the right band (interior rows) by the guarded kernel.
template <int WR, int WC, int KC, int OPT>
static void launch_sgemm32(hfpx_ctx* ctx, float* C, int64_t ldc, const float* A, int64_t lda,
                           const float* B, int64_t ldb, int64_t M, int64_t N, int64_t K) {
  if constexpr ((OPT & SG_INT) == 0) {
    launch_sgemm32_grid<WR, WC, KC, OPT>(ctx, C, ldc, A, lda, B, ldb, M, N, K);
  } else {
    constexpr int EDGE = OPT & ~SG_INT & ~SG_SWZ & ~SG_PIPE;
    const int64_t Mi = M / (64 * WR) * (64 * WR), Ni = N / (64 * WC) * (64 * WC);
    launch_sgemm32_grid<WR, WC, KC, OPT>(ctx, C, ldc, A, lda, B, ldb, Mi, Ni, K);
    launch_sgemm32_grid<WR, WC, KC, EDGE>(ctx, C + Mi, ldc, A + Mi, lda, B, ldb, M - Mi, N, K);
    launch_sgemm32_grid<WR, WC, KC, EDGE>(ctx, C + Ni * ldc, ldc, A, lda, B + Ni * ldb, ldb, Mi, N - Ni, K);
  }
}

int gemm_f32_variants() { return 18; }

// FP64 128 x 128 product: interior tiles by the unguarded kernel, then the
// bottom band (all columns) and the right band (interior rows) guarded.
// variant (HFPX_KNOB_GEMM_F64): 0 = default (3), 2 = plain walk, C read
// and written at the end; 3 = XCD-grouped walk, C preloaded; 4 = 3 with the
// plain walk; 5 = 3 with bands of 4 row tiles.
static int launch_dgemm128(hfpx_ctx* ctx, double* C, int64_t ldc, const double* A, int64_t lda, const double* B,
                           int64_t ldb, int64_t M, int64_t N, int64_t K, int variant) {
  if (variant == 0) variant = 3;
  const int band = variant == 5 ? 4 : SG_GROUP;
  const int64_t Mi = M / DG_T * DG_T, Ni = N / DG_T * DG_T;
  auto grid = [](int64_t m, int64_t n) { return ((m + DG_T - 1) / DG_T) * ((n + DG_T - 1) / DG_T); };
  struct Part {
    double* c;
    const double* a;
    const double* b;
    int64_t m, n;
    bool interior;
  } parts[3] = {{C, A, B, Mi, Ni, true},
                {C + Mi, A + Mi, B, M - Mi, N, false},
                {C + Ni * ldc, A, B + Ni * ldb, Mi, N - Ni, false}};
  for (const Part& p : parts) {
    if (p.m <= 0 || p.n <= 0) continue;
    const int64_t nt = grid(p.m, p.n);
    if (nt > 0x7fffffff) return set_error(ctx, HFPX_ERR_INVALID, "gemm: too many tiles");
    const dim3 g((unsigned)nt);
    if (variant == 2) {
      if (p.interior)
        hipLaunchKernelGGL(k_dgemm128<SG_INT>, g, dim3(256), 0, ctx->stream, p.c, ldc, p.a, lda, p.b, ldb, p.m, p.n,
                           K, band);
      else
        hipLaunchKernelGGL(k_dgemm128<0>, g, dim3(256), 0, ctx->stream, p.c, ldc, p.a, lda, p.b, ldb, p.m, p.n, K,
                           band);
    } else if (variant == 4) {
      if (p.interior)
        hipLaunchKernelGGL(k_dgemm128<SG_INT | SG_CPRE>, g, dim3(256), 0, ctx->stream, p.c, ldc, p.a, lda, p.b, ldb,
                           p.m, p.n, K, band);
      else
        hipLaunchKernelGGL(k_dgemm128<SG_CPRE>, g, dim3(256), 0, ctx->stream, p.c, ldc, p.a, lda, p.b, ldb, p.m, p.n,
                           K, band);
    } else {
      if (p.interior)
        hipLaunchKernelGGL(k_dgemm128<SG_INT | SG_SWZ | SG_CPRE>, g, dim3(256), 0, ctx->stream, p.c, ldc, p.a, lda,
                           p.b, ldb, p.m, p.n, K, band);
      else
        hipLaunchKernelGGL(k_dgemm128<SG_SWZ | SG_CPRE>, g, dim3(256), 0, ctx->stream, p.c, ldc, p.a, lda, p.b, ldb,
                           p.m, p.n, K, band);
    }
  }
  HFPX_TRY_HIP(ctx, hipGetLastError());
  return HFPX_OK;
}

template <typename T>
int gemm_sub(hfpx_ctx* ctx, T* C, int64_t ldc, const T* A, int64_t lda, const T* B,
                    int64_t ldb, int64_t M, int64_t N, int64_t K) {
  if (M <= 0 || N <= 0 || K <= 0) return HFPX_OK;
  const bool skinny = N <= 64;
  const int gtn = skinny ? 64 : 128;
  dim3 grid((unsigned)((M + GT - 1) / GT), (unsigned)((N + gtn - 1) / gtn));
  if ((N + 63) / 64 > 65535) return set_error(ctx, HFPX_ERR_INVALID, "gemm: N too large");
  if (ldc > (int64_t)1 << 24) return set_error(ctx, HFPX_ERR_INVALID, "gemm: ldc above 2^24");
  const int tok = prof_begin(ctx, sizeof(T) == 8 ? HFPX_KC_GEMM_F64 : HFPX_KC_GEMM_F32);
  int var = ctx->knob[HFPX_KNOB_GEMM_F32];
  if (var == 0) {
    // automatic choice: the bf16-split product wherever it has tiles to
    // fill (faster on every LU shape measured, 1.5x on the large trailing
    // updates), with 128 x 256 tiles and the pipelined loop on the large
    // square updates (+5 to 7 % for the tiles, slower on narrow products;
    // +5 % for the pipelining); on the other products the split fused into
    // the product (the split pass costs more than the split redone per
    // tile when one side has few tiles; the same bits; 72,000-DoF mixed
    // solve 1.869 -> 1.852 s), while on the large updates the per-tile
    // split costs more (70,976^2 x 1,024: 59.4 against 51.3 ms); small
    // products on FP32 MFMA, C preloaded for short K
    var = (M >= 256 && N >= 128 && K >= 64) ? ((M >= 4096 && N >= 4096 && K >= 512) ? 15 : 17)
                                             : (K <= 384 ? 7 : 2);
  }
  bool done = false;
  if constexpr (sizeof(T) == 4) {
    if (!skinny && (var >= 13 && var <= 17)) {
      // 13 .. 15: split pass + product (128 x 128, 128 x 256, pipelined
      // 128 x 256); 16, 17: the split fused into the product (128 x 256,
      // 128 x 128)
      const int wide = var == 16 ? 3 : (var == 17 ? 4 : var - 13);
      int rc = gemm_x3(ctx, C, ldc, A, lda, B, ldb, M, N, K, wide);
      if (rc) return rc;
      done = true;
    } else if (!skinny && var > 1) {
      switch (var) {
        case 2: launch_sgemm32<2, 2, 16, 0>(ctx, C, ldc, A, lda, B, ldb, M, N, K); break;
        case 3: launch_sgemm32<2, 2, 16, SG_INT>(ctx, C, ldc, A, lda, B, ldb, M, N, K); break;
        case 4: launch_sgemm32<2, 2, 16, SG_INT | SG_PIPE>(ctx, C, ldc, A, lda, B, ldb, M, N, K); break;
        case 5: launch_sgemm32<2, 2, 16, SG_INT | SG_PIPE | SG_SWZ>(ctx, C, ldc, A, lda, B, ldb, M, N, K); break;
        case 7: launch_sgemm32<2, 2, 16, SG_CPRE>(ctx, C, ldc, A, lda, B, ldb, M, N, K); break;
        case 8: launch_sgemm32<2, 2, 16, SG_INT | SG_SWZ | SG_CPRE>(ctx, C, ldc, A, lda, B, ldb, M, N, K); break;
        case 9: launch_sgemm32<4, 2, 16, SG_INT | SG_SWZ>(ctx, C, ldc, A, lda, B, ldb, M, N, K); break;
        case 10: launch_sgemm32<2, 4, 16, SG_INT | SG_SWZ>(ctx, C, ldc, A, lda, B, ldb, M, N, K); break;
        case 11: launch_sgemm32<4, 2, 16, SG_INT | SG_SWZ | SG_CPRE>(ctx, C, ldc, A, lda, B, ldb, M, N, K); break;
        case 12: launch_sgemm32<4, 2, 16, SG_INT | SG_SWZ | SG_PIPE>(ctx, C, ldc, A, lda, B, ldb, M, N, K); break;
        default: launch_sgemm32<2, 2, 16, SG_INT | SG_SWZ>(ctx, C, ldc, A, lda, B, ldb, M, N, K); break;
      }
      done = true;
    }
  }
  if constexpr (sizeof(T) == 8) {
    if (!done && skinny && K >= SPLITK_MIN_K) {
      int rc = dgemm_stream_sub(ctx, (double*)C, ldc, (const double*)A, lda, (const double*)B, ldb, M, N, K,
                                nullptr);
      if (rc) return rc;
      done = true;
    }
  }
  if (!done && skinny && K >= SPLITK_MIN_K) {
    // few columns, long K (the residuals of the mixed solve): slices of K
    // write partial products -A_z B_z to a workspace, summed in slice order
    const int64_t nrt = (M + GT - 1) / GT;
    int64_t ks = std::min<int64_t>(SPLITK_MAX, std::max<int64_t>(1, (4 * ctx->num_cus + nrt - 1) / nrt));
    const int64_t kchunk = ((K + ks - 1) / ks + KC - 1) / KC * KC;
    ks = (K + kchunk - 1) / kchunk;
    const size_t slice = (size_t)M * N;
    int rc = ensure_gemm_work(ctx, sizeof(T) * slice * ks);
    if (rc) return rc;
    T* P = (T*)ctx->gemm_work;
    HFPX_TRY_HIP(ctx, hipMemsetAsync(P, 0, sizeof(T) * slice * ks, ctx->stream));
    hipLaunchKernelGGL((k_gemm_sub<T, 64>), dim3(grid.x, grid.y, (unsigned)ks), dim3(256), 0, ctx->stream, P, M, A,
                       lda, B, ldb, M, N, K, kchunk, (int64_t)slice);
    const int64_t tot = M * N;
    hipLaunchKernelGGL(k_splitk_sum<T>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, ctx->stream, C, ldc, P,
                       M, N, (int)ks);
    done = true;
  }
  if constexpr (sizeof(T) == 8) {
    // the LU's Schur-complement updates: interior tiles unguarded, the edge
    // bands guarded; HFPX_KNOB_GEMM_F64 selects the walk / epilogue
    const int v64 = ctx->knob[HFPX_KNOB_GEMM_F64];
    if (!done && !skinny && v64 != 1 && M >= 256 && N >= 128) {
      int rc = launch_dgemm128(ctx, (double*)C, ldc, (const double*)A, lda, (const double*)B, ldb, M, N, K, v64);
      if (rc) return rc;
      done = true;
    }
  }
  if (!done) {
    if (skinny)
      hipLaunchKernelGGL((k_gemm_sub<T, 64>), grid, dim3(256), 0, ctx->stream, C, ldc, A, lda, B, ldb, M, N, K, K,
                         (int64_t)0);
    else
      hipLaunchKernelGGL((k_gemm_sub<T, 128>), grid, dim3(256), 0, ctx->stream, C, ldc, A, lda, B, ldb, M, N, K, K,
                         (int64_t)0);
  }
  // algorithmic: 2MNK flops; bytes: C read + written, A and B read once
  prof_end(ctx, tok, sizeof(T) == 8 ? HFPX_KC_GEMM_F64 : HFPX_KC_GEMM_F32, 2.0 * M * N * K,
           (double)sizeof(T) * (2.0 * M * N + (double)M * K + (double)K * N));
  HFPX_TRY_HIP(ctx, hipGetLastError());
  return HFPX_OK;
}


template int gemm_sub<double>(hfpx_ctx*, double*, int64_t, const double*, int64_t, const double*, int64_t,
                              int64_t, int64_t, int64_t);
template int gemm_sub<float>(hfpx_ctx*, float*, int64_t, const float*, int64_t, const float*, int64_t,
                             int64_t, int64_t, int64_t);

int gemm_sub_any(hfpx_ctx* ctx, int dtype, void* C, int64_t ldc, const void* A, int64_t lda,
                 const void* B, int64_t ldb, int64_t M, int64_t N, int64_t K) {
  if (dtype == HFPX_F64)
    return gemm_sub<double>(ctx, (double*)C, ldc, (const double*)A, lda, (const double*)B, ldb, M, N, K);
  return gemm_sub<float>(ctx, (float*)C, ldc, (const float*)A, lda, (const float*)B, ldb, M, N, K);
}

}  // namespace hfpx

namespace hfpx {

// Batched Schur-complement update of step (k, j) of the batched LU over
// the active cracks list[0 .. nact): one workgroup per BGT x BGT tile that
// exists, no empty workgroups.  tstart[e] (e = 0 .. nact) is the prefix sum
// of the tile counts along the list; the workgroups are dealt to the XCDs
// in contiguous ranges of that walk, and a crack's tiles are walked in bands
// of SG_GROUP row tiles.  FP32 tiles run on v_mfma_f32_32x32x2f32 (interior
// tiles unguarded; BG_OPT = SG_CPRE or 0), FP64 tiles on
// v_mfma_f64_16x16x4f64.
template <typename T, int BG_OPT>
__global__ __launch_bounds__(256) void k_bgemm(T* __restrict__ base, int which, const BCrack* __restrict__ cr,
                                               const int32_t* __restrict__ list,
                                               const int32_t* __restrict__ tstart, int nact, int64_t k, int64_t j,
                                               int outer) {
  __shared__ int s_e;
  const int64_t nid = sg_xcd_order(blockIdx.x, tstart[nact]);
  // exactly one entry e has tstart[e] <= nid < tstart[e + 1]
  for (int e = threadIdx.x; e < nact; e += 256)
    if (tstart[e] <= nid && nid < tstart[e + 1]) s_e = e;
  __syncthreads();
  const int e = s_e;
  const BCrack& c = cr[list[e]];
  const int64_t n = c.n;
  int64_t r0, c0, p0, M, N, K;
  bgemm_shape(n, k, j, outer != 0, r0, c0, p0, M, N, K);
  const int64_t tm = (M + BGT - 1) / BGT, tn = (N + BGT - 1) / BGT;
  int64_t tile_r, tile_c;
  sg_band_tile(nid - tstart[e], tm, tn, tile_r, tile_c);
  const int64_t row0 = tile_r * BGT, col0 = tile_c * BGT;
  T* a = base + bcopy_off(c, which);
  T* C = a + r0 + c0 * n;
  const T* A = a + r0 + p0 * n;
  const T* B = a + p0 + c0 * n;
  if constexpr (sizeof(T) == 4) {
    typedef Sg32<2, 2, 16> S;
    static_assert(S::GM == BGT && S::GN == BGT, "batched tile");
    __shared__ float sA[2 * S::LA];
    __shared__ float sB[2 * S::LB];
    if (row0 + BGT <= M && col0 + BGT <= N)
      sgemm32_tile<2, 2, 16, SG_INT | BG_OPT>(sA, sB, C, n, A, n, B, n, M, N, K, row0, col0);
    else
      sgemm32_tile<2, 2, 16, BG_OPT>(sA, sB, C, n, A, n, B, n, M, N, K, row0, col0);
  } else {
    static_assert(GT == BGT, "batched tile");
    gemm_tile<T, BGT>(C, n, A, n, B, n, M, N, K, row0, col0);
  }
}

template <typename T>
int launch_bgemm(hfpx_ctx* ctx, T* base, int which, const BCrack* cr, const int32_t* list, int nact,
                 const int32_t* tstart, int64_t tiles, int64_t k, int64_t j, bool outer) {
  if (nact <= 0 || tiles <= 0) return HFPX_OK;
  if (tiles > 0x7fffffff) return set_error(ctx, HFPX_ERR_INVALID, "batched update: too many tiles");
  const int tok = prof_begin(ctx, sizeof(T) == 8 ? HFPX_KC_GEMM_F64 : HFPX_KC_GEMM_F32);
  if (ctx->knob[HFPX_KNOB_BGEMM_F32] == 1)
    hipLaunchKernelGGL((k_bgemm<T, 0>), dim3((unsigned)tiles), dim3(256), 0, ctx->stream, base, which, cr, list,
                       tstart, nact, k, j, outer ? 1 : 0);
  else
    hipLaunchKernelGGL((k_bgemm<T, SG_CPRE>), dim3((unsigned)tiles), dim3(256), 0, ctx->stream, base, which, cr,
                       list, tstart, nact, k, j, outer ? 1 : 0);
  prof_end(ctx, tok, sizeof(T) == 8 ? HFPX_KC_GEMM_F64 : HFPX_KC_GEMM_F32, 0.0, 0.0);
  HFPX_TRY_HIP(ctx, hipGetLastError());
  return HFPX_OK;
}

template int launch_bgemm<double>(hfpx_ctx*, double*, int, const BCrack*, const int32_t*, int, const int32_t*,
                                  int64_t, int64_t, int64_t, bool);
template int launch_bgemm<float>(hfpx_ctx*, float*, int, const BCrack*, const int32_t*, int, const int32_t*,
                                 int64_t, int64_t, int64_t, bool);

}  // namespace hfpx
