// hfpx_sweep.hip -- FP32 triangular sweeps of the LU solve for many right-hand
// sides (the preconditioner applications of the mixed-precision solve).
//
// Replaces, for the FP32 factors, the forward/back substitution of
// LAPACKE_dgetrs (il/linear_algebra/dense/factorization/LU.h:106-138).
//
// Blocks of SB = 128 rows.  One launch inverts every diagonal block of L
// (unit lower) and of U; each sweep step is then one launch: block c's
// solution rows are final in B, every workgroup subtracts T(tile, c) X_c from
// its 128-row tile (MFMA), and the workgroup whose tile is the next block
// multiplies its freshly updated rows by that block's inverse, which makes
// them final in place.  A sweep is nblk launches whose critical path is one
// tile update plus one 128 x 128 product; the factor is read once per sweep.
#include "hfpx_internal.h"

namespace hfpx {

typedef float floatx16 __attribute__((ext_vector_type(16)));

constexpr int SB = 128;  // rows per block
constexpr int SWN = 64;  // right-hand sides per workgroup

// Inverse of diagonal block blockIdx.x (rows/columns [128 b, 128 b + kb)),
// padded with the identity to 128 x 128, column-major into inv + b SB^2.
// Thread t forms column t by column-oriented substitution in LDS.
template <bool LOWER>
__global__ __launch_bounds__(SB) void k_block_inv(const float* __restrict__ LU, int64_t lda, int64_t n,
                                                  float* __restrict__ inv) {
  constexpr int P = SB + 1;
  __shared__ float sT[SB * P];  // sT[i + j P] = T(i, j)
  __shared__ float sX[SB * P];  // sX[i + t P] = inverse(i, t)
  const int64_t b = blockIdx.x, k = b * SB;
  const int kb = (int)std::min<int64_t>(SB, n - k);
  const int t = threadIdx.x;
  for (int e = t; e < SB * SB; e += SB) {
    const int i = e % SB, j = e / SB;
    float v;
    if (i < kb && j < kb)
      v = LU[(k + i) + (k + j) * lda];
    else
      v = (i == j) ? 1.f : 0.f;
    sT[i + j * P] = v;
    sX[i + j * P] = (i == j) ? 1.f : 0.f;
  }
  __syncthreads();
  float* x = sX + t * P;
  if (LOWER) {
    // unit lower: column t is zero above row t
    for (int j = t; j < SB; ++j) {
      const float xj = x[j];
      for (int i = j + 1; i < SB; ++i) x[i] -= sT[i + j * P] * xj;
    }
  } else {
    // upper: column t is zero below row t
    for (int j = t; j >= 0; --j) {
      const float xj = x[j] / sT[j + j * P];
      x[j] = xj;
      for (int i = 0; i < j; ++i) x[i] -= sT[i + j * P] * xj;
    }
  }
  __syncthreads();
  float* out = inv + b * SB * SB;
  for (int e = t; e < SB * SB; e += SB) out[e] = sX[(e % SB) + (e / SB) * P];
}

// One sweep step (see the file comment).  LOWER: forward with unit L, tiles
// c+1+t; UPPER: backward with U, tiles c-1-t.  c = -1 (LOWER) or nblk
// (UPPER) is the first step: no update, the first block is only solved.
// Eight waves, one 32 x 32 block of the 128 x 64 tile each; every lane
// fetches all of its factor (and inverse) operands before the first MFMA, so
// a step costs one memory round trip plus 64 MFMAs per phase.  Products use
// the transposed MFMA form (v_mfma_f32_32x32x2f32): lane l holds row
// 32 (wave % 4) + l % 32 of the tile and column 8 (v/4) + 4 (l/32) + v%4 of
// the wave's 32-column block, so tile loads and stores are column segments.
constexpr int SWT = 512;  // threads per sweep workgroup

template <bool LOWER>
__global__ __launch_bounds__(SWT) void k_sweep(const float* __restrict__ LU, int64_t lda, int64_t n, int64_t nblk,
                                               const float* __restrict__ inv, float* __restrict__ B, int64_t ldb,
                                               int64_t nrhs, int64_t c) {
  __shared__ float sX[SB][SWN + 1];  // solution rows of block c: sX[k][j]
  __shared__ float sC[SB][SWN + 1];  // updated tile of the solving workgroup: sC[i][j]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, l32 = lane & 31, hk = lane >> 5;
  const int64_t t = blockIdx.x;
  const int64_t tile = LOWER ? c + 1 + t : c - 1 - t;
  const int64_t r0 = tile * SB;
  const int rows = (int)std::min<int64_t>(SB, n - r0);
  const int64_t j0 = (int64_t)blockIdx.y * SWN;
  const int ncol = (int)std::min<int64_t>(SWN, nrhs - j0);
  const bool has_update = LOWER ? (c >= 0) : (c < nblk);
  const int i = 32 * (wave & 3) + l32;  // tile row of this lane
  const int jb = wave >> 2;             // 32-column block of this wave
  const bool rok = i < rows;

  floatx16 acc;
#pragma unroll
  for (int v = 0; v < 16; ++v) acc[v] = 0.f;

  if (has_update) {
    const int64_t k0 = c * SB;
    const int kc = (int)std::min<int64_t>(SB, n - k0);
    const float* lrow = LU + (r0 + i) + k0 * lda;
    float lv[SB / 2];
#pragma unroll
    for (int s = 0; s < SB / 2; ++s) {
      const int k = 2 * s + hk;
      lv[s] = (rok && k < kc) ? lrow[(int64_t)k * lda] : 0.f;
    }
    for (int e = tid; e < SB * SWN; e += SWT) {
      const int k = e % SB, j = e / SB;
      sX[k][j] = (k < kc && j < ncol) ? B[(k0 + k) + (j0 + j) * ldb] : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < SB / 2; ++s)
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(sX[2 * s + hk][32 * jb + l32], lv[s], acc, 0, 0, 0);
  }

  if (t != 0) {
    // plain update of this tile
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const int j = 32 * jb + 8 * (v >> 2) + 4 * hk + (v & 3);
      if (rok && j < ncol) B[(r0 + i) + (j0 + j) * ldb] -= acc[v];
    }
    return;
  }
  // the next block: updated rows -> LDS, then times the block inverse
  const float* iv = inv + tile * SB * SB + i;  // inverse(i, k) at iv[k SB]
  float ivv[SB / 2];
#pragma unroll
  for (int s = 0; s < SB / 2; ++s) ivv[s] = iv[(2 * s + hk) * SB];
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    const int j = 32 * jb + 8 * (v >> 2) + 4 * hk + (v & 3);
    sC[i][j] = (rok && j < ncol) ? B[(r0 + i) + (j0 + j) * ldb] - acc[v] : 0.f;
  }
  __syncthreads();
  floatx16 y;
#pragma unroll
  for (int v = 0; v < 16; ++v) y[v] = 0.f;
#pragma unroll
  for (int s = 0; s < SB / 2; ++s)
    y = __builtin_amdgcn_mfma_f32_32x32x2f32(sC[2 * s + hk][32 * jb + l32], ivv[s], y, 0, 0, 0);
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    const int j = 32 * jb + 8 * (v >> 2) + 4 * hk + (v & 3);
    if (rok && j < ncol) B[(r0 + i) + (j0 + j) * ldb] = y[v];
  }
}

size_t sweep_f32_work_bytes(int64_t n) {
  const int64_t nblk = (n + SB - 1) / SB;
  return sizeof(float) * 2 * (size_t)nblk * SB * SB;
}

// B := U^{-1} L^{-1} B for the FP32 factors (row interchanges already
// applied to B); work holds sweep_f32_work_bytes(n).
int sweep_f32(hfpx_ctx* ctx, const float* LU, int64_t lda, int64_t n, float* B, int64_t ldb, int64_t nrhs,
              void* work) {
  if (n <= 0 || nrhs <= 0) return HFPX_OK;
  const int64_t nblk = (n + SB - 1) / SB;
  const unsigned gy = (unsigned)((nrhs + SWN - 1) / SWN);
  float* invL = (float*)work;
  float* invU = invL + (size_t)nblk * SB * SB;
  int tok = prof_begin(ctx, HFPX_KC_TRSM);
  hipLaunchKernelGGL(k_block_inv<true>, dim3((unsigned)nblk), dim3(SB), 0, ctx->stream, LU, lda, n, invL);
  hipLaunchKernelGGL(k_block_inv<false>, dim3((unsigned)nblk), dim3(SB), 0, ctx->stream, LU, lda, n, invU);
  prof_end(ctx, tok, HFPX_KC_TRSM, 2.0 * nblk * SB * SB * SB / 3.0, 0.0);
  HFPX_TRY_HIP(ctx, hipGetLastError());
  tok = prof_begin(ctx, HFPX_KC_TRSM);
  for (int64_t c = -1; c <= nblk - 2; ++c) {
    const unsigned gx = (unsigned)(c < 0 ? 1 : nblk - 1 - c);
    hipLaunchKernelGGL(k_sweep<true>, dim3(gx, gy), dim3(SWT), 0, ctx->stream, LU, lda, n, nblk, invL, B, ldb,
                       nrhs, c);
  }
  for (int64_t c = nblk; c >= 1; --c) {
    const unsigned gx = (unsigned)(c == nblk ? 1 : c);
    hipLaunchKernelGGL(k_sweep<false>, dim3(gx, gy), dim3(SWT), 0, ctx->stream, LU, lda, n, nblk, invU, B, ldb,
                       nrhs, c);
  }
  // algorithmic: 2 n^2 nrhs flops, the factor read once (4 n^2 bytes)
  prof_end(ctx, tok, HFPX_KC_TRSM, 2.0 * n * n * nrhs, 4.0 * n * n);
  HFPX_TRY_HIP(ctx, hipGetLastError());
  return HFPX_OK;
}

}  // namespace hfpx
