// hfpx_gemm_x3.hip -- FP32 C -= A B on the bf16 MFMA units of gfx950 by
// three-way operand splitting, for the large trailing updates of the FP32
// LU.
//
// Every FP32 operand x is written exactly as x = h + m + l + r with h, m, l
// bf16 (h = bf16(x), m = bf16(x - h), l = bf16(x - h - m); each difference
// is exact in FP32) and |r| <= 2^-24 |x| or so.  With |m| ~ 2^-9 |x| and
// |l| ~ 2^-18 |x|, the product a b is h_a h_b + h_a m_b + m_a h_b + h_a l_b +
// l_a h_b + m_a m_b up to terms of relative size 2^-27, below FP32 rounding:
// six bf16 products (each exact, accumulated in FP32 by the MFMA) give an
// FP32-accurate product at up to 16/6 the rate of v_mfma_f32_32x32x2f32.
//
// Two passes: the split kernels write the parts of A (M x K) and B (K x N)
// in blocks of 128 rows (columns of B) by 16 k, each block [part][row][k]
// contiguous (12 KB) and zero-padded, B negated; the product kernel preloads
// the tile of C into its accumulators and adds the six products of the
// negated B parts and the A parts.  No direct reference counterpart: inside
// LAPACKE_dgetrf (il/linear_algebra/dense/factorization/LU.h:79-138) this is
// the dgemm of the trailing update, done here in FP32 for the mixed solve.
#include <algorithm>

#include "hfpx_internal.h"

namespace hfpx {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16_x3 __attribute__((ext_vector_type(16)));

constexpr int X3_T = 128;                  // rows (A) / columns (B) of a block
constexpr int X3_KC = 16;                  // k of a block (one MFMA k-step)
constexpr int X3_PIECES = 3 * X3_T * 2;    // 16-byte pieces of a block (768)
constexpr int X3_BLK = 3 * X3_T * X3_KC;   // bf16 per block
constexpr int X3_ROW = 24;                 // LDS pitch of a 16-k row (bf16): 48 B keeps
                                           // the 16-byte reads of 8 lanes on distinct banks

__device__ __forceinline__ void split3(float x, __bf16& h, __bf16& m, __bf16& l) {
  h = (__bf16)x;
  const float r1 = x - (float)h;
  m = (__bf16)r1;
  const float r2 = r1 - (float)m;
  l = (__bf16)r2;
}

// Row tile it, chunk kc of A (M x K, lda) -> one block [part][row][k] at
// blk, sign * A.  Thread t writes piece t of each part: row t / 2, k 8 (t % 2)
// .. + 7.  s is the caller's LDS staging.
__device__ __forceinline__ void split_a_block(const float* __restrict__ A, int64_t lda, int64_t M, int64_t K,
                                              int64_t it, int64_t kc, bf16x8* __restrict__ blk, float sign,
                                              float (*s)[X3_T + 1]) {
  const int t = threadIdx.x;
#pragma unroll
  for (int q = 0; q < X3_T * X3_KC / 256; ++q) {
    const int e = t + 256 * q, i = e % X3_T, k = e / X3_T;
    const int64_t gi = it * X3_T + i, gk = kc * X3_KC + k;
    s[k][i] = (gi < M && gk < K) ? A[gi + gk * lda] : 0.f;
  }
  __syncthreads();
  const int il = t >> 1, half = t & 1;
  bf16x8 ph, pm, pl;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    __bf16 h, m, l;
    split3(sign * s[8 * half + e][il], h, m, l);
    ph[e] = h;
    pm[e] = m;
    pl[e] = l;
  }
  blk[t] = ph;
  blk[256 + t] = pm;
  blk[512 + t] = pl;
}

// Column tile jt, chunk kc of B (K x N, ldb) -> one block [part][column][k].
__device__ __forceinline__ void split_b_block(const float* __restrict__ B, int64_t ldb, int64_t K, int64_t N,
                                              int64_t jt, int64_t kc, bf16x8* __restrict__ blk, float sign) {
  const int t = threadIdx.x;
  const int jl = t >> 1, half = t & 1;
  const int64_t gj = jt * X3_T + jl, k0 = kc * X3_KC + 8 * half;
  bf16x8 ph, pm, pl;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const float x = (gj < N && k0 + e < K) ? B[(k0 + e) + gj * ldb] : 0.f;
    __bf16 h, m, l;
    split3(sign * x, h, m, l);
    ph[e] = h;
    pm[e] = m;
    pl[e] = l;
  }
  blk[t] = ph;
  blk[256 + t] = pm;
  blk[512 + t] = pl;
}

// A (M x K, lda) -> blocks [row tile it][chunk kc], sign * A.
__global__ __launch_bounds__(256) void k_split_a(const float* __restrict__ A, int64_t lda, int64_t M, int64_t K,
                                                 int nkc, __bf16* __restrict__ out, float sign) {
  __shared__ float s[X3_KC][X3_T + 1];
  const int64_t it = blockIdx.x, kc = blockIdx.y;
  split_a_block(A, lda, M, K, it, kc, (bf16x8*)(out + ((size_t)it * nkc + kc) * X3_BLK), sign, s);
}

// B (K x N, ldb) -> blocks [column tile jt][chunk kc], sign * B.
__global__ __launch_bounds__(256) void k_split_b(const float* __restrict__ B, int64_t ldb, int64_t K, int64_t N,
                                                 int nkc, __bf16* __restrict__ out, float sign) {
  const int64_t jt = blockIdx.x, kc = blockIdx.y;
  split_b_block(B, ldb, K, N, jt, kc, (bf16x8*)(out + ((size_t)jt * nkc + kc) * X3_BLK), sign);
}

// C (M x N) += As Bs over nkc chunks: tiles of 128 rows by 64 WCB columns,
// 4 waves of 64 x 32 WCB (2 x WCB MFMA blocks of 32 x 32), one 16-k chunk
// per LDS buffer (double buffered, one barrier per chunk; loading two chunks
// ahead measured slower).  WCB = 2 reads 0.5 LDS operands per MFMA, WCB = 4
// 0.375.  ROW is the LDS pitch of a 16-k row (bf16): 24 (48 B) or 16
// (unpadded, less LDS per workgroup; the two 16-byte halves of row r are
// then swapped when bit 2 xor bit 3 of r is set, so that the 16-byte
// operand reads of 16 consecutive lanes, rows r .. r + 15, cover the 64
// banks once, and those of 8 lanes 32 banks, where the plain 32-byte pitch
// puts rows r and r + 8 on the same banks: measured, 2.8e9 conflict cycles
// in a 70,976 x 70,976 x 1,024 product).  As the FP32 kernels, the MFMA forms
// the transposed product: operand a is a block of Bs^T (C columns), operand
// b a block of As^T (C rows), so lane l % 32 runs down 32 consecutive rows
// of C.  v_mfma_f32_32x32x16_bf16: lane l supplies a[l%32][8(l/32) + e] and
// b[8(l/32) + e][l%32], e = 0..7; result element v of lane l is
// D[8(v/4) + 4(l/32) + v%4][l%32].  Tiles are walked as in the FP32 kernel
// (each XCD a contiguous range, in bands of row tiles), the band height
// chosen by the host so that the band's A parts fit an XCD's L2.  Bounded
// to two waves per SIMD, the compiler keeps accumulators and operands in
// VGPRs.  Bs has an even number of column blocks when WCB = 4 (the split
// kernel zero-fills the last one).
template <int ROW, int WCB>
struct X3Lds {
  static constexpr int NBB = 64 * WCB / X3_T;  // B blocks per tile (1 or 2)
  __bf16 a[2][3 * X3_T * ROW];
  __bf16 b[2][NBB][3 * X3_T * ROW];
};

// One tile of C at (row0, col0): pa = the row tile's A blocks (nkc of them),
// pb = the column tile's B blocks (NBB x nkc, consecutive), lds = the
// caller's LDS.
// PIPE: the software-pipelined loop (one scheduling region per chunk, the
// last chunk loaded and staged again into the idle buffer instead of a
// branch): the B operands of column block tc + 1 are read from LDS, into a
// second register set, while the twelve products of block tc issue, and
// the next chunk's LDS stores go between the last block's products, so
// that a wave waits for LDS at most once per chunk (after the barrier)
// where the plain loop waits before every column block.
template <int ROW, int WCB, bool PIPE = false>
__device__ __forceinline__ void x3_tile(X3Lds<ROW, WCB>& lds, float* __restrict__ C, int64_t ldc,
                                        const __bf16* __restrict__ pa, const __bf16* __restrict__ pb, int64_t M,
                                        int64_t N, int nkc, int64_t row0, int64_t col0) {
  constexpr int GN = 64 * WCB;                 // tile columns
  constexpr int NBB = X3Lds<ROW, WCB>::NBB;
  auto& sA = lds.a;
  auto& sB = lds.b;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int l32 = lane & 31, hk = lane >> 5;
  const int wr = (wave & 1) * 64, wc = (wave >> 1) * 32 * WCB;

  // accumulators start as the tile of C (32-bit offsets from its corner)
  float* __restrict__ Ct = C + row0 + col0 * ldc;
  const int ld = (int)ldc;
  const int off0 = wr + l32 + (wc + 4 * hk) * ld;
  const int mrem = (int)((M - row0) < X3_T ? (M - row0) : X3_T);
  const int nrem = (int)((N - col0) < GN ? (N - col0) : GN);
  floatx16_x3 acc[WCB][2];
#pragma unroll
  for (int tc = 0; tc < WCB; ++tc)
#pragma unroll
    for (int tr = 0; tr < 2; ++tr) {
      const int i = wr + 32 * tr + l32;
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int jc = wc + 32 * tc + 8 * (v >> 2) + 4 * hk + (v & 3);
        const int off = off0 + 32 * tr + (32 * tc + 8 * (v >> 2) + (v & 3)) * ld;
        acc[tc][tr][v] = (i < mrem && jc < nrem) ? Ct[off] : 0.f;
      }
    }

  // staging: piece t + 256 q of a block is part q, row t / 2, k 8 (t % 2) .. + 7
  typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
  const u32x4* pa4 = (const u32x4*)pa;
  const u32x4* pb4 = (const u32x4*)pb;
  u32x4 ra[3], rb[NBB][3];
  auto gload = [&](int kc) {
#pragma unroll
    for (int q = 0; q < 3; ++q) ra[q] = pa4[(size_t)kc * X3_PIECES + t + 256 * q];
#pragma unroll
    for (int h = 0; h < NBB; ++h)
#pragma unroll
      for (int q = 0; q < 3; ++q) rb[h][q] = pb4[((size_t)h * nkc + kc) * X3_PIECES + t + 256 * q];
  };
  // half h of row r sits at 8 (h ^ sw(r)), sw(r) = bit 2 ^ bit 3 of r; the
  // operand rows of a lane are l32 plus multiples of 32
  constexpr int SWZ = ROW == 16 ? 1 : 0;
  const int soff = (t >> 1) * ROW + 8 * ((t & 1) ^ (SWZ & ((t >> 3) ^ (t >> 4))));
  const int rh = 8 * (hk ^ (SWZ & ((l32 >> 2) ^ (l32 >> 3))));  // this lane's operand half
  // piece q of the staged chunk: A parts q < 3, then B block h part q % 3
  auto lstore_piece = [&](int buf, int q) {
    if (q < 3)
      *(u32x4*)&sA[buf][q * X3_T * ROW + soff] = ra[q];
    else
      *(u32x4*)&sB[buf][(q - 3) / 3][((q - 3) % 3) * X3_T * ROW + soff] = rb[(q - 3) / 3][(q - 3) % 3];
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int q = 0; q < 3; ++q) *(u32x4*)&sA[buf][q * X3_T * ROW + soff] = ra[q];
#pragma unroll
    for (int h = 0; h < NBB; ++h)
#pragma unroll
      for (int q = 0; q < 3; ++q) *(u32x4*)&sB[buf][h][q * X3_T * ROW + soff] = rb[h][q];
  };

  gload(0);
  lstore(0);
  __syncthreads();
  if constexpr (PIPE) {
    // operand (B part, A part) of product q of a pair, small terms first
    constexpr int PB[6] = {1, 2, 0, 1, 0, 0}, PA[6] = {1, 0, 2, 0, 1, 0};
    constexpr int NLD = 3 * (1 + X3Lds<ROW, WCB>::NBB);  // staged pieces per chunk
    for (int kc = 0; kc < nkc; ++kc) {
      const int buf = kc & 1;
      gload(kc + 1 < nkc ? kc + 1 : kc);
      __builtin_amdgcn_sched_barrier(0);
      bf16x8 fb[2][3], fa[2][3];
#pragma unroll
      for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int u = 0; u < 2; ++u)
          fb[u][p] = *(const bf16x8*)&sA[buf][(p * X3_T + wr + 32 * u + l32) * ROW + rh];
      auto read_fa = [&](int tc, int p, bf16x8& f) {
        const int jb = wc + 32 * tc, h = jb / X3_T, jl = jb % X3_T;
        f = *(const bf16x8*)&sB[buf][h][(p * X3_T + jl + l32) * ROW + rh];
      };
#pragma unroll
      for (int p = 0; p < 3; ++p) read_fa(0, p, fa[0][p]);
      __builtin_amdgcn_sched_barrier(0);
      // program order is issue order: product i of block tc (the two row
      // blocks alternate), then block tc + 1's operand read i (i < 3), or
      // in the last block the LDS store i - 3 of the next chunk
#pragma unroll
      for (int tc = 0; tc < WCB; ++tc) {
#pragma unroll
        for (int i = 0; i < 12; ++i) {
          const int tr = i & 1, q = i >> 1;
          acc[tc][tr] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[tc & 1][PB[q]], fb[tr][PA[q]], acc[tc][tr], 0,
                                                                0, 0);
          __builtin_amdgcn_sched_barrier(0);
          if (tc + 1 < WCB && i < 3) {
            read_fa(tc + 1, i, fa[(tc + 1) & 1][i]);
            __builtin_amdgcn_sched_barrier(0);
          }
          if (tc + 1 == WCB && i >= 12 - NLD) {
            lstore_piece(buf ^ 1, i - (12 - NLD));
            __builtin_amdgcn_sched_barrier(0);
          }
        }
      }
      __syncthreads();
    }
  } else
  for (int kc = 0; kc < nkc; ++kc) {
    const int buf = kc & 1;
    if (kc + 1 < nkc) gload(kc + 1);
    bf16x8 fb[2][3];
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
      for (int u = 0; u < 2; ++u)
        fb[u][p] = *(const bf16x8*)&sA[buf][(p * X3_T + wr + 32 * u + l32) * ROW + rh];
#pragma unroll
    for (int tc = 0; tc < WCB; ++tc) {
      // the B operands of one 32-column block at a time (register pressure)
      const int jb = wc + 32 * tc, h = jb / X3_T, jl = jb % X3_T;
      bf16x8 fa[3];
#pragma unroll
      for (int p = 0; p < 3; ++p) fa[p] = *(const bf16x8*)&sB[buf][h][(p * X3_T + jl + l32) * ROW + rh];
#pragma unroll
      for (int tr = 0; tr < 2; ++tr) {
        floatx16_x3 c = acc[tc][tr];
        // small terms first
        c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[1], fb[tr][1], c, 0, 0, 0);  // m m
        c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[2], fb[tr][0], c, 0, 0, 0);  // l h
        c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[0], fb[tr][2], c, 0, 0, 0);  // h l
        c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[1], fb[tr][0], c, 0, 0, 0);  // m h
        c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[0], fb[tr][1], c, 0, 0, 0);  // h m
        c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[0], fb[tr][0], c, 0, 0, 0);  // h h
        acc[tc][tr] = c;
      }
    }
    if (kc + 1 < nkc) lstore(buf ^ 1);
    __syncthreads();
  }
#pragma unroll
  for (int tc = 0; tc < WCB; ++tc)
#pragma unroll
    for (int tr = 0; tr < 2; ++tr) {
      const int i = wr + 32 * tr + l32;
      if (i >= mrem) continue;
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int jc = wc + 32 * tc + 8 * (v >> 2) + 4 * hk + (v & 3);
        const int off = off0 + 32 * tr + (32 * tc + 8 * (v >> 2) + (v & 3)) * ld;
        if (jc < nrem) Ct[off] = acc[tc][tr][v];
      }
    }
}

// The same tile with the split fused into the load stage: the kernel reads
// the FP32 operands themselves (4 bytes per entry instead of the parts'
// 6, and no split pass), splits each staged piece in registers into its
// three bf16 parts and stores them to LDS in the layout of the split blocks.
// The parts, and so the products, are bitwise those of the split kernels.
// Thread t stages row row0 + t / 2 of A and column col0 + 128 h + t / 2 of
// B (h < NBB), k 8 (t % 2) .. + 7 of each chunk.  Software-pipelined as
// PIPE above; each staged piece is split at the LDS store slot of its
// first part (one piece's parts live at a time).  VECB: B's columns are
// 16-byte aligned (two 16-byte loads per piece instead of eight 4-byte
// ones).
template <int ROW, int WCB, bool VECB>
__device__ __forceinline__ void x3f_tile(X3Lds<ROW, WCB>& lds, float* __restrict__ C, int64_t ldc,
                                         const float* __restrict__ A, int64_t lda, const float* __restrict__ B,
                                         int64_t ldb, int64_t M, int64_t N, int64_t K, int nkc, int64_t row0,
                                         int64_t col0) {
  constexpr int GN = 64 * WCB;
  constexpr int NBB = X3Lds<ROW, WCB>::NBB;
  auto& sA = lds.a;
  auto& sB = lds.b;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int l32 = lane & 31, hk = lane >> 5;
  const int wr = (wave & 1) * 64, wc = (wave >> 1) * 32 * WCB;

  float* __restrict__ Ct = C + row0 + col0 * ldc;
  const int ld = (int)ldc;
  const int off0 = wr + l32 + (wc + 4 * hk) * ld;
  const int mrem = (int)((M - row0) < X3_T ? (M - row0) : X3_T);
  const int nrem = (int)((N - col0) < GN ? (N - col0) : GN);
  floatx16_x3 acc[WCB][2];
#pragma unroll
  for (int tc = 0; tc < WCB; ++tc)
#pragma unroll
    for (int tr = 0; tr < 2; ++tr) {
      const int i = wr + 32 * tr + l32;
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int jc = wc + 32 * tc + 8 * (v >> 2) + 4 * hk + (v & 3);
        const int off = off0 + 32 * tr + (32 * tc + 8 * (v >> 2) + (v & 3)) * ld;
        acc[tc][tr][v] = (i < mrem && jc < nrem) ? Ct[off] : 0.f;
      }
    }

  // raw FP32 staging of this thread's pieces
  const int pr = t >> 1, pk = 8 * (t & 1);
  const int64_t ga = row0 + pr;
  const bool a_in = ga < M;
  float ra[8], rb[NBB][8];
  auto gload = [&](int kc) {
    const int64_t k0 = (int64_t)kc * X3_KC + pk;
    const bool kfull = k0 + 8 <= K;
    const float* p = A + ga + k0 * lda;
    if (a_in && kfull) {
#pragma unroll
      for (int e = 0; e < 8; ++e) ra[e] = p[e * lda];
    } else {
#pragma unroll
      for (int e = 0; e < 8; ++e) ra[e] = (a_in && k0 + e < K) ? p[e * lda] : 0.f;
    }
#pragma unroll
    for (int h = 0; h < NBB; ++h) {
      const int64_t gj = col0 + X3_T * h + pr;
      const float* q = B + k0 + gj * ldb;
      if (gj < N && kfull) {
        if constexpr (VECB) {
          const float4 x0 = *(const float4*)q, x1 = *(const float4*)(q + 4);
          rb[h][0] = x0.x; rb[h][1] = x0.y; rb[h][2] = x0.z; rb[h][3] = x0.w;
          rb[h][4] = x1.x; rb[h][5] = x1.y; rb[h][6] = x1.z; rb[h][7] = x1.w;
        } else {
#pragma unroll
          for (int e = 0; e < 8; ++e) rb[h][e] = q[e];
        }
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) rb[h][e] = (gj < N && k0 + e < K) ? q[e] : 0.f;
      }
    }
  };
  // the three parts of one staged piece at a time (B negated): piece g = 0
  // is A's, g = 1 + h block h of B's
  bf16x8 pp[3];
  auto split_piece = [&](int g) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      __bf16 h, m, l;
      if (g == 0)
        split3(ra[e], h, m, l);
      else
        split3(-rb[g - 1][e], h, m, l);
      pp[0][e] = h;
      pp[1][e] = m;
      pp[2][e] = l;
    }
  };
  constexpr int SWZ = ROW == 16 ? 1 : 0;
  const int soff = (t >> 1) * ROW + 8 * ((t & 1) ^ (SWZ & ((t >> 3) ^ (t >> 4))));
  const int rh = 8 * (hk ^ (SWZ & ((l32 >> 2) ^ (l32 >> 3))));
  // store q of a chunk: part q % 3 of piece q / 3 (split at its first part)
  auto lstore_piece = [&](int buf, int q) {
    if (q % 3 == 0) split_piece(q / 3);
    if (q < 3)
      *(bf16x8*)&sA[buf][q * X3_T * ROW + soff] = pp[q];
    else
      *(bf16x8*)&sB[buf][(q - 3) / 3][((q - 3) % 3) * X3_T * ROW + soff] = pp[(q - 3) % 3];
  };
  constexpr int NLD = 3 * (1 + NBB);  // staged pieces per chunk

  gload(0);
#pragma unroll
  for (int q = 0; q < NLD; ++q) lstore_piece(0, q);
  __syncthreads();
  constexpr int PB[6] = {1, 2, 0, 1, 0, 0}, PA[6] = {1, 0, 2, 0, 1, 0};
  for (int kc = 0; kc < nkc; ++kc) {
    const int buf = kc & 1;
    gload(kc + 1 < nkc ? kc + 1 : kc);
    __builtin_amdgcn_sched_barrier(0);
    bf16x8 fb[2][3], fa[2][3];
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
      for (int u = 0; u < 2; ++u)
        fb[u][p] = *(const bf16x8*)&sA[buf][(p * X3_T + wr + 32 * u + l32) * ROW + rh];
    auto read_fa = [&](int tc, int p, bf16x8& f) {
      const int jb = wc + 32 * tc, h = jb / X3_T, jl = jb % X3_T;
      f = *(const bf16x8*)&sB[buf][h][(p * X3_T + jl + l32) * ROW + rh];
    };
#pragma unroll
    for (int p = 0; p < 3; ++p) read_fa(0, p, fa[0][p]);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int tc = 0; tc < WCB; ++tc) {
#pragma unroll
      for (int i = 0; i < 12; ++i) {
        const int tr = i & 1, q = i >> 1;
        acc[tc][tr] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[tc & 1][PB[q]], fb[tr][PA[q]], acc[tc][tr], 0,
                                                              0, 0);
        __builtin_amdgcn_sched_barrier(0);
        if (tc + 1 < WCB && i < 3) {
          read_fa(tc + 1, i, fa[(tc + 1) & 1][i]);
          __builtin_amdgcn_sched_barrier(0);
        }
        if (tc + 1 == WCB && i >= 12 - NLD) {
          lstore_piece(buf ^ 1, i - (12 - NLD));
          __builtin_amdgcn_sched_barrier(0);
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int tc = 0; tc < WCB; ++tc)
#pragma unroll
    for (int tr = 0; tr < 2; ++tr) {
      const int i = wr + 32 * tr + l32;
      if (i >= mrem) continue;
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int jc = wc + 32 * tc + 8 * (v >> 2) + 4 * hk + (v & 3);
        const int off = off0 + 32 * tr + (32 * tc + 8 * (v >> 2) + (v & 3)) * ld;
        if (jc < nrem) Ct[off] = acc[tc][tr][v];
      }
    }
}

template <int ROW, int WCB, bool VECB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_x3gemm_f(
    float* __restrict__ C, int64_t ldc, const float* __restrict__ A, int64_t lda, const float* __restrict__ B,
    int64_t ldb, int64_t M, int64_t N, int64_t K, int nkc, int64_t tm, int64_t tn, int64_t band) {
  __shared__ __attribute__((aligned(16))) X3Lds<ROW, WCB> lds;
  int64_t tile_r, tile_c;
  sg_band_tile(sg_xcd_order(blockIdx.x, tm * tn), tm, tn, tile_r, tile_c, band);
  x3f_tile<ROW, WCB, VECB>(lds, C, ldc, A, lda, B, ldb, M, N, K, nkc, tile_r * X3_T, tile_c * 64 * WCB);
}

template <int ROW, int WCB, bool PIPE = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_x3gemm(
    float* __restrict__ C, int64_t ldc, const __bf16* __restrict__ As, const __bf16* __restrict__ Bs, int64_t M,
    int64_t N, int nkc, int64_t tm, int64_t tn, int64_t band) {
  __shared__ __attribute__((aligned(16))) X3Lds<ROW, WCB> lds;
  int64_t tile_r, tile_c;
  sg_band_tile(sg_xcd_order(blockIdx.x, tm * tn), tm, tn, tile_r, tile_c, band);
  constexpr int NBB = X3Lds<ROW, WCB>::NBB;
  x3_tile<ROW, WCB, PIPE>(lds, C, ldc, As + (size_t)tile_r * nkc * X3_BLK, Bs + (size_t)tile_c * NBB * nkc * X3_BLK, M,
                    N, nkc, tile_r * X3_T, tile_c * 64 * WCB);
}

// ---------------------------------------------------------------------
// Batched form: the outer Schur-complement updates of the FP32 group of the
// batched LU (bgemm_shape with outer = true, K = BNB).  For one launch,
// rstart / cstart (nact + 1 entries) are the prefix sums of the active
// cracks' 128-row tiles of A and 128-column tiles of B along the list; the
// split buffer holds the A blocks of all of them, then the B blocks.
// ---------------------------------------------------------------------
// entry e of the list with start[e] <= id < start[e + 1] (one exists)
__device__ __forceinline__ int find_entry(const int32_t* __restrict__ start, int nact, int64_t id, int* s_e) {
  for (int e = threadIdx.x; e < nact; e += 256)
    if (start[e] <= id && id < start[e + 1]) *s_e = e;
  __syncthreads();
  return *s_e;
}

__global__ __launch_bounds__(256) void k_bsplit_a(float* __restrict__ base, int which, const BCrack* __restrict__ cr,
                                                  const int32_t* __restrict__ list, const int32_t* __restrict__ rstart,
                                                  int nact, int64_t k, int nkc, __bf16* __restrict__ out) {
  __shared__ float s[X3_KC][X3_T + 1];
  __shared__ int s_e;
  const int64_t rb = blockIdx.x, kc = blockIdx.y;
  const int e = find_entry(rstart, nact, rb, &s_e);
  const BCrack& c = cr[list[e]];
  int64_t r0, c0, p0, M, N, K;
  bgemm_shape(c.n, k, k, true, r0, c0, p0, M, N, K);
  const float* a = base + bcopy_off(c, which);
  split_a_block(a + r0 + p0 * c.n, c.n, M, K, rb - rstart[e], kc, (bf16x8*)(out + ((size_t)rb * nkc + kc) * X3_BLK),
                1.0f, s);
}

__global__ __launch_bounds__(256) void k_bsplit_b(float* __restrict__ base, int which, const BCrack* __restrict__ cr,
                                                  const int32_t* __restrict__ list, const int32_t* __restrict__ cstart,
                                                  int nact, int64_t k, int nkc, __bf16* __restrict__ out) {
  __shared__ int s_e;
  const int64_t cb = blockIdx.x, kc = blockIdx.y;
  const int e = find_entry(cstart, nact, cb, &s_e);
  const BCrack& c = cr[list[e]];
  int64_t r0, c0, p0, M, N, K;
  bgemm_shape(c.n, k, k, true, r0, c0, p0, M, N, K);
  const float* a = base + bcopy_off(c, which);
  split_b_block(a + p0 + c0 * c.n, c.n, K, N, cb - cstart[e], kc, (bf16x8*)(out + ((size_t)cb * nkc + kc) * X3_BLK),
                -1.0f);
}

// one workgroup per 128 x 128 tile of the launch (tstart: tile prefix sums)
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_bx3gemm(
    float* __restrict__ base, int which, const BCrack* __restrict__ cr, const int32_t* __restrict__ list,
    const int32_t* __restrict__ tstart, const int32_t* __restrict__ rstart, const int32_t* __restrict__ cstart,
    int nact, int64_t k, int nkc, const __bf16* __restrict__ As, const __bf16* __restrict__ Bs) {
  __shared__ __attribute__((aligned(16))) X3Lds<24, 2> lds;
  __shared__ int s_e;
  const int64_t nid = sg_xcd_order(blockIdx.x, tstart[nact]);
  const int e = find_entry(tstart, nact, nid, &s_e);
  const BCrack& c = cr[list[e]];
  int64_t r0, c0, p0, M, N, K;
  bgemm_shape(c.n, k, k, true, r0, c0, p0, M, N, K);
  const int64_t tm = (M + X3_T - 1) / X3_T, tn = (N + X3_T - 1) / X3_T;
  int64_t tile_r, tile_c;
  sg_band_tile(nid - tstart[e], tm, tn, tile_r, tile_c);
  float* a = base + bcopy_off(c, which);
  x3_tile<24, 2>(lds, a + r0 + c0 * c.n, c.n, As + (size_t)(rstart[e] + tile_r) * nkc * X3_BLK,
                 Bs + (size_t)(cstart[e] + tile_c) * nkc * X3_BLK, M, N, nkc, tile_r * X3_T, tile_c * X3_T);
}

size_t bgemm_x3_split_bytes(int64_t rtot, int64_t ctot) {
  return sizeof(__bf16) * (size_t)(rtot + ctot) * (BNB / X3_KC) * X3_BLK;
}

int launch_bgemm_x3(hfpx_ctx* ctx, float* base, int which, const BCrack* cr, const int32_t* list, int nact,
                    const int32_t* tstart, int64_t tiles, const int32_t* rstart, int64_t rtot, const int32_t* cstart,
                    int64_t ctot, int64_t k, __bf16* split) {
  if (nact <= 0 || tiles <= 0) return HFPX_OK;
  constexpr int nkc = BNB / X3_KC;
  __bf16* As = split;
  __bf16* Bs = split + (size_t)rtot * nkc * X3_BLK;
  const int tok = prof_begin(ctx, HFPX_KC_GEMM_F32);
  hipLaunchKernelGGL(k_bsplit_a, dim3((unsigned)rtot, (unsigned)nkc), dim3(256), 0, ctx->stream, base, which, cr,
                     list, rstart, nact, k, nkc, As);
  hipLaunchKernelGGL(k_bsplit_b, dim3((unsigned)ctot, (unsigned)nkc), dim3(256), 0, ctx->stream, base, which, cr,
                     list, cstart, nact, k, nkc, Bs);
  hipLaunchKernelGGL(k_bx3gemm, dim3((unsigned)tiles), dim3(256), 0, ctx->stream, base, which, cr, list, tstart,
                     rstart, cstart, nact, k, nkc, (const __bf16*)As, (const __bf16*)Bs);
  prof_end(ctx, tok, HFPX_KC_GEMM_F32, 0.0, 0.0);
  HFPX_TRY_HIP(ctx, hipGetLastError());
  return HFPX_OK;
}

// Grow-only workspace of the split operands.
int ensure_x3_work(hfpx_ctx* ctx, size_t bytes) {
  if (ctx->x3_work_bytes >= bytes) return HFPX_OK;
  if (ctx->x3_work) {
    HFPX_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    HFPX_TRY_HIP(ctx, hipFree(ctx->x3_work));
    ctx->x3_work = nullptr;
    ctx->x3_work_bytes = 0;
  }
  HFPX_TRY_HIP(ctx, hipMalloc(&ctx->x3_work, bytes));
  ctx->x3_work_bytes = bytes;
  return HFPX_OK;
}

size_t gemm_x3_work_bytes(int64_t M, int64_t N, int64_t K) {
  // B blocks rounded up to an even count (the 256-column tiles)
  const int64_t tm = (M + X3_T - 1) / X3_T, tnb = ((N + 2 * X3_T - 1) / (2 * X3_T)) * 2,
                nkc = (K + X3_KC - 1) / X3_KC;
  return sizeof(__bf16) * (size_t)(tm + tnb) * nkc * X3_BLK;
}

// The fused-split products (wide 3: 128 x 256 tiles, 4: 128 x 128): no
// split pass and no workspace.
static int gemm_x3_fused(hfpx_ctx* ctx, float* C, int64_t ldc, const float* A, int64_t lda, const float* B,
                         int64_t ldb, int64_t M, int64_t N, int64_t K, bool w256) {
  const int64_t gn = w256 ? 256 : 128;
  const int64_t tm = (M + X3_T - 1) / X3_T, tn = (N + gn - 1) / gn, nkc = (K + X3_KC - 1) / X3_KC;
  if (tm * tn > INT32_MAX || nkc > 65535) return set_error(ctx, HFPX_ERR_INVALID, "gemm: too large");
  // band height: the band's A rows (128 x K FP32) in the 4 MB L2 of an XCD
  int64_t band = ctx->knob[HFPX_KNOB_X3_BAND];
  if (band <= 0) band = std::max<int64_t>(1, std::min<int64_t>(16, (4 << 20) / ((int64_t)X3_T * K * 4)));
  const bool vec = ((uintptr_t)B % 16 == 0) && (ldb % 4 == 0);
  const dim3 grid((unsigned)(tm * tn));
  if (w256) {
    if (vec)
      hipLaunchKernelGGL((k_x3gemm_f<16, 4, true>), grid, dim3(256), 0, ctx->stream, C, ldc, A, lda, B, ldb, M, N, K,
                         (int)nkc, tm, tn, band);
    else
      hipLaunchKernelGGL((k_x3gemm_f<16, 4, false>), grid, dim3(256), 0, ctx->stream, C, ldc, A, lda, B, ldb, M, N,
                         K, (int)nkc, tm, tn, band);
  } else {
    if (vec)
      hipLaunchKernelGGL((k_x3gemm_f<24, 2, true>), grid, dim3(256), 0, ctx->stream, C, ldc, A, lda, B, ldb, M, N, K,
                         (int)nkc, tm, tn, band);
    else
      hipLaunchKernelGGL((k_x3gemm_f<24, 2, false>), grid, dim3(256), 0, ctx->stream, C, ldc, A, lda, B, ldb, M, N,
                         K, (int)nkc, tm, tn, band);
  }
  HFPX_TRY_HIP(ctx, hipGetLastError());
  return HFPX_OK;
}

int gemm_x3(hfpx_ctx* ctx, float* C, int64_t ldc, const float* A, int64_t lda, const float* B, int64_t ldb,
            int64_t M, int64_t N, int64_t K, int wide) {
  if (M <= 0 || N <= 0 || K <= 0) return HFPX_OK;
  if (ldc > (int64_t)1 << 24) return set_error(ctx, HFPX_ERR_INVALID, "gemm: ldc above 2^24");
  if (wide == 3 || wide == 4) return gemm_x3_fused(ctx, C, ldc, A, lda, B, ldb, M, N, K, wide == 3);
  const int64_t tm = (M + X3_T - 1) / X3_T, nkc = (K + X3_KC - 1) / X3_KC;
  const int64_t tnb = ((N + 2 * X3_T - 1) / (2 * X3_T)) * 2;  // B blocks, even
  const int64_t tn = wide ? tnb / 2 : (N + X3_T - 1) / X3_T;  // product tiles along N
  if (tm > 65535 || tnb > 65535 || nkc > 65535) return set_error(ctx, HFPX_ERR_INVALID, "gemm: too large");
  int rc = ensure_x3_work(ctx, gemm_x3_work_bytes(M, N, K));
  if (rc) return rc;
  __bf16* As = (__bf16*)ctx->x3_work;
  __bf16* Bs = As + (size_t)tm * nkc * X3_BLK;
  hipLaunchKernelGGL(k_split_a, dim3((unsigned)tm, (unsigned)nkc), dim3(256), 0, ctx->stream, A, lda, M, K,
                     (int)nkc, As, 1.0f);
  hipLaunchKernelGGL(k_split_b, dim3((unsigned)tnb, (unsigned)nkc), dim3(256), 0, ctx->stream, B, ldb, K, N,
                     (int)nkc, Bs, -1.0f);
  // band height: the band's A parts (128 rows x K, 6 bytes per entry) in
  // the 4 MB L2 of an XCD, between 1 and 16 row tiles
  int64_t band = ctx->knob[HFPX_KNOB_X3_BAND];
  if (band <= 0) band = std::max<int64_t>(1, std::min<int64_t>(16, (4 << 20) / ((int64_t)X3_T * nkc * X3_KC * 6)));
  if (wide == 2)
    hipLaunchKernelGGL((k_x3gemm<16, 4, true>), dim3((unsigned)(tm * tn)), dim3(256), 0, ctx->stream, C, ldc, As, Bs,
                       M, N, (int)nkc, tm, tn, band);
  else if (wide)
    hipLaunchKernelGGL((k_x3gemm<16, 4>), dim3((unsigned)(tm * tn)), dim3(256), 0, ctx->stream, C, ldc, As, Bs, M,
                       N, (int)nkc, tm, tn, band);
  else
    hipLaunchKernelGGL((k_x3gemm<24, 2>), dim3((unsigned)(tm * tn)), dim3(256), 0, ctx->stream, C, ldc, As, Bs, M,
                       N, (int)nkc, tm, tn, band);
  HFPX_TRY_HIP(ctx, hipGetLastError());
  return HFPX_OK;
}

}  // namespace hfpx
