// hfpx_batch.hip -- many independent cracks of uneven size: batched assembly,
// batched LU (FP64 for the small cracks, FP32 with FP64 iterative refinement
// for the large ones), batched solves, the whole sequence captured once in a
// hipGraph and replayed.
//
// No reference counterpart: the reference solves one crack per run
// (src/main.cpp:89-175); each crack here is that pipeline with fix_tip = 0
// (make_3dbem_matrix_s, add_s_inf_to_3dbem_rhs and il::LU,
// src/Solvers/system_assembly.cpp:30-274, LU.h:79-138).
//
// Layout: crack c's FP64 matrix is n_c x n_c column-major (ld n_c) at
// A64 + a_off, its FP32 copy (mixed cracks) at A32 + a32_off; its vectors
// (rhs, solution, pivots) at x_off.  The table is ordered with the mixed
// cracks first, so their vectors form one contiguous range.
//
// The LU is step-synchronous across cracks: outer blocks of BNB = 128
// columns, panels of BIB = 32.  Every launch covers all cracks that still
// have work in that step (lists sorted by size, largest first, so the
// active cracks are a prefix); a crack's panel is factorised by one
// workgroup (its column block lives in L2), swaps and triangular solves run
// one thread per column, the Schur-complement updates are one grouped MFMA
// launch over the tiles of all cracks.
#include <algorithm>
#include <cmath>
#include <numeric>

#include "hfpx_internal.h"

namespace hfpx {
// One batched Schur-complement update launch: its tile prefix table starts
// at entry off of the batch's table, and it runs tiles workgroups.
struct BTab {
  int64_t off = 0, tiles = 0;
  int64_t roff = 0, rtot = 0;  // prefix of the cracks' 128-row tiles of A (split product)
  int64_t coff = 0, ctot = 0;  // prefix of their 128-column tiles of B
};
}  // namespace hfpx

struct hfpx_batch {
  int ncr = 0;
  int n32 = 0, n64 = 0;            // cracks solved mixed / in FP64
  int64_t nmax32 = 0, nmax64 = 0;  // largest n of each group
  int64_t n_el_total = 0, x32_total = 0, x_total = 0;
  int64_t total_wg = 0;
  std::vector<hfpx::BCrack> host_cr;  // table order
  std::vector<int> input_of;          // table index -> input index
  std::vector<int> table_of;          // input index -> table index
  std::vector<int64_t> n32_sorted, n64_sorted;  // sizes along the lists
  hfpx::BCrack* cr = nullptr;
  int32_t* list32 = nullptr;  // table indices, n descending
  int32_t* list64 = nullptr;
  int32_t* list_all = nullptr;
  double* nods = nullptr;
  int32_t* conn = nullptr;
  hfpx::SrcElem<double>* elem = nullptr;
  double* cp = nullptr;
  double* A64 = nullptr;
  float* A32 = nullptr;
  double* F64 = nullptr;    // working copies of the FP64 group (their LU factors)
  int64_t f_first = 0, f_total = 0;  // A64 range of the FP64 group, copied to F64
  int32_t* ipiv = nullptr;
  int32_t* perm = nullptr;  // per crack: the interchanges composed, b_permuted[i] = b[perm[i]]
  int32_t* info = nullptr;  // per crack (table order)
  double *b = nullptr, *x = nullptr, *r = nullptr, *rs = nullptr;
  float* x32 = nullptr;
  double* berr = nullptr;   // per crack (table order)
  int32_t* tiles = nullptr;  // tile prefix tables of the batched updates
  __bf16* xsplit = nullptr;  // bf16 parts of the FP32 outer updates (largest launch)
  std::vector<hfpx::BTab> tab32, tab64;  // update launches of each group, in issue order
  // the mixed group's solves: dependency-flow sweeps over (crack, 128-row
  // block) tickets (hfpx_sweep.hip)
  int64_t ntk32 = 0;          // tickets = row blocks of the mixed group
  int32_t* tk32 = nullptr;    // (table index, level) per ticket
  int64_t* bofs = nullptr;    // per table index: first block in the block tables
  float* inv32 = nullptr;     // block inverses: L (ntk32 x 128^2), then U
  uint64_t* bflags = nullptr; // sweep flags, tickets, error words
  float* y32 = nullptr;       // permuted right-hand sides being solved
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_refine = -1;
};

namespace hfpx {

constexpr int BATCH_MAX_N = 9216;  // vectors of one crack fit in LDS (FP64)
constexpr int PT = 1024;           // threads of the per-crack workgroups

// ---------------------------------------------------------------------
// panel [j, j + jb) of every active crack: one workgroup each, partial
// pivoting (first largest |a|, LAPACK's choice), the column block read and
// written in place (it is L2-resident)
// ---------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(PT) void k_bpanel(T* __restrict__ base, int which, const BCrack* __restrict__ cr,
                                               const int32_t* __restrict__ list, int32_t* __restrict__ ipiv,
                                               int32_t* __restrict__ info, int64_t k, int64_t j) {
  const int idx = list[blockIdx.x];
  const BCrack& c = cr[idx];
  const int64_t n = c.n;
  int64_t kb, jb;
  if (!bstep(n, k, j, kb, jb)) return;
  T* a = base + bcopy_off(c, which);
  int32_t* piv = ipiv + c.x_off;
  __shared__ T red_v[PT / 64];
  __shared__ int32_t red_i[PT / 64];
  __shared__ T prow[BIB];
  const int t = threadIdx.x, lane = t & 63;
  for (int J = 0; J < jb; ++J) {
    const int64_t col = j + J;
    T* acol = a + col * n;
    T bv = T(-1);
    int32_t bi = -1;
    for (int64_t rr = col + t; rr < n; rr += PT) {
      const T v = fabs(acol[rr]);
      if (v > bv) {
        bv = v;
        bi = (int32_t)rr;
      }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const T ov = __shfl_xor(bv, off);
      const int32_t oi = __shfl_xor(bi, off);
      if (oi >= 0 && (bi < 0 || ov > bv || (ov == bv && oi < bi))) {
        bv = ov;
        bi = oi;
      }
    }
    if (lane == 0) {
      red_v[t >> 6] = bv;
      red_i[t >> 6] = bi;
    }
    __syncthreads();
    T gv = T(-1);
    int32_t gi = -1;
    for (int q = 0; q < PT / 64; ++q) {
      const T ov = red_v[q];
      const int32_t oi = red_i[q];
      if (oi >= 0 && (gi < 0 || ov > gv || (ov == gv && oi < gi))) {
        gv = ov;
        gi = oi;
      }
    }
    const int64_t p = gi;
    if (t == 0) {
      piv[col] = (int32_t)(p + 1);
      if (gv == T(0) && info[idx] == 0) info[idx] = (int32_t)(col + 1);
    }
    if (p != col && t < jb) {
      T* cc = a + (j + t) * n;
      const T tmp = cc[col];
      cc[col] = cc[p];
      cc[p] = tmp;
    }
    __syncthreads();
    if (t < jb) prow[t] = a[col + (j + t) * n];
    __syncthreads();
    const T pv = prow[J];
    const T inv = (pv == T(0)) ? T(0) : T(1) / pv;
    if (pv != T(0)) {
      for (int64_t rr = col + 1 + t; rr < n; rr += PT) {
        const T l = acol[rr] * inv;
        acol[rr] = l;
        for (int q = J + 1; q < jb; ++q) a[rr + (j + q) * n] -= l * prow[q];
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------
// row interchanges of step (k, j): inner (outer = 0) applies ipiv[j, j+jb)
// to the columns [k, k + kb) outside the panel; outer applies ipiv[k, k+kb)
// to the columns [0, n) outside the outer block.  One thread per column.
// ---------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_bswap(T* __restrict__ base, int which, const BCrack* __restrict__ cr,
                                               const int32_t* __restrict__ list, const int32_t* __restrict__ ipiv,
                                               int64_t k, int64_t j, int outer) {
  const BCrack& c = cr[list[blockIdx.y]];
  const int64_t n = c.n;
  int64_t kb, jb;
  if (!bstep(n, k, outer ? k : j, kb, jb)) return;
  const int64_t s0 = outer ? k : j, s1 = outer ? k + kb : j + jb;
  const int64_t lo = outer ? 0 : k, hi = outer ? n : k + kb;
  const int64_t col = lo + (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (col >= hi || (col >= s0 && col < s1)) return;
  T* ac = base + bcopy_off(c, which) + col * n;
  const int32_t* piv = ipiv + c.x_off;
  for (int64_t s = s0; s < s1; ++s) {
    const int64_t p = piv[s] - 1;
    if (p != s) {
      const T tmp = ac[s];
      ac[s] = ac[p];
      ac[p] = tmp;
    }
  }
}

// ---------------------------------------------------------------------
// unit-lower triangular solve of step (k, j): inner: rows [j, j + jb),
// columns [j + jb, k + kb); outer: rows [k, k + kb), columns [k + kb, n).
// The triangle (<= 128 rows) in LDS, one thread per column, two leaves of 64
// rows with the coupling applied between them.
// ---------------------------------------------------------------------
constexpr int BLF = 64;

template <typename T>
__global__ __launch_bounds__(256) void k_btrsm(T* __restrict__ base, int which, const BCrack* __restrict__ cr,
                                               const int32_t* __restrict__ list, int64_t k, int64_t j, int outer) {
  __shared__ T sL[BNB][BNB + 1];
  const BCrack& c = cr[list[blockIdx.y]];
  const int64_t n = c.n;
  int64_t kb, jb;
  if (!bstep(n, k, outer ? k : j, kb, jb)) return;
  const int64_t r0 = outer ? k : j;
  const int kk = (int)(outer ? kb : jb);
  const int64_t lo = outer ? k + kb : j + jb, hi = outer ? n : k + kb;
  if (lo + (int64_t)blockIdx.x * 256 >= hi) return;
  T* a = base + bcopy_off(c, which);
  for (int e = threadIdx.x; e < BNB * BNB; e += 256) {
    const int i = e % BNB, q = e / BNB;
    sL[i][q] = (i < kk && q < kk) ? a[(r0 + i) + (r0 + q) * n] : T(0);
  }
  __syncthreads();
  const int64_t col = lo + (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (col >= hi) return;
  T* bc = a + r0 + col * n;
  T x[BLF];
#pragma unroll
  for (int i = 0; i < BLF; ++i) x[i] = (i < kk) ? bc[i] : T(0);
#pragma unroll
  for (int i = 1; i < BLF; ++i) {
    if (i < kk) {
      T acc = x[i];
#pragma unroll
      for (int q = 0; q < i; ++q) acc -= sL[i][q] * x[q];
      x[i] = acc;
    }
  }
#pragma unroll
  for (int i = 0; i < BLF; ++i)
    if (i < kk) bc[i] = x[i];
  if (kk <= BLF) return;
  // rows 64 .. kk-1: subtract the first leaf, then solve the second leaf
  for (int i = BLF; i < kk; ++i) {
    T acc = bc[i];
#pragma unroll
    for (int q = 0; q < BLF; ++q) acc -= sL[i][q] * x[q];
    bc[i] = acc;
  }
#pragma unroll
  for (int i = 0; i < BLF; ++i) x[i] = (BLF + i < kk) ? bc[BLF + i] : T(0);
#pragma unroll
  for (int i = 1; i < BLF; ++i) {
    if (BLF + i < kk) {
      T acc = x[i];
#pragma unroll
      for (int q = 0; q < i; ++q) acc -= sL[BLF + i][BLF + q] * x[q];
      x[i] = acc;
    }
  }
#pragma unroll
  for (int i = 0; i < BLF; ++i)
    if (BLF + i < kk) bc[BLF + i] = x[i];
}

// ---------------------------------------------------------------------
// v := U^{-1} L^{-1} P v for every crack of the list (one workgroup each, v
// in LDS).  Blocks of 32 columns: wave 0 solves the diagonal block with
// shuffles (its rows preloaded into registers), then every thread updates its
// rows with the block's 32 values.
// ---------------------------------------------------------------------
// The sequential interchanges ipiv of each listed crack composed into one
// gather permutation: after them, row i holds the original row perm[i]
// (thread 0 walks them in order on an index array in LDS).
__global__ __launch_bounds__(PT) void k_bperm(const BCrack* __restrict__ cr, const int32_t* __restrict__ list,
                                              const int32_t* __restrict__ ipiv, int32_t* __restrict__ perm) {
  __shared__ int32_t idx[BATCH_MAX_N];
  const BCrack& c = cr[list[blockIdx.x]];
  const int64_t n = c.n;
  const int32_t* piv = ipiv + c.x_off;
  const int t = threadIdx.x;
  for (int64_t i = t; i < n; i += PT) idx[i] = (int32_t)i;
  __syncthreads();
  if (t == 0)
    for (int64_t s = 0; s < n; ++s) {
      const int32_t p = piv[s] - 1;
      if (p != s && p >= 0 && p < n) {
        const int32_t tmp = idx[s];
        idx[s] = idx[p];
        idx[p] = tmp;
      }
    }
  __syncthreads();
  for (int64_t i = t; i < n; i += PT) perm[c.x_off + i] = idx[i];
}

// In-place x := U^-1 L^-1 P x for every listed crack, one workgroup each,
// the vector in LDS: the interchanges as one parallel gather (perm), then
// 32-column blocks: wave 0 solves the diagonal block with shuffles, then
// every thread updates its rows below (above) it.
template <typename T>
__global__ __launch_bounds__(PT) void k_bsolve(const T* __restrict__ base, int which, const BCrack* __restrict__ cr,
                                               const int32_t* __restrict__ list, const int32_t* __restrict__ perm,
                                               T* __restrict__ vec) {
  __shared__ T v[BATCH_MAX_N];
  const BCrack& c = cr[list[blockIdx.x]];
  const int64_t n = c.n;
  const T* a = base + bcopy_off(c, which);
  const int32_t* pm = perm + c.x_off;
  T* g = vec + c.x_off;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  for (int64_t i = t; i < n; i += PT) v[i] = g[pm[i]];
  __syncthreads();
  // forward, unit lower
  for (int64_t c0 = 0; c0 < n; c0 += 32) {
    const int cb = (int)std::min<int64_t>(32, n - c0);
    if (wave == 0) {
      T rowv[32];
#pragma unroll
      for (int q = 0; q < 32; ++q) rowv[q] = (lane < cb && q < cb) ? a[(c0 + lane) + (c0 + q) * n] : T(0);
      T vi = (lane < cb) ? v[c0 + lane] : T(0);
#pragma unroll
      for (int q = 0; q < 32; ++q) {
        const T xq = __shfl(vi, q);
        if (lane > q) vi -= rowv[q] * xq;
      }
      if (lane < cb) v[c0 + lane] = vi;
    }
    __syncthreads();
    T xb[32];
#pragma unroll
    for (int q = 0; q < 32; ++q) xb[q] = (q < cb) ? v[c0 + q] : T(0);
    for (int64_t rr = c0 + cb + t; rr < n; rr += PT) {
      T acc = v[rr];
#pragma unroll
      for (int q = 0; q < 32; ++q)
        if (q < cb) acc -= a[rr + (c0 + q) * n] * xb[q];
      v[rr] = acc;
    }
    __syncthreads();
  }
  // backward, upper
  for (int64_t c1 = n; c1 > 0; c1 -= 32) {
    const int64_t c0 = std::max<int64_t>(0, c1 - 32);
    const int cb = (int)(c1 - c0);
    if (wave == 0) {
      T rowv[32];
#pragma unroll
      for (int q = 0; q < 32; ++q) rowv[q] = (lane < cb && q < cb) ? a[(c0 + lane) + (c0 + q) * n] : T(0);
      T vi = (lane < cb) ? v[c0 + lane] : T(0);
#pragma unroll
      for (int q = 31; q >= 0; --q) {
        if (q < cb) {
          if (lane == q) vi = vi / rowv[q];
          const T xq = __shfl(vi, q);
          if (lane < q) vi -= rowv[q] * xq;
        }
      }
      if (lane < cb) v[c0 + lane] = vi;
    }
    __syncthreads();
    T xb[32];
#pragma unroll
    for (int q = 0; q < 32; ++q) xb[q] = (q < cb) ? v[c0 + q] : T(0);
    for (int64_t rr = t; rr < c0; rr += PT) {
      T acc = v[rr];
#pragma unroll
      for (int q = 0; q < 32; ++q)
        if (q < cb) acc -= a[rr + (c0 + q) * n] * xb[q];
      v[rr] = acc;
    }
    __syncthreads();
  }
  for (int64_t i = t; i < n; i += PT) g[i] = v[i];
}

// r = b - A x (FP64) and, with rs != nullptr, rs = row sums of |A|; one
// thread per row, blockIdx.y = crack of the list
__global__ __launch_bounds__(256) void k_bresid(const double* __restrict__ A64, const BCrack* __restrict__ cr,
                                                const int32_t* __restrict__ list, const double* __restrict__ b,
                                                const double* __restrict__ x, double* __restrict__ r,
                                                double* __restrict__ rs) {
  const BCrack& c = cr[list[blockIdx.y]];
  const int64_t n = c.n;
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double* a = A64 + c.a_off + i;
  const double* xc = x + c.x_off;
  double acc = 0.0, sabs = 0.0;
  for (int64_t q = 0; q < n; ++q) {
    const double v = a[q * n];
    acc += v * xc[q];
    sabs += fabs(v);
  }
  r[c.x_off + i] = b[c.x_off + i] - acc;
  if (rs) rs[c.x_off + i] = sabs;
}

// per crack: normwise backward error max|r| / (max rs * max|x| + max|b|)
__global__ __launch_bounds__(256) void k_bberr(const BCrack* __restrict__ cr, const double* __restrict__ r,
                                               const double* __restrict__ x, const double* __restrict__ b,
                                               const double* __restrict__ rs, double* __restrict__ berr) {
  const BCrack& c = cr[blockIdx.x];
  __shared__ double red[4][4];
  double m[4] = {0, 0, 0, 0};
  for (int64_t i = threadIdx.x; i < c.n; i += 256) {
    const int64_t o = c.x_off + i;
    m[0] = nan_max(m[0], fabs(r[o]));
    m[1] = nan_max(m[1], rs[o]);
    m[2] = nan_max(m[2], fabs(x[o]));
    m[3] = nan_max(m[3], fabs(b[o]));
  }
  for (int q = 0; q < 4; ++q)
    for (int off = 32; off > 0; off >>= 1) m[q] = nan_max(m[q], __shfl_xor(m[q], off));
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0)
    for (int q = 0; q < 4; ++q) red[wave][q] = m[q];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 4; ++w)
      for (int q = 0; q < 4; ++q) m[q] = nan_max(m[q], red[w][q]);
    const double den = m[1] * m[2] + m[3];
    berr[blockIdx.x] = (den > 0 || den != den) ? m[0] / den : (m[0] == 0.0 ? 0.0 : m[0] / den);
  }
}

__global__ void k_bfill_i32(int32_t* __restrict__ p, int64_t cnt, int32_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < cnt) p[i] = v;
}
// mode 0: f = (float) d; mode 1: d = (double) f; mode 2: d += (double) f; mode 3: d2 = d
__global__ void k_bvec(double* __restrict__ d, float* __restrict__ f, const double* __restrict__ d2, int64_t cnt,
                       int mode) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cnt) return;
  if (mode == 0)
    f[i] = (float)d2[i];
  else if (mode == 1)
    d[i] = (double)f[i];
  else if (mode == 2)
    d[i] += (double)f[i];
  else
    d[i] = d2[i];
}

// ---------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------
static int bvec(hfpx_ctx* ctx, double* d, float* f, const double* d2, int64_t cnt, int mode) {
  if (cnt <= 0) return HFPX_OK;
  hipLaunchKernelGGL(k_bvec, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, ctx->stream, d, f, d2, cnt, mode);
  HFPX_TRY_HIP(ctx, hipGetLastError());
  return HFPX_OK;
}

// number of list entries (sizes descending) with n > v
static int count_above(const std::vector<int64_t>& sizes, int64_t v) {
  int c = 0;
  while (c < (int)sizes.size() && sizes[c] > v) ++c;
  return c;
}

// The batched LU's update launches, in the order batch_lu issues them: for
// each, the prefix sums of the cracks' tile counts (bgemm_tiles) along the
// active prefix of the list are appended to tab.  False if one launch would
// exceed 2^31 - 1 tiles.
static bool plan_updates(const std::vector<int64_t>& sizes, std::vector<int32_t>& tab, std::vector<BTab>& out) {
  if (sizes.empty()) return true;
  const int64_t nmax = sizes[0];
  auto add = [&](int nact, int64_t k, int64_t j, bool outer) {
    BTab t;
    // three prefix tables: tiles of the product, row tiles of A, column
    // tiles of B
    int64_t tot[3];
    int64_t* offs[3] = {&t.off, &t.roff, &t.coff};
    for (int w = 0; w < 3; ++w) {
      *offs[w] = (int64_t)tab.size();
      int64_t sum = 0;
      tab.push_back(0);
      for (int e = 0; e < nact; ++e) {
        int64_t r0, c0, p0, M, N, K;
        if (bgemm_shape(sizes[e], k, j, outer, r0, c0, p0, M, N, K))
          sum += w == 0 ? ((M + BGT - 1) / BGT) * ((N + BGT - 1) / BGT) : (w == 1 ? (M + BGT - 1) / BGT
                                                                                  : (N + BGT - 1) / BGT);
        if (sum > 0x7fffffff) return false;
        tab.push_back((int32_t)sum);
      }
      tot[w] = sum;
    }
    t.tiles = tot[0];
    t.rtot = tot[1];
    t.ctot = tot[2];
    out.push_back(t);
    return true;
  };
  for (int64_t k = 0; k < nmax; k += BNB) {
    for (int64_t j = k; j < k + BNB; j += BIB) {
      const int na = count_above(sizes, j);
      if (na == 0) break;
      if (!add(na, k, j, false)) return false;
    }
    const int nrest = count_above(sizes, k + BNB);
    if (nrest > 0 && !add(nrest, k, k, true)) return false;
  }
  return true;
}

template <typename T>
static int batch_lu(hfpx_ctx* ctx, hfpx_batch* bt, T* base, int which, const int32_t* list,
                    const std::vector<int64_t>& sizes, const std::vector<BTab>& plan) {
  if (sizes.empty()) return HFPX_OK;
  const int64_t nmax = sizes[0];
  const int u = which;
  size_t li = 0;  // next entry of plan (same loop structure as plan_updates)
  auto update = [&](int nact, int64_t k, int64_t j, bool outer) -> int {
    if (li >= plan.size()) return set_error(ctx, HFPX_ERR_INVALID, "batched LU: update plan out of step");
    const BTab& t = plan[li++];
    if constexpr (sizeof(T) == 4) {
      // FP32 outer updates: the bf16-split product unless an FP32 epilogue is asked for
      const int kn = ctx->knob[HFPX_KNOB_BGEMM_F32];
      if (outer && (kn == 0 || kn == 3) && bt->xsplit)
        return launch_bgemm_x3(ctx, base, which, bt->cr, list, nact, bt->tiles + t.off, t.tiles,
                               bt->tiles + t.roff, t.rtot, bt->tiles + t.coff, t.ctot, k, bt->xsplit);
    }
    return launch_bgemm<T>(ctx, base, which, bt->cr, list, nact, bt->tiles + t.off, t.tiles, k, j, outer);
  };
  int rc;
  for (int64_t k = 0; k < nmax; k += BNB) {
    const int nact = count_above(sizes, k);
    for (int64_t j = k; j < k + BNB; j += BIB) {
      const int na = count_above(sizes, j);
      if (na == 0) break;
      {
        const int tok = prof_begin(ctx, HFPX_KC_PANEL);
        hipLaunchKernelGGL(k_bpanel<T>, dim3((unsigned)na), dim3(PT), 0, ctx->stream, base, u, bt->cr, list,
                           bt->ipiv, bt->info, k, j);
        prof_end(ctx, tok, HFPX_KC_PANEL, 0.0, 0.0);
      }
      hipLaunchKernelGGL(k_bswap<T>, dim3(1, (unsigned)na), dim3(256), 0, ctx->stream, base, u, bt->cr, list,
                         bt->ipiv, k, j, 0);
      hipLaunchKernelGGL(k_btrsm<T>, dim3(1, (unsigned)na), dim3(256), 0, ctx->stream, base, u, bt->cr, list, k, j,
                         0);
      HFPX_TRY_HIP(ctx, hipGetLastError());
      if ((rc = update(na, k, j, false))) return rc;
    }
    hipLaunchKernelGGL(k_bswap<T>, dim3((unsigned)((nmax + 255) / 256), (unsigned)nact), dim3(256), 0, ctx->stream,
                       base, u, bt->cr, list, bt->ipiv, k, k, 1);
    const int nrest = count_above(sizes, k + BNB);
    if (nrest > 0) {
      const int64_t rest = nmax - k - BNB;
      hipLaunchKernelGGL(k_btrsm<T>, dim3((unsigned)((rest + 255) / 256), (unsigned)nrest), dim3(256), 0,
                         ctx->stream, base, u, bt->cr, list, k, k, 1);
      HFPX_TRY_HIP(ctx, hipGetLastError());
      if ((rc = update(nrest, k, k, true))) return rc;
    }
    HFPX_TRY_HIP(ctx, hipGetLastError());
  }
  if (li != plan.size()) return set_error(ctx, HFPX_ERR_INVALID, "batched LU: update plan out of step");
  return HFPX_OK;
}

// y[x_off + i] = x[x_off + perm[x_off + i]] (inv = false) or back,
// x[...] = y[...] unpermuted positions (inv = true: a plain copy), for
// every crack of the list
__global__ void k_bgather_f(const BCrack* __restrict__ cr, const int32_t* __restrict__ list,
                            const int32_t* __restrict__ perm, const float* __restrict__ x, float* __restrict__ y) {
  const BCrack& c = cr[list[blockIdx.y]];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= c.n) return;
  y[c.x_off + i] = x[c.x_off + perm[c.x_off + i]];
}

__global__ void k_bcopy_f(const float* __restrict__ a, float* __restrict__ b, int64_t cnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < cnt) b[i] = a[i];
}

// x32 := U^-1 L^-1 P x32 for every crack of the mixed group: the
// interchanges as one gather, then the persistent dependency-flow sweeps
// with the block inverses of this factorisation
static int batch_sweep_vec(hfpx_ctx* ctx, hfpx_batch* bt) {
  if (bt->n32 <= 0) return HFPX_OK;
  hipLaunchKernelGGL(k_bgather_f, dim3((unsigned)((bt->nmax32 + 255) / 256), (unsigned)bt->n32), dim3(256), 0,
                     ctx->stream, bt->cr, bt->list32, bt->perm, bt->x32, bt->y32);
  HFPX_TRY_HIP(ctx, hipGetLastError());
  int rc = launch_bsweeps(ctx, bt->A32, bt->cr, bt->tk32, bt->ntk32, bt->bofs, bt->inv32,
                          bt->inv32 + bt->ntk32 * 128 * 128, bt->y32, bt->bflags);
  if (rc) return rc;
  hipLaunchKernelGGL(k_bcopy_f, dim3((unsigned)((bt->x32_total + 255) / 256)), dim3(256), 0, ctx->stream, bt->y32,
                     bt->x32, bt->x32_total);
  HFPX_TRY_HIP(ctx, hipGetLastError());
  return HFPX_OK;
}

template <typename T>
static int batch_solve_vec(hfpx_ctx* ctx, hfpx_batch* bt, const T* base, int which, const int32_t* list, int cnt,
                           T* vec) {
  if (cnt <= 0) return HFPX_OK;
  hipLaunchKernelGGL(k_bsolve<T>, dim3((unsigned)cnt), dim3(PT), 0, ctx->stream, base, which, bt->cr,
                     list, bt->perm, vec);
  HFPX_TRY_HIP(ctx, hipGetLastError());
  return HFPX_OK;
}

// The whole batch: assembly, right-hand sides, both LU groups, solves and
// n_refine refinement steps of the mixed group, backward errors.
static int batch_enqueue(hfpx_ctx* ctx, hfpx_batch* bt, int n_refine) {
  int rc;
  if ((rc = launch_bassemble(ctx, bt->elem, bt->cp, 6 * bt->n_el_total, bt->cr, bt->ncr, bt->total_wg, bt->A64,
                             bt->A32)))
    return rc;
  if ((rc = launch_brhs(ctx, bt->elem, bt->n_el_total, bt->cr, bt->ncr, bt->b))) return rc;
  hipLaunchKernelGGL(k_bfill_i32, dim3((unsigned)((bt->ncr + 255) / 256)), dim3(256), 0, ctx->stream, bt->info,
                     (int64_t)bt->ncr, 0);
  // FP64 group: LU in a working copy (the assembled matrix stays for the
  // backward error), solve in place in x
  if ((rc = bvec(ctx, bt->F64, nullptr, bt->A64 + bt->f_first, bt->f_total, 3))) return rc;
  if ((rc = batch_lu<double>(ctx, bt, bt->F64, B_F64, bt->list64, bt->n64_sorted, bt->tab64))) return rc;
  if (bt->n64 > 0)
    hipLaunchKernelGGL(k_bperm, dim3((unsigned)bt->n64), dim3(PT), 0, ctx->stream, bt->cr, bt->list64, bt->ipiv,
                       bt->perm);
  if ((rc = bvec(ctx, bt->x + bt->x32_total, nullptr, bt->b + bt->x32_total, bt->x_total - bt->x32_total, 3)))
    return rc;
  if ((rc = batch_solve_vec<double>(ctx, bt, bt->F64, B_F64, bt->list64, bt->n64, bt->x))) return rc;
  // mixed group: FP32 LU, x0 = LU^-1 b, then refinement against the FP64 matrix
  if ((rc = batch_lu<float>(ctx, bt, bt->A32, B_A32, bt->list32, bt->n32_sorted, bt->tab32))) return rc;
  if (bt->n32 > 0)
    hipLaunchKernelGGL(k_bperm, dim3((unsigned)bt->n32), dim3(PT), 0, ctx->stream, bt->cr, bt->list32, bt->ipiv,
                       bt->perm);
  if ((rc = launch_bblock_inv(ctx, bt->A32, bt->cr, bt->tk32, bt->ntk32, bt->bofs, bt->inv32,
                              bt->inv32 + bt->ntk32 * 128 * 128)))
    return rc;
  if ((rc = bvec(ctx, nullptr, bt->x32, bt->b, bt->x32_total, 0))) return rc;
  if ((rc = batch_sweep_vec(ctx, bt))) return rc;
  if ((rc = bvec(ctx, bt->x, bt->x32, nullptr, bt->x32_total, 1))) return rc;
  const unsigned gr32 = (unsigned)((bt->nmax32 + 255) / 256);
  for (int it = 0; it < n_refine && bt->n32 > 0; ++it) {
    hipLaunchKernelGGL(k_bresid, dim3(gr32, (unsigned)bt->n32), dim3(256), 0, ctx->stream, bt->A64, bt->cr,
                       bt->list32, bt->b, bt->x, bt->r, nullptr);
    if ((rc = bvec(ctx, nullptr, bt->x32, bt->r, bt->x32_total, 0))) return rc;
    if ((rc = batch_sweep_vec(ctx, bt))) return rc;
    if ((rc = bvec(ctx, bt->x, bt->x32, nullptr, bt->x32_total, 2))) return rc;
  }
  // backward errors of every crack
  const int64_t nmax = std::max(bt->nmax32, bt->nmax64);
  hipLaunchKernelGGL(k_bresid, dim3((unsigned)((nmax + 255) / 256), (unsigned)bt->ncr), dim3(256), 0, ctx->stream,
                     bt->A64, bt->cr, bt->list_all, bt->b, bt->x, bt->r, bt->rs);
  hipLaunchKernelGGL(k_bberr, dim3((unsigned)bt->ncr), dim3(256), 0, ctx->stream, bt->cr, bt->r, bt->x, bt->b,
                     bt->rs, bt->berr);
  HFPX_TRY_HIP(ctx, hipGetLastError());
  return HFPX_OK;
}

}  // namespace hfpx

using namespace hfpx;

extern "C" {

int hfpx_batch_create(hfpx_ctx* ctx, int n_cracks, const int64_t* n_nodes, const int64_t* n_el,
                      const double* nods_host, const int64_t* conn_host, double beta, const double* shear_m,
                      const double* poiss_r, const double* sinf_host, int64_t mixed_min_n, hfpx_batch** out) {
  if (!ctx || !out || n_cracks < 1 || !n_nodes || !n_el || !nods_host || !conn_host || !shear_m || !poiss_r ||
      !sinf_host)
    return ctx ? set_error(ctx, HFPX_ERR_INVALID, "batch_create: invalid arguments") : HFPX_ERR_INVALID;
  *out = nullptr;
  HFPX_TRY_HIP(ctx, hipSetDevice(ctx->device));
  // validate and lay out: mixed cracks first (their vectors contiguous), each
  // group in input order
  std::vector<int64_t> node_off(n_cracks + 1, 0), el_in(n_cracks + 1, 0);
  for (int i = 0; i < n_cracks; ++i) {
    const int64_t n = 18 * n_el[i];
    if (n_el[i] < 1 || n_nodes[i] < 3 || n > BATCH_MAX_N)
      return set_error(ctx, HFPX_ERR_INVALID, "batch_create: crack %d has %lld elements (1 .. %d allowed)", i,
                       (long long)n_el[i], BATCH_MAX_N / 18);
    if (!(poiss_r[i] < 1.0) || !(shear_m[i] > 0.0))
      return set_error(ctx, HFPX_ERR_INVALID, "batch_create: crack %d: need shear modulus > 0, Poisson ratio < 1", i);
    node_off[i + 1] = node_off[i] + n_nodes[i];
    el_in[i + 1] = el_in[i] + n_el[i];
  }
  for (int64_t q = 0; q < 3 * node_off[n_cracks]; ++q)
    if (!std::isfinite(nods_host[q])) return set_error(ctx, HFPX_ERR_INVALID, "batch_create: non-finite coordinate");
  for (int i = 0; i < n_cracks; ++i)
    for (int64_t e = el_in[i]; e < el_in[i + 1]; ++e)
      for (int v = 0; v < 3; ++v) {
        const int64_t id = conn_host[3 * e + v];
        if (id < 0 || id >= n_nodes[i])
          return set_error(ctx, HFPX_ERR_INVALID, "batch_create: crack %d element %lld: node %lld out of range", i,
                           (long long)(e - el_in[i]), (long long)id);
      }
  for (int i = 0; i < n_cracks; ++i)
    for (int64_t e = el_in[i]; e < el_in[i + 1]; ++e) {
      const double* nb = nods_host + 3 * node_off[i];
      if (degenerate_triangle(nb + 3 * conn_host[3 * e], nb + 3 * conn_host[3 * e + 1], nb + 3 * conn_host[3 * e + 2]))
        return set_error(ctx, HFPX_ERR_INVALID, "batch_create: crack %d element %lld is degenerate", i,
                         (long long)(e - el_in[i]));
    }
  hfpx_batch* bt = new hfpx_batch();
  bt->ncr = n_cracks;
  std::vector<int> order;
  for (int i = 0; i < n_cracks; ++i)
    if (18 * n_el[i] >= mixed_min_n) order.push_back(i);
  bt->n32 = (int)order.size();
  for (int i = 0; i < n_cracks; ++i)
    if (18 * n_el[i] < mixed_min_n) order.push_back(i);
  bt->n64 = n_cracks - bt->n32;
  bt->input_of = order;
  bt->table_of.assign(n_cracks, 0);
  int64_t el_off = 0, a_off = 0, a32_off = 0, x_off = 0, wg_off = 0;
  int64_t f_first = 0;  // A64 offset where the FP64 group starts
  for (int ti = 0; ti < bt->n32; ++ti) f_first += 18 * n_el[order[ti]] * 18 * n_el[order[ti]];
  std::vector<int32_t> conn_t;
  conn_t.reserve(3 * el_in[n_cracks]);
  for (int ti = 0; ti < n_cracks; ++ti) {
    const int i = order[ti];
    bt->table_of[i] = ti;
    BCrack c;
    c.el_off = el_off;
    c.n_el = n_el[i];
    c.n = 18 * n_el[i];
    c.a_off = a_off;
    c.a32_off = ti < bt->n32 ? a32_off : -1;
    c.f_off = ti < bt->n32 ? -1 : a_off - f_first;
    c.x_off = x_off;
    c.wg_off = wg_off;
    c.G = shear_m[i];
    c.nu = poiss_r[i];
    for (int q = 0; q < 6; ++q) c.sinf[q] = sinf_host[6 * i + q];
    bt->host_cr.push_back(c);
    for (int64_t e = el_in[i]; e < el_in[i + 1]; ++e)
      for (int v = 0; v < 3; ++v) conn_t.push_back((int32_t)(node_off[i] + conn_host[3 * e + v]));
    el_off += c.n_el;
    a_off += c.n * c.n;
    if (ti < bt->n32) a32_off += c.n * c.n;
    x_off += c.n;
    if (ti < bt->n32) bt->x32_total = x_off;
    wg_off += ((6 * c.n_el + 63) / 64) * ((c.n_el + 3) / 4);
  }
  bt->n_el_total = el_off;
  bt->f_first = f_first;
  bt->f_total = a_off - f_first;
  bt->x_total = x_off;
  bt->total_wg = wg_off;
  // lists by size, largest first
  auto sorted = [&](int from, int to, std::vector<int64_t>& sizes) {
    std::vector<int32_t> l;
    for (int ti = from; ti < to; ++ti) l.push_back(ti);
    std::stable_sort(l.begin(), l.end(), [&](int32_t p, int32_t q) { return bt->host_cr[p].n > bt->host_cr[q].n; });
    for (int32_t ti : l) sizes.push_back(bt->host_cr[ti].n);
    return l;
  };
  std::vector<int32_t> l32 = sorted(0, bt->n32, bt->n32_sorted), l64 = sorted(bt->n32, n_cracks, bt->n64_sorted);
  bt->nmax32 = bt->n32 ? bt->n32_sorted[0] : 0;
  bt->nmax64 = bt->n64 ? bt->n64_sorted[0] : 0;
  std::vector<int32_t> lall(n_cracks);
  std::iota(lall.begin(), lall.end(), 0);
  auto fail = [&](int rc) {
    hfpx_batch_destroy(bt);
    return rc;
  };
  std::vector<int32_t> tab;
  if (!plan_updates(bt->n32_sorted, tab, bt->tab32) || !plan_updates(bt->n64_sorted, tab, bt->tab64))
    return fail(set_error(ctx, HFPX_ERR_INVALID, "batch: too many tiles in one update"));
  // the mixed group's sweep tickets: level-major, largest crack first in a
  // level; block tables by table index
  std::vector<int32_t> tk;
  std::vector<int64_t> bofs(n_cracks, 0);
  {
    int64_t off = 0, maxb = 0;
    for (int ti = 0; ti < bt->n32; ++ti) {
      bofs[ti] = off;
      const int64_t nb = (bt->host_cr[ti].n + 127) / 128;
      off += nb;
      maxb = std::max(maxb, nb);
    }
    for (int64_t q = 0; q < maxb; ++q)
      for (int32_t ti : l32)
        if ((bt->host_cr[ti].n + 127) / 128 > q) {
          tk.push_back(ti);
          tk.push_back((int32_t)q);
        }
    bt->ntk32 = off;
  }
  size_t split_bytes = 0;  // the largest FP32 outer update's bf16 parts
  for (const BTab& t : bt->tab32) split_bytes = std::max(split_bytes, bgemm_x3_split_bytes(t.rtot, t.ctot));
  auto dalloc = [&](void** p, size_t bytes) -> int {
    HFPX_TRY_HIP(ctx, hipMalloc(p, std::max<size_t>(bytes, 256)));
    return HFPX_OK;
  };
  int rc;
  const int64_t nn = node_off[n_cracks];
  if ((rc = dalloc((void**)&bt->cr, sizeof(BCrack) * n_cracks)) ||
      (rc = dalloc((void**)&bt->list32, sizeof(int32_t) * (bt->n32 + 1))) ||
      (rc = dalloc((void**)&bt->list64, sizeof(int32_t) * (bt->n64 + 1))) ||
      (rc = dalloc((void**)&bt->list_all, sizeof(int32_t) * n_cracks)) ||
      (rc = dalloc((void**)&bt->nods, sizeof(double) * 3 * nn)) ||
      (rc = dalloc((void**)&bt->conn, sizeof(int32_t) * 3 * el_off)) ||
      (rc = dalloc((void**)&bt->elem, sizeof(SrcElem<double>) * el_off)) ||
      (rc = dalloc((void**)&bt->cp, sizeof(double) * 36 * el_off)) ||
      (rc = dalloc((void**)&bt->A64, sizeof(double) * a_off)) ||
      (rc = dalloc((void**)&bt->A32, sizeof(float) * a32_off)) ||
      (rc = dalloc((void**)&bt->F64, sizeof(double) * bt->f_total)) ||
      (rc = dalloc((void**)&bt->ipiv, sizeof(int32_t) * x_off)) ||
      (rc = dalloc((void**)&bt->perm, sizeof(int32_t) * x_off)) ||
      (rc = dalloc((void**)&bt->info, sizeof(int32_t) * n_cracks)) ||
      (rc = dalloc((void**)&bt->b, sizeof(double) * x_off)) || (rc = dalloc((void**)&bt->x, sizeof(double) * x_off)) ||
      (rc = dalloc((void**)&bt->r, sizeof(double) * x_off)) || (rc = dalloc((void**)&bt->rs, sizeof(double) * x_off)) ||
      (rc = dalloc((void**)&bt->x32, sizeof(float) * std::max<int64_t>(1, bt->x32_total))) ||
      (rc = dalloc((void**)&bt->berr, sizeof(double) * n_cracks)) ||
      (rc = dalloc((void**)&bt->tiles, sizeof(int32_t) * tab.size())) ||
      (split_bytes && (rc = dalloc((void**)&bt->xsplit, split_bytes))) ||
      (rc = dalloc((void**)&bt->tk32, sizeof(int32_t) * std::max<size_t>(1, tk.size()))) ||
      (rc = dalloc((void**)&bt->bofs, sizeof(int64_t) * n_cracks)) ||
      (rc = dalloc((void**)&bt->inv32, sizeof(float) * 2 * (size_t)bt->ntk32 * 128 * 128)) ||
      (rc = dalloc((void**)&bt->bflags, sizeof(uint64_t) * (2 * (size_t)bt->ntk32 + 3))) ||
      (rc = dalloc((void**)&bt->y32, sizeof(float) * std::max<int64_t>(1, bt->x32_total))))
    return fail(rc);
  auto up = [&](void* d, const void* h, size_t bytes) -> int {
    if (bytes) HFPX_TRY_HIP(ctx, hipMemcpy(d, h, bytes, hipMemcpyHostToDevice));
    return HFPX_OK;
  };
  std::vector<double> nods3(3 * nn);
  for (int64_t q = 0; q < 3 * nn; ++q) nods3[q] = nods_host[q];
  if ((rc = up(bt->cr, bt->host_cr.data(), sizeof(BCrack) * n_cracks)) ||
      (rc = up(bt->list32, l32.data(), sizeof(int32_t) * l32.size())) ||
      (rc = up(bt->list64, l64.data(), sizeof(int32_t) * l64.size())) ||
      (rc = up(bt->list_all, lall.data(), sizeof(int32_t) * n_cracks)) ||
      (rc = up(bt->nods, nods3.data(), sizeof(double) * 3 * nn)) ||
      (rc = up(bt->conn, conn_t.data(), sizeof(int32_t) * conn_t.size())) ||
      (rc = up(bt->tiles, tab.data(), sizeof(int32_t) * tab.size())) ||
      (rc = up(bt->tk32, tk.data(), sizeof(int32_t) * tk.size())) ||
      (rc = up(bt->bofs, bofs.data(), sizeof(int64_t) * n_cracks)))
    return fail(rc);
  // element geometry of all cracks at once (the element table is in crack order)
  hfpx_mesh m;
  m.n_el = el_off;
  m.beta = beta;
  m.nods = bt->nods;
  m.conn = bt->conn;
  m.elem = bt->elem;
  m.cp = bt->cp;
  if ((rc = launch_geometry(ctx, &m))) return fail(rc);
  if (hipStreamSynchronize(ctx->stream) != hipSuccess) return fail(set_error(ctx, HFPX_ERR_HIP, "batch geometry"));
  *out = bt;
  return HFPX_OK;
}

int hfpx_batch_destroy(hfpx_batch* bt) {
  if (!bt) return HFPX_OK;
  if (bt->exec) (void)hipGraphExecDestroy(bt->exec);
  if (bt->graph) (void)hipGraphDestroy(bt->graph);
  void* ps[] = {bt->cr, bt->list32, bt->list64, bt->list_all, bt->nods, bt->conn, bt->elem, bt->cp, bt->A64, bt->A32, bt->F64,
                bt->ipiv, bt->perm, bt->info, bt->b, bt->x, bt->r, bt->rs, bt->x32, bt->berr, bt->tiles, bt->xsplit,
                bt->tk32, bt->bofs, bt->inv32, bt->bflags, bt->y32};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  delete bt;
  return HFPX_OK;
}

// after a solve: a bounded wait of the last sweeps that gave up (the error
// words follow the flags and tickets)
static int batch_sweep_status(hfpx_ctx* ctx, hfpx_batch* bt) {
  if (bt->ntk32 <= 0) return HFPX_OK;
  int32_t e[2] = {0, 0};
  HFPX_TRY_HIP(ctx, hipMemcpy(e, bt->bflags + 2 * bt->ntk32 + 2, sizeof(e), hipMemcpyDeviceToHost));
  if (e[0] || e[1]) return set_error(ctx, HFPX_ERR_HIP, "batch: a sweep dependency wait timed out");
  return HFPX_OK;
}

int hfpx_batch_solve(hfpx_ctx* ctx, hfpx_batch* bt, int n_refine, int use_graph) {
  if (!ctx || !bt || n_refine < 0 || n_refine > 50) return HFPX_ERR_INVALID;
  HFPX_TRY_HIP(ctx, hipSetDevice(ctx->device));
  int rc;
  if (!use_graph) {
    if ((rc = batch_enqueue(ctx, bt, n_refine))) return rc;
    HFPX_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return batch_sweep_status(ctx, bt);
  }
  if (!bt->exec || bt->graph_refine != n_refine) {
    if (bt->exec) (void)hipGraphExecDestroy(bt->exec);
    if (bt->graph) (void)hipGraphDestroy(bt->graph);
    bt->exec = nullptr;
    bt->graph = nullptr;
    HFPX_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const bool prof = ctx->prof;
    ctx->prof = false;  // no event records inside the captured graph
    HFPX_TRY_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    rc = batch_enqueue(ctx, bt, n_refine);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    ctx->prof = prof;
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    HFPX_TRY_HIP(ctx, e);
    bt->graph = g;
    HFPX_TRY_HIP(ctx, hipGraphInstantiate(&bt->exec, g, nullptr, nullptr, 0));
    bt->graph_refine = n_refine;
  }
  HFPX_TRY_HIP(ctx, hipGraphLaunch(bt->exec, ctx->stream));
  HFPX_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return batch_sweep_status(ctx, bt);
}

int hfpx_batch_assemble(hfpx_ctx* ctx, hfpx_batch* bt) {
  if (!ctx || !bt) return HFPX_ERR_INVALID;
  HFPX_TRY_HIP(ctx, hipSetDevice(ctx->device));
  int rc;
  if ((rc = launch_bassemble(ctx, bt->elem, bt->cp, 6 * bt->n_el_total, bt->cr, bt->ncr, bt->total_wg, bt->A64,
                             bt->A32)))
    return rc;
  if ((rc = launch_brhs(ctx, bt->elem, bt->n_el_total, bt->cr, bt->ncr, bt->b))) return rc;
  HFPX_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return HFPX_OK;
}

int hfpx_batch_rhs(hfpx_ctx* ctx, hfpx_batch* bt, int crack, double* b_host) {
  if (!ctx || !bt || crack < 0 || crack >= bt->ncr || !b_host) return HFPX_ERR_INVALID;
  const BCrack& c = bt->host_cr[bt->table_of[crack]];
  HFPX_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  HFPX_TRY_HIP(ctx, hipMemcpy(b_host, bt->b + c.x_off, sizeof(double) * c.n, hipMemcpyDeviceToHost));
  return HFPX_OK;
}

int hfpx_batch_solution(hfpx_ctx* ctx, hfpx_batch* bt, int crack, double* x_host, int64_t* n_out, double* berr,
                        int64_t* info) {
  if (!ctx || !bt || crack < 0 || crack >= bt->ncr) return HFPX_ERR_INVALID;
  const int ti = bt->table_of[crack];
  const BCrack& c = bt->host_cr[ti];
  if (n_out) *n_out = c.n;
  HFPX_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (x_host) HFPX_TRY_HIP(ctx, hipMemcpy(x_host, bt->x + c.x_off, sizeof(double) * c.n, hipMemcpyDeviceToHost));
  if (berr) HFPX_TRY_HIP(ctx, hipMemcpy(berr, bt->berr + ti, sizeof(double), hipMemcpyDeviceToHost));
  if (info) {
    int32_t v = 0;
    HFPX_TRY_HIP(ctx, hipMemcpy(&v, bt->info + ti, sizeof(int32_t), hipMemcpyDeviceToHost));
    *info = v;
  }
  return HFPX_OK;
}

int hfpx_batch_matrix(hfpx_ctx* ctx, hfpx_batch* bt, int crack, double* A_host, int64_t* n_out) {
  if (!ctx || !bt || crack < 0 || crack >= bt->ncr) return HFPX_ERR_INVALID;
  const BCrack& c = bt->host_cr[bt->table_of[crack]];
  if (n_out) *n_out = c.n;
  HFPX_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (A_host)
    HFPX_TRY_HIP(ctx, hipMemcpy(A_host, bt->A64 + c.a_off, sizeof(double) * c.n * c.n, hipMemcpyDeviceToHost));
  return HFPX_OK;
}

}  // extern "C"
