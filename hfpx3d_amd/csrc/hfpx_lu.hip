// hfpx_lu.hip -- dense LU with partial pivoting and triangular solves on
// gfx950, FP64 and FP32.
//
// Replaces il::LU<il::Array2D<double>> (il/linear_algebra/dense/factorization/
// LU.h:79-138), i.e. LAPACKE_dgetrf / LAPACKE_dgetrs: same result semantics
// (P A = L U in place, unit lower L, 1-based sequential row interchanges in
// ipiv, info = first exactly-zero pivot column), FP32 variants for the
// mixed-precision solve.
//
// Algorithm: right-looking blocked LU.  Each outer block of NB columns is
// factorised recursively: left half, row interchanges + unit-lower
// triangular solve + Schur-complement update of the right half, right half,
// down to panels of IB columns.  A panel is factorised by one cooperative
// kernel whose workgroups keep their rows in registers and meet once per
// column at a grid barrier (pivot search -> interchange -> rank-1 update).
// After each outer block the trailing matrix gets the interchanges, the
// triangular solve for its U rows and one update C -= L21 U12 with K = NB on
// MFMA (hfpx_gemm.hip).
#include <algorithm>
#include <type_traits>

#include "hfpx_internal.h"

namespace hfpx {

__device__ __forceinline__ void fence_agent_release() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
}
__device__ __forceinline__ void fence_agent_acquire() {
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// ---------------------------------------------------------------------
// Grid barrier for co-resident (cooperatively launched) workgroups.  The
// arrival counter only grows: the host passes its value before the launch
// (base), so barrier b of a launch with nwg workgroups completes when the
// counter reaches base + (b + 1) nwg.
// Memory ordering is paid once per workgroup, not per wave: the workgroup
// barrier makes every wave's stores happen-before thread 0, whose agent-scope
// release fence publishes them (one L2 write-back per workgroup); thread 0
// then arrives and polls with relaxed atomics and ends with an agent-scope
// acquire fence, which the closing workgroup barrier extends to the other
// waves.  The wait is bounded: after ~2^24 polls thread 0 raises *err and
// leaves (co-residency, guaranteed by the launch, makes this unreachable),
// and every 256 polls it looks at *err, so that once one wait has failed
// every other wait of the factorisation gives up at once instead of
// running to its own bound.
// ---------------------------------------------------------------------
__device__ __forceinline__ bool lu_failed(const int32_t* err) {
  return __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
}

__device__ __forceinline__ void grid_barrier(uint64_t* ctr, uint64_t target, int32_t* err) {
  __syncthreads();
  if (threadIdx.x == 0) {
    fence_agent_release();
    __hip_atomic_fetch_add(ctr, (uint64_t)1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t polls = 0;
    while (__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      __builtin_amdgcn_s_sleep(1);
      ++polls;
      if ((polls & 255u) == 0 && lu_failed(err)) break;
      if (polls > (1u << 24)) {
        __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
    }
    fence_agent_acquire();
  }
  __syncthreads();
}

// A panel launched after a wait of the same factorisation has failed does
// no work: it records identity interchanges (so later launches stay
// well-defined) and returns; the host reports the failure.
__device__ __forceinline__ bool panel_skip(const int32_t* err, int64_t j0, int jb, int32_t* ipiv) {
  if (!lu_failed(err)) return false;
  if (blockIdx.x == 0 && (int)threadIdx.x < jb) ipiv[j0 + threadIdx.x] = (int32_t)(j0 + threadIdx.x + 1);
  return true;
}

// ---------------------------------------------------------------------
// Cooperative panel factorisation of A[j0:n, j0:j0+jb] (jb <= IB).
// Thread t of workgroup w holds row w*PWG + t (relative to j0) of all jb
// columns in registers.  Per column: workgroup arg-max -> candidate (value,
// row, and the row itself) published -> grid barrier -> every wave reduces
// the candidates redundantly -> interchange + rank-1 update in registers.
// Pivot choice is LAPACK's: the largest |a|, the first such row on ties.
// ---------------------------------------------------------------------
constexpr int PWG = 1024;          // threads (rows) per panel workgroup
constexpr int PANEL_MAX_WG = 1024;  // candidate slots

template <typename T>
struct PanelScratch {
  T val[2][PANEL_MAX_WG];        // per-workgroup candidate |pivot| (by column parity)
  int32_t idx[2][PANEL_MAX_WG];  // per-workgroup candidate row (relative to j0)
};

template <typename T>
struct PanelCtx {
  int nwg, w, t, jb;
  int64_t j0;
  int32_t row;  // row (relative to j0) held by this thread, -1 if none
  PanelScratch<T>* ps;
  T* rows;
  int32_t* err;
  uint64_t* ctr;
  uint64_t base;
  T* red_v;
  int32_t* red_i;
  T* s_piv;
  int32_t* s_ipiv;  // pivots of the panel's columns (LDS, written out at the end)
  int32_t* s_info;  // first column with a zero pivot, -1: none
};

// (bv, bi) := better of itself and (ov, oi): larger value, then smaller row;
// a row of -1 marks "no candidate".
template <typename T>
__device__ __forceinline__ bool pick(T& bv, int32_t& bi, T ov, int32_t oi) {
  if (oi >= 0 && (bi < 0 || ov > bv || (ov == bv && oi < bi))) {
    bv = ov;
    bi = oi;
    return true;
  }
  return false;
}

// Wave-wide maxima on DPP lane moves (no LDS round trips): xor 1 and 2
// inside quads (quad_perm), then row_half_mirror and row_mirror inside rows
// of 16, then row_bcast15 / row_bcast31 fold the rows into lane 63, which
// is broadcast.  Lanes a move does not write keep 0, the identity.
template <int CTRL, int ROWS>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROWS, 0xf, false);
}
template <int CTRL, int ROWS>
__device__ __forceinline__ uint64_t dpp_u64(uint64_t v) {
  return ((uint64_t)dpp_u32<CTRL, ROWS>((uint32_t)(v >> 32)) << 32) | dpp_u32<CTRL, ROWS>((uint32_t)v);
}
__device__ __forceinline__ uint64_t wave_max_u64(uint64_t v) {
  uint64_t o;
  o = dpp_u64<0xB1, 0xf>(v);  v = o > v ? o : v;   // quad_perm [1,0,3,2]
  o = dpp_u64<0x4E, 0xf>(v);  v = o > v ? o : v;   // quad_perm [2,3,0,1]
  o = dpp_u64<0x141, 0xf>(v); v = o > v ? o : v;   // row_half_mirror
  o = dpp_u64<0x140, 0xf>(v); v = o > v ? o : v;   // row_mirror
  o = dpp_u64<0x142, 0xa>(v); v = o > v ? o : v;   // row_bcast15 -> rows 1, 3
  o = dpp_u64<0x143, 0xc>(v); v = o > v ? o : v;   // row_bcast31 -> rows 2, 3
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), 63);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, 63);
  return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
  uint32_t o;
  o = dpp_u32<0xB1, 0xf>(v);  v = o > v ? o : v;
  o = dpp_u32<0x4E, 0xf>(v);  v = o > v ? o : v;
  o = dpp_u32<0x141, 0xf>(v); v = o > v ? o : v;
  o = dpp_u32<0x140, 0xf>(v); v = o > v ? o : v;
  o = dpp_u32<0x142, 0xa>(v); v = o > v ? o : v;
  o = dpp_u32<0x143, 0xc>(v); v = o > v ? o : v;
  return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

// Wave-wide arg-max of the lanes' (bv >= 0, row bi; bi < 0 = none), the
// same order as pick: largest value, then smallest row.  Non-negative IEEE
// values order as their bit patterns, so FP32 is one maximum of the 64-bit
// key bits(bv):~bi; FP64 takes the largest bits(bv), then the largest ~bi
// among the lanes holding it.  The result is uniform over the wave.
template <typename T>
__device__ __forceinline__ void wave_argmax(T& bv, int32_t& bi) {
  const bool valid = bi >= 0;
  if constexpr (sizeof(T) == 4) {
    uint64_t k = valid ? (((uint64_t)__float_as_uint(bv) << 32) | (uint32_t)~bi) : 0;
    k = wave_max_u64(k);
    bv = k ? __uint_as_float((uint32_t)(k >> 32)) : T(-1);
    bi = k ? (int32_t)~(uint32_t)k : -1;
  } else {
    const uint64_t bits = valid ? (uint64_t)__double_as_longlong(bv) : 0;
    const uint64_t mb = wave_max_u64(bits);
    uint32_t k2 = (valid && bits == mb) ? ~(uint32_t)bi : 0u;
    k2 = wave_max_u32(k2);
    bv = k2 ? __longlong_as_double((long long)mb) : T(-1);
    bi = k2 ? (int32_t)~k2 : -1;
  }
}

// Workgroup arg-max of column J over the rows the threads hold (uniform
// result; one workgroup barrier).
template <typename T, int IB>
__device__ __forceinline__ void wg_argmax(const T (&a)[IB], int J, int32_t row, int t, T* red_v, int32_t* red_i,
                                          T& bv, int32_t& bi) {
  const int lane = t & 63;
  bv = T(-1);
  bi = -1;
  if (row >= J) {
    bv = fabs(a[J]);
    bi = row;
  }
  wave_argmax(bv, bi);
  if (lane == 0) {
    red_v[t >> 6] = bv;
    red_i[t >> 6] = bi;
  }
  __syncthreads();
  bv = T(-1);
  bi = -1;
  if (lane < PWG / 64) {
    bv = red_v[lane];
    bi = red_i[lane];
  }
  wave_argmax(bv, bi);
}

// The panel's pivots are kept in LDS and written once at the end:
// s_ipiv[J] = 1-based global pivot row of column J, s_info = first column
// with an exactly zero pivot (-1: none).
template <typename T>
__device__ __forceinline__ void panel_finish(int w, int t, int jb, int64_t j0, const int32_t* s_ipiv,
                                             const int32_t* s_info, int32_t* ipiv, int32_t* info) {
  if (w != 0) return;
  __syncthreads();
  if (t < jb) ipiv[j0 + t] = s_ipiv[t];
  if (t == 0 && *s_info >= 0 && *info == 0) *info = (int32_t)(j0 + *s_info + 1);
}

template <typename T, int IB, int J>
__device__ __forceinline__ void panel_step(T (&a)[IB], const PanelCtx<T>& P) {
  if constexpr (J < IB) {
    if (J < P.jb) {
      const int par = J & 1;
      const int t = P.t, lane = t & 63;
      // 1. candidate of this workgroup
      T bv;
      int32_t bi;
      wg_argmax<T, IB>(a, J, P.row, t, P.red_v, P.red_i, bv, bi);
      T* cand = P.rows + ((size_t)par * (P.nwg + 1) + P.w) * IB;
      T* rowj = P.rows + ((size_t)par * (P.nwg + 1) + P.nwg) * IB;
      if (t == 0) {
        P.ps->val[par][P.w] = bv;
        P.ps->idx[par][P.w] = bi;
      }
      if (P.row >= 0 && P.row == bi) {
#pragma unroll
        for (int c = 0; c < IB; ++c) cand[c] = a[c];
      }
      if (P.row == J) {
#pragma unroll
        for (int c = 0; c < IB; ++c) rowj[c] = a[c];
      }
      grid_barrier(P.ctr, P.base + (uint64_t)(J + 1) * P.nwg, P.err);
      // 2. global pivot, reduced redundantly by every wave: four slots per
      // lane, all loads issued before any is used
      T gv = T(-1);
      int32_t gi = -1;
      for (int q0 = 0; q0 < P.nwg; q0 += 256) {
        T cv[4];
        int32_t ci[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const int q = q0 + lane + 64 * k;
          ci[k] = -1;
          if (q < P.nwg) {
            cv[k] = P.ps->val[par][q];
            ci[k] = P.ps->idx[par][q];
          }
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) pick(gv, gi, cv[k], ci[k]);
      }
      wave_argmax(gv, gi);
      const int gw = gi >= 0 ? gi / PWG : 0;  // the workgroup holding the pivot row
      const T* prow = P.rows + ((size_t)par * (P.nwg + 1) + gw) * IB;
      if (t < IB) P.s_piv[t] = prow[t];
      __syncthreads();
      const T pv = P.s_piv[J];
      const bool zero = (pv == T(0));
      if (t == 0) {
        P.s_ipiv[J] = (int32_t)(P.j0 + (zero ? J : gi) + 1);
        if (zero && *P.s_info < 0) *P.s_info = J;
      }
      const T inv = zero ? T(0) : T(1) / pv;
      // 3. interchange + rank-1 update of this thread's row
      if (P.row >= J) {
        if (!zero && P.row == J) {
#pragma unroll
          for (int c = 0; c < IB; ++c) a[c] = P.s_piv[c];
        } else {
          if (!zero && P.row == gi) {
#pragma unroll
            for (int c = 0; c < IB; ++c) a[c] = rowj[c];
          }
          const T l = a[J] * inv;
          a[J] = l;
#pragma unroll
          for (int c = J + 1; c < IB; ++c) a[c] -= l * P.s_piv[c];
        }
      }
    }
    panel_step<T, IB, J + 1>(a, P);
  }
}

template <typename T, int IB>
__global__ __launch_bounds__(PWG) void k_panel(T* __restrict__ A, int64_t lda, int64_t n, int64_t j0, int jb,
                                               int32_t* __restrict__ ipiv, int32_t* __restrict__ info,
                                               int32_t* __restrict__ err, PanelScratch<T>* __restrict__ ps,
                                               T* __restrict__ rows /* [2][nwg+1][IB] */, uint64_t* ctr,
                                               uint64_t base) {
  if (panel_skip(err, j0, jb, ipiv)) return;
  const int nwg = gridDim.x, w = blockIdx.x, t = threadIdx.x;
  const int64_t m = n - j0;
  const int64_t r = (int64_t)w * PWG + t;
  __shared__ T red_v[PWG / 64];
  __shared__ int32_t red_i[PWG / 64];
  __shared__ T s_piv[IB];
  __shared__ int32_t s_ipiv[IB], s_info;
  if (t == 0) s_info = -1;  // read by thread 0 only until the closing barrier
  T a[IB];
  const bool ok = r < m;
#pragma unroll
  for (int c = 0; c < IB; ++c) a[c] = (ok && c < jb) ? A[(j0 + r) + (j0 + c) * lda] : T(0);
  PanelCtx<T> P{nwg, w, t, jb, j0, ok ? (int32_t)r : -1, ps, rows, err, ctr, base,
                red_v, red_i, s_piv, s_ipiv, &s_info};
  panel_step<T, IB, 0>(a, P);
  if (ok)
#pragma unroll
    for (int c = 0; c < IB; ++c)
      if (c < jb) A[(j0 + r) + (j0 + c) * lda] = a[c];
  panel_finish<T>(w, t, jb, j0, s_ipiv, &s_info, ipiv, info);
}

// ---------------------------------------------------------------------
// The same panel factorisation without a grid barrier.  Per column every
// workgroup publishes its candidate row (the full row of jb values) and
// its index as 64-bit words whose high half is the column's sequence number
// and low half 32 bits of payload; workgroup 0 also publishes the current
// row J (the row the pivot row is exchanged with).  Every workgroup then
// polls, with agent-scope atomic loads, the column-J value and the index of
// every candidate until each carries the current sequence number (staged in
// LDS), picks the pivot, and reads the pivot row's words (and, in the
// pivot's workgroup, row J's) the same way.  A word is single-copy atomic,
// so it validates itself: no fences and no arrival counter, two short
// round trips per column.  The last four waves publish and the first ones
// poll, so that no polling wave waits on its own stores.  Two buffers
// alternate by column parity: a workgroup writes column J + 2 only after
// reading every word of column J + 1 it needs, which each workgroup
// publishes only after it has read what it needs of column J.  Sequence
// numbers are the global column + 1, unique within one factorisation (the
// buffer is cleared before it), so stale words are never taken for current
// ones.  Polls are bounded like the grid barrier: past ~2^22 polls *err is
// raised and the panel completes with garbage (reported by the host), never
// a hang.  Measured on MI355X the sc1 polls cost more than the barrier's
// fences, so this variant is selectable (HFPX_KNOB_PANEL = 2), not the
// default.
// ---------------------------------------------------------------------
constexpr int PX_MAX_WG = 256;  // LDS staging bound (one 1024-thread workgroup per CU)

template <typename T, int IB>
struct PxFmt {
  static constexpr int W = (int)(sizeof(T) / 4);  // payload words per value
  static constexpr int WORDS = IB * W + 1;        // row words + index word
};

template <typename T>
__device__ __forceinline__ T px_val(const uint32_t* w, int c) {
  if constexpr (sizeof(T) == 4)
    return __uint_as_float(w[c]);
  else
    return __hiloint2double((int)w[2 * c + 1], (int)w[2 * c]);
}

template <typename T>
__device__ __forceinline__ void px_put(uint32_t* w, int c, T v) {
  if constexpr (sizeof(T) == 4) {
    w[c] = __float_as_uint(v);
  } else {
    const uint64_t u = (uint64_t)__double_as_longlong(v);
    w[2 * c] = (uint32_t)u;
    w[2 * c + 1] = (uint32_t)(u >> 32);
  }
}

// Payload of the tagged word at p once it carries sequence number seq
// (bounded wait; *err raised on time-out).
__device__ __forceinline__ uint32_t px_get(uint64_t* p, uint32_t seq, int32_t* err) {
  uint64_t v = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  uint32_t polls = 0;
  while ((uint32_t)(v >> 32) != seq) {
    __builtin_amdgcn_s_sleep(1);
    ++polls;
    if ((polls & 255u) == 0 && lu_failed(err)) break;
    if (polls > (1u << 22)) {
      __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      break;
    }
    v = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  return (uint32_t)v;
}

template <typename T, int IB>
struct PxCtx {
  int nwg, w, t, jb;
  int64_t j0;
  int32_t row;
  uint64_t* xw;  // [2][nwg + 1][WORDS]
  int32_t* err;
  T* red_v;
  int32_t* red_i;
  uint32_t* s_cand;  // IB W words: this workgroup's candidate row
  uint32_t* s_rowj;  // IB W words: row J (workgroup 0)
  uint32_t* s_all;   // (nwg + 1) WORDS staged payloads
  int32_t* s_ipiv;
  int32_t* s_info;
};

constexpr int PX_PUB = PWG - 128;   // first publishing thread (waves 14, 15)
constexpr int PX_ROWJ = PWG - 256;  // first thread publishing row J (waves 12, 13)

template <typename T, int IB, int J>
__device__ __forceinline__ void panel_step_x(T (&a)[IB], const PxCtx<T, IB>& P) {
  if constexpr (J < IB) {
    if (J < P.jb) {
      typedef PxFmt<T, IB> F;
      constexpr int RW = IB * F::W;  // row words
      static_assert(RW + 1 <= 128 && RW <= 128, "publisher waves");
      const int t = P.t, lane = t & 63;
      const uint32_t seq = (uint32_t)(P.j0 + J + 1);
      uint64_t* buf = P.xw + (size_t)(J & 1) * (P.nwg + 1) * F::WORDS;
      // 1. candidate of this workgroup
      T bv;
      int32_t bi;
      wg_argmax<T, IB>(a, J, P.row, t, P.red_v, P.red_i, bv, bi);
      if (P.row >= 0 && P.row == bi) {
#pragma unroll
        for (int c = 0; c < IB; ++c) px_put<T>(P.s_cand, c, a[c]);
      }
      if (P.row == J) {
#pragma unroll
        for (int c = 0; c < IB; ++c) px_put<T>(P.s_rowj, c, a[c]);
      }
      __syncthreads();
      // 2. publish
      const uint64_t tag = (uint64_t)seq << 32;
      if (t >= PX_PUB) {
        const int k = t - PX_PUB;
        if (k < RW)
          __hip_atomic_store(buf + (size_t)P.w * F::WORDS + k, tag | P.s_cand[k], __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
        else if (k == RW)
          __hip_atomic_store(buf + (size_t)P.w * F::WORDS + RW, tag | (uint32_t)bi, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
      } else if (P.w == 0 && t >= PX_ROWJ && t < PX_ROWJ + RW) {
        const int k = t - PX_ROWJ;
        __hip_atomic_store(buf + (size_t)P.nwg * F::WORDS + k, tag | P.s_rowj[k], __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
      }
      // 3a. column J and the index of every candidate, one word per thread
      constexpr int PA = F::W + 1;  // words per slot in this phase
      if (t < P.nwg * PA) {
        const int q = t / PA, k = t % PA;
        const int wd = k < F::W ? J * F::W + k : RW;
        P.s_all[q * F::WORDS + wd] = px_get(buf + (size_t)q * F::WORDS + wd, seq, P.err);
      }
      __syncthreads();
      // 4. global pivot, reduced redundantly by every wave from LDS
      T gv = T(-1);
      int32_t gi = -1;
#pragma unroll
      for (int k = 0; k < PX_MAX_WG / 64; ++k) {
        const int q = lane + 64 * k;
        if (q < P.nwg) {
          const uint32_t* sq = P.s_all + q * F::WORDS;
          pick(gv, gi, (T)fabs(px_val<T>(sq, J)), (int32_t)sq[RW]);
        }
      }
      wave_argmax(gv, gi);
      const int gw = gi >= 0 ? gi / PWG : 0;  // the workgroup holding the pivot row
      // 3b. the pivot row (held locally by its own workgroup) and, in that
      // workgroup only, row J (held locally by workgroup 0)
      uint32_t* prow = P.s_all + gw * F::WORDS;
      const uint32_t* rowj = P.w == 0 ? P.s_rowj : P.s_all + P.nwg * F::WORDS;
      if (t < RW) {
        if (gw == P.w)
          prow[t] = P.s_cand[t];
        else if (t / F::W != J)
          prow[t] = px_get(buf + (size_t)gw * F::WORDS + t, seq, P.err);
      } else if (t >= PWG / 2 && t < PWG / 2 + RW && gw == P.w && P.w != 0) {
        const int k = t - PWG / 2;
        P.s_all[P.nwg * F::WORDS + k] = px_get(buf + (size_t)P.nwg * F::WORDS + k, seq, P.err);
      }
      __syncthreads();
      const T pv = px_val<T>(prow, J);
      const bool zero = (pv == T(0));
      if (t == 0) {
        P.s_ipiv[J] = (int32_t)(P.j0 + (zero ? J : gi) + 1);
        if (zero && *P.s_info < 0) *P.s_info = J;
      }
      const T inv = zero ? T(0) : T(1) / pv;
      // 5. interchange + rank-1 update of this thread's row
      if (P.row >= J) {
        if (!zero && P.row == J) {
#pragma unroll
          for (int c = 0; c < IB; ++c) a[c] = px_val<T>(prow, c);
        } else {
          if (!zero && P.row == gi) {
#pragma unroll
            for (int c = 0; c < IB; ++c) a[c] = px_val<T>(rowj, c);
          }
          const T l = a[J] * inv;
          a[J] = l;
#pragma unroll
          for (int c = J + 1; c < IB; ++c) a[c] -= l * px_val<T>(prow, c);
        }
      }
    }
    panel_step_x<T, IB, J + 1>(a, P);
  }
}

template <typename T, int IB>
__global__ __launch_bounds__(PWG) void k_panel_x(T* __restrict__ A, int64_t lda, int64_t n, int64_t j0, int jb,
                                                 int32_t* __restrict__ ipiv, int32_t* __restrict__ info,
                                                 int32_t* __restrict__ err, uint64_t* __restrict__ xw) {
  typedef PxFmt<T, IB> F;
  static_assert(PX_MAX_WG * (F::W + 1) <= PX_ROWJ, "polling threads overlap the publishing waves");
  if (panel_skip(err, j0, jb, ipiv)) return;
  const int nwg = gridDim.x, w = blockIdx.x, t = threadIdx.x;
  const int64_t m = n - j0;
  const int64_t r = (int64_t)w * PWG + t;
  __shared__ T red_v[PWG / 64];
  __shared__ int32_t red_i[PWG / 64];
  __shared__ uint32_t s_cand[IB * F::W], s_rowj[IB * F::W];
  __shared__ uint32_t s_all[(PX_MAX_WG + 1) * F::WORDS];
  __shared__ int32_t s_ipiv[IB], s_info;
  if (t == 0) s_info = -1;
  T a[IB];
  const bool ok = r < m;
#pragma unroll
  for (int c = 0; c < IB; ++c) a[c] = (ok && c < jb) ? A[(j0 + r) + (j0 + c) * lda] : T(0);
  PxCtx<T, IB> P{nwg, w, t, jb, j0, ok ? (int32_t)r : -1, xw, err, red_v, red_i, s_cand, s_rowj, s_all,
                 s_ipiv, &s_info};
  panel_step_x<T, IB, 0>(a, P);
  if (ok)
#pragma unroll
    for (int c = 0; c < IB; ++c)
      if (c < jb) A[(j0 + r) + (j0 + c) * lda] = a[c];
  panel_finish<T>(w, t, jb, j0, s_ipiv, &s_info, ipiv, info);
}

// ---------------------------------------------------------------------
// Row interchanges.  A block of up to SWAP_BLK sequential interchanges
// ipiv[s0 .. s0+nb) is one permutation of at most 2 SWAP_BLK rows; it is
// computed once (k_perm_pairs, one workgroup per block) as pairs
// new A[dst] = old A[src], then applied to many columns in parallel, one
// wave per column, all loads before any store (k_apply_pairs), or block
// after block within one workgroup per column (k_apply_pairs_seq).
// ---------------------------------------------------------------------
constexpr int SWAP_BLK = 256;
constexpr int PAIR_MAX = 2 * SWAP_BLK;

struct PermPairs {
  int32_t count;
  int32_t pad[63];
  int32_t dst[PAIR_MAX];
  int32_t src[PAIR_MAX];
};

// The pairs of the interchanges ipiv[s0, s0 + nb) (nb <= SWAP_BLK) into
// dst / src / *count, by the whole workgroup (256 threads); the tables are
// the caller's LDS.
struct PairLds {
  int32_t piv[SWAP_BLK], borig[SWAP_BLK];
  int32_t hkey[1024], horig[1024];  // rows outside the block (open addressing)
  int32_t olist[SWAP_BLK];
};

__device__ __forceinline__ void build_pairs(const int32_t* __restrict__ ipiv, int64_t s0, int nb, int64_t n,
                                            PairLds& L, int32_t* dst, int32_t* src, int32_t* count) {
  int32_t* piv = L.piv;
  int32_t* borig = L.borig;
  int32_t* hkey = L.hkey;
  int32_t* horig = L.horig;
  int32_t* olist = L.olist;
  const int t = threadIdx.x;
  if (t < nb) {
    piv[t] = ipiv[s0 + t] - 1;
    borig[t] = (int32_t)(s0 + t);
  }
  for (int i = t; i < 1024; i += 256) hkey[i] = -1;
  __syncthreads();
  if (t != 0) return;
  // borig[i] / horig[h]: original row of the value now held by that row
  int nout = 0;
  for (int i = 0; i < nb; ++i) {
    const int32_t r1 = (int32_t)(s0 + i), r2 = piv[i];
    if (r2 == r1 || r2 < 0 || r2 >= n) continue;  // no-op or invalid entry
    int32_t* slot;
    if (r2 >= s0 && r2 < s0 + nb) {
      slot = &borig[r2 - s0];
    } else {
      uint32_t h = ((uint32_t)r2 * 2654435761u) >> 22;
      while (hkey[h] != -1 && hkey[h] != r2) h = (h + 1) & 1023;
      if (hkey[h] == -1) {
        hkey[h] = r2;
        horig[h] = r2;
        olist[nout++] = (int32_t)h;
      }
      slot = &horig[h];
    }
    const int32_t tmp = borig[i];
    borig[i] = *slot;
    *slot = tmp;
  }
  int cnt = 0;
  for (int i = 0; i < nb; ++i)
    if (borig[i] != (int32_t)(s0 + i)) {
      dst[cnt] = (int32_t)(s0 + i);
      src[cnt] = borig[i];
      ++cnt;
    }
  for (int q = 0; q < nout; ++q) {
    const int h = olist[q];
    if (horig[h] != hkey[h]) {
      dst[cnt] = hkey[h];
      src[cnt] = horig[h];
      ++cnt;
    }
  }
  *count = cnt;
}

__global__ __launch_bounds__(256) void k_perm_pairs(const int32_t* __restrict__ ipiv, int64_t k0, int64_t k1,
                                                    int64_t n, PermPairs* __restrict__ out) {
  __shared__ PairLds L;
  const int64_t s0 = k0 + (int64_t)blockIdx.x * SWAP_BLK;
  const int nb = (int)std::min<int64_t>(SWAP_BLK, k1 - s0);
  PermPairs* pp = out + blockIdx.x;
  build_pairs(ipiv, s0, nb, n, L, pp->dst, pp->src, &pp->count);
}

// Small interchange blocks (k1 - k0 <= SWAP_SMALL) in one launch: every
// workgroup builds the pairs in its LDS (thread 0 walks at most SWAP_SMALL
// interchanges) and applies them to its four columns as k_apply_pairs does.
constexpr int SWAP_SMALL = 64;

template <typename T>
__global__ __launch_bounds__(256) void k_swap_small(T* __restrict__ A, int64_t lda, const int32_t* __restrict__ ipiv,
                                                    int64_t k0, int64_t k1, int64_t n, int64_t c0, int64_t c1,
                                                    int64_t c2, int64_t c3) {
  __shared__ PairLds L;
  __shared__ int32_t sdst[2 * SWAP_SMALL], ssrc[2 * SWAP_SMALL], scnt;
  build_pairs(ipiv, k0, (int)(k1 - k0), n, L, sdst, ssrc, &scnt);
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int64_t ci = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int64_t n1 = c1 - c0;
  if (ci >= n1 + (c3 - c2)) return;
  const int64_t col = ci < n1 ? c0 + ci : c2 + (ci - n1);
  const int cnt = scnt;
  T v[2 * SWAP_SMALL / 64];
  int32_t d[2 * SWAP_SMALL / 64];
#pragma unroll
  for (int q = 0; q < 2 * SWAP_SMALL / 64; ++q) {
    const int i = lane + 64 * q;
    if (i < cnt) {
      d[q] = sdst[i];
      v[q] = A[ssrc[i] + col * lda];
    }
  }
  // every load of this wave has returned before any store (one wave per
  // column: the wave's own loads and stores are the only ones on it)
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (int q = 0; q < 2 * SWAP_SMALL / 64; ++q) {
    const int i = lane + 64 * q;
    if (i < cnt) A[d[q] + col * lda] = v[q];
  }
}

// Columns [c0, c1) and [c2, c3), one wave each.
template <typename T>
__global__ __launch_bounds__(256) void k_apply_pairs(T* __restrict__ A, int64_t lda,
                                                     const PermPairs* __restrict__ pp, int64_t c0, int64_t c1,
                                                     int64_t c2, int64_t c3) {
  const int lane = threadIdx.x & 63;
  const int64_t ci = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int64_t n1 = c1 - c0;
  const bool active = ci < n1 + (c3 - c2);
  const int64_t col = ci < n1 ? c0 + ci : c2 + (ci - n1);
  const int cnt = pp->count;
  T v[PAIR_MAX / 64];
  int32_t d[PAIR_MAX / 64];
#pragma unroll
  for (int q = 0; q < PAIR_MAX / 64; ++q) {
    const int i = lane + 64 * q;
    if (active && i < cnt) {
      d[q] = pp->dst[i];
      v[q] = A[pp->src[i] + col * lda];
    }
  }
  __syncthreads();  // every load of this wave has returned before any store
#pragma unroll
  for (int q = 0; q < PAIR_MAX / 64; ++q) {
    const int i = lane + 64 * q;
    if (active && i < cnt) A[d[q] + col * lda] = v[q];
  }
}

// All nblk pair blocks in order on column blockIdx.x of B (one workgroup).
template <typename T>
__global__ __launch_bounds__(256) void k_apply_pairs_seq(T* __restrict__ B, int64_t ldb,
                                                         const PermPairs* __restrict__ pp, int nblk) {
  T* col = B + (int64_t)blockIdx.x * ldb;
  const int t = threadIdx.x;
  for (int b = 0; b < nblk; ++b) {
    const int cnt = pp[b].count;
    T v0 = T(0), v1 = T(0);
    int32_t d0 = -1, d1 = -1;
    if (t < cnt) {
      d0 = pp[b].dst[t];
      v0 = col[pp[b].src[t]];
    }
    if (t + 256 < cnt) {
      d1 = pp[b].dst[t + 256];
      v1 = col[pp[b].src[t + 256]];
    }
    __syncthreads();
    if (d0 >= 0) col[d0] = v0;
    if (d1 >= 0) col[d1] = v1;
    __syncthreads();
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
  {
    // all loads of the triangle in flight before the first LDS store
    constexpr int LQ = LB * LB / 256;
    T lv[LQ];
#pragma unroll
    for (int q = 0; q < LQ; ++q) {
      const int e = threadIdx.x + 256 * q, i = e % LB, j = e / LB;
      lv[q] = (i < kb && j < kb) ? Lm[i + (int64_t)j * ldl] : T(0);
    }
#pragma unroll
    for (int q = 0; q < LQ; ++q) {
      const int e = threadIdx.x + 256 * q;
      sL[e % LB][e / LB] = lv[q];
    }
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

// FP32 leaf with B staged through LDS: a workgroup of two waves owns
// TRSM_F32_TC = 128 columns; each wave loads and stores whole 64-row column
// segments (256 contiguous bytes) instead of one lane per column touching 64
// cache lines per load.  The unit-lower solve runs by columns: once x_j is
// final, the updates x_i -= L_ij x_j of the rows below it are independent,
// so the dependent chain is 63 steps long instead of 2,016 (the row form's
// running sum), with each x_i still reduced in the same order j = 0, 1, ...
// and therefore to the same bits.  The triangle sits in LDS transposed
// (column j contiguous, rows 16-byte aligned) so the reads of one column
// merge.  The upper solve keeps the row form of k_trsm_leaf.  50 KB of LDS:
// three workgroups per CU.  The unit-lower leaves of the factorisation take
// the lane form below (k_trsm_leaf_lane), in FP32 and FP64.
constexpr int TRSM_F32_TC = 128;
constexpr int64_t TRSM_LANE_NARROW = 1024;  // columns up to which the lane leaf takes one column per wave
constexpr int64_t TRSM_LANE_WIDE = 24576;   // FP32 columns above which the staged leaf is faster

template <bool LOWER>
__global__ __launch_bounds__(TRSM_F32_TC) void k_trsm_leaf_f32(const float* __restrict__ Lm, int64_t ldl,
                                                               float* __restrict__ B, int64_t ldb, int kb,
                                                               int64_t ncols) {
  constexpr int TC = TRSM_F32_TC;
  constexpr int NW = TC / 64;
  // LOWER: sM[j][i] = L(i, j); upper: sM[i][j] = U(i, j)
  __shared__ __align__(16) float sM[LB][LB + 4];
  __shared__ float sB[TC][LB + 1];
  constexpr int LQ = LB * LB / TC, BQ = TC / NW;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t c0 = (int64_t)blockIdx.x * TC;
  const int nc = (int)((ncols - c0) < TC ? (ncols - c0) : TC);
  // every global load of the triangle and of the tile is issued before the
  // first is used: one memory latency, not one per loop trip
  {
    float lv[LQ];
#pragma unroll
    for (int q = 0; q < LQ; ++q) {
      const int e = tid + TC * q, i = e % LB, j = e / LB;
      lv[q] = (i < kb && j < kb) ? Lm[i + (int64_t)j * ldl] : 0.0f;
    }
    float bv[BQ];
#pragma unroll
    for (int q = 0; q < BQ; ++q) {
      const int c = wave + NW * q;
      bv[q] = (c < nc && lane < kb) ? B[(c0 + c) * ldb + lane] : 0.0f;
    }
#pragma unroll
    for (int q = 0; q < LQ; ++q) {
      const int e = tid + TC * q, i = e % LB, j = e / LB;
      if (LOWER)
        sM[j][i] = lv[q];
      else
        sM[i][j] = lv[q];
    }
#pragma unroll
    for (int q = 0; q < BQ; ++q) sB[wave + NW * q][lane] = bv[q];
  }
  __syncthreads();
  if (tid < nc) {
    float x[LB];
#pragma unroll
    for (int i = 0; i < LB; ++i) x[i] = sB[tid][i];
    if (LOWER) {
#pragma unroll
      for (int j = 0; j < LB - 1; ++j) {
        // the guard keeps the scheduler from hoisting every LDS read of the
        // unrolled loop nest at once (which spills)
        if (j < kb) {
          const float xj = x[j];
#pragma unroll
          for (int i = j + 1; i < LB; ++i) x[i] -= sM[j][i] * xj;
        }
      }
    } else {
#pragma unroll
      for (int i = LB - 1; i >= 0; --i) {
        if (i < kb) {
          float acc = x[i];
#pragma unroll
          for (int j = i + 1; j < LB; ++j) acc -= sM[i][j] * x[j];
          x[i] = acc / sM[i][i];
        }
      }
    }
#pragma unroll
    for (int i = 0; i < LB; ++i) sB[tid][i] = x[i];
  }
  __syncthreads();
  for (int c = wave; c < nc; c += NW)
    if (lane < kb) B[(c0 + c) * ldb + lane] = sB[c][lane];
}

// Unit-lower leaf in lane form: a wave owns CPW columns of B and lane i row
// i, holding L(i, 0 .. i-1) in registers; step j broadcasts x_j with
// v_readlane (scalar registers) and every lane below j applies
// x_i -= L_ij x_j.  No LDS and no barrier: the dependent chain is 63
// readlane + FMA steps, so a small solve (the leaves inside an outer block,
// at most a few hundred columns) is bound by its loads rather than by the
// per-step LDS latency of the staged form.  Each x_i is reduced in the same
// order j = 0, 1, ... with fused multiply-adds, as in the staged FP32 form.
template <typename T>
__device__ __forceinline__ T lane_bcast(T v, int j) {
  if constexpr (sizeof(T) == 4) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
  } else {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffff), j);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), j);
    return __hiloint2double(hi, lo);
  }
}

template <typename T, int CPW>
__global__ __launch_bounds__(256) void k_trsm_leaf_lane(const T* __restrict__ Lm, int64_t ldl, T* __restrict__ B,
                                                        int64_t ldb, int kb, int64_t ncols) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t c0 = ((int64_t)blockIdx.x * 4 + wave) * CPW;
  if (c0 >= ncols) return;  // whole waves only; no barrier below
  const bool rin = lane < kb;
  // unconditional loads (a masked lane reads L(0, 0)), then the mask: the 64
  // loads issue back to back without a branch each
  const uint32_t ld = (uint32_t)ldl;
  T l[LB];
#pragma unroll
  for (int j = 0; j < LB; ++j) {
    const bool ok = j < lane && rin;
    const T v = Lm[ok ? (uint32_t)lane + (uint32_t)j * ld : 0u];
    l[j] = ok ? v : T(0);
  }
  T x[CPW];
#pragma unroll
  for (int c = 0; c < CPW; ++c) x[c] = (rin && c0 + c < ncols) ? B[lane + (c0 + c) * ldb] : T(0);
#pragma unroll
  for (int j = 0; j < LB - 1; ++j) {
    if (j + 1 >= kb) break;  // uniform
#pragma unroll
    for (int c = 0; c < CPW; ++c) {
      const T u = std::fma(-l[j], lane_bcast(x[c], j), x[c]);
      x[c] = lane > j ? u : x[c];
    }
  }
#pragma unroll
  for (int c = 0; c < CPW; ++c)
    if (rin && c0 + c < ncols) B[lane + (c0 + c) * ldb] = x[c];
}

// ---------------------------------------------------------------------
// host drivers
// ---------------------------------------------------------------------
// Pairs for the interchanges ipiv[k0, k1), one PermPairs per SWAP_BLK.
static int perm_pairs(hfpx_ctx* ctx, const int32_t* ipiv, int64_t k0, int64_t k1, int64_t n, PermPairs* pp) {
  if (k1 <= k0) return HFPX_OK;
  const unsigned nblk = (unsigned)((k1 - k0 + SWAP_BLK - 1) / SWAP_BLK);
  const int tok = prof_begin(ctx, HFPX_KC_LASWP);
  hipLaunchKernelGGL(k_perm_pairs, dim3(nblk), dim3(256), 0, ctx->stream, ipiv, k0, k1, n, pp);
  prof_end(ctx, tok, HFPX_KC_LASWP, 0.0, 0.0);
  HFPX_TRY_HIP(ctx, hipGetLastError());
  return HFPX_OK;
}

// Apply one PermPairs block to columns [c0, c1) and [c2, c3) of A.
template <typename T>
static int apply_pairs(hfpx_ctx* ctx, T* A, int64_t lda, const PermPairs* pp, int64_t c0, int64_t c1, int64_t c2,
                       int64_t c3) {
  const int64_t ncol = std::max<int64_t>(0, c1 - c0) + std::max<int64_t>(0, c3 - c2);
  if (ncol <= 0) return HFPX_OK;
  if (c1 < c0) c1 = c0;
  if (c3 < c2) c3 = c2;
  const int tok = prof_begin(ctx, HFPX_KC_LASWP);
  hipLaunchKernelGGL(k_apply_pairs<T>, dim3((unsigned)((ncol + 3) / 4)), dim3(256), 0, ctx->stream, A, lda, pp, c0,
                     c1, c2, c3);
  prof_end(ctx, tok, HFPX_KC_LASWP, 0.0, 2.0 * sizeof(T) * ncol * 2 * SWAP_BLK);
  HFPX_TRY_HIP(ctx, hipGetLastError());
  return HFPX_OK;
}

template <typename T, bool LOWER>
static int trsm_leaf(hfpx_ctx* ctx, const T* L, int64_t ldl, T* B, int64_t ldb, int kb, int64_t ncols) {
  const int tok = prof_begin(ctx, HFPX_KC_TRSM);
  if (LOWER && sizeof(T) == 4 && ncols > TRSM_LANE_WIDE) {
    // very wide FP32 solves: the lane form is VALU-bound there (three
    // instructions per column and step for 64 rows), the staged form's
    // thread per column does one FMA per entry (same bits)
    hipLaunchKernelGGL((k_trsm_leaf_f32<true>), dim3((unsigned)((ncols + TRSM_F32_TC - 1) / TRSM_F32_TC)),
                       dim3(TRSM_F32_TC), 0, ctx->stream, (const float*)L, ldl, (float*)B, ldb, kb, ncols);
  } else if constexpr (LOWER) {
    // lane form: one column per wave for the small solves inside an outer
    // block, eight per wave (L loaded once per eight columns) for the wide
    // ones
    if (ncols <= TRSM_LANE_NARROW)
      hipLaunchKernelGGL((k_trsm_leaf_lane<T, 1>), dim3((unsigned)((ncols + 3) / 4)), dim3(256), 0, ctx->stream,
                         L, ldl, B, ldb, kb, ncols);
    else
      hipLaunchKernelGGL((k_trsm_leaf_lane<T, 8>), dim3((unsigned)((ncols + 31) / 32)), dim3(256), 0,
                         ctx->stream, L, ldl, B, ldb, kb, ncols);
  } else if constexpr (sizeof(T) == 4)
    hipLaunchKernelGGL((k_trsm_leaf_f32<LOWER>), dim3((unsigned)((ncols + TRSM_F32_TC - 1) / TRSM_F32_TC)),
                       dim3(TRSM_F32_TC), 0, ctx->stream, L, ldl, B, ldb, kb, ncols);
  else
    hipLaunchKernelGGL((k_trsm_leaf<T, LOWER>), dim3((unsigned)((ncols + 255) / 256)), dim3(256), 0, ctx->stream,
                       L, ldl, B, ldb, kb, ncols);
  prof_end(ctx, tok, HFPX_KC_TRSM, (double)kb * kb * ncols, 0.0);
  HFPX_TRY_HIP(ctx, hipGetLastError());
  return HFPX_OK;
}

// B (kb x ncols) := L^{-1} B, L unit lower kb x kb.  Recursive halving down
// to leaves of LB rows, so that the MFMA updates in between have M = kb/2,
// kb/4, ... rather than LB.
template <typename T>
static int trsm_lunit(hfpx_ctx* ctx, const T* L, int64_t ldl, T* B, int64_t ldb, int kb, int64_t ncols) {
  if (ncols <= 0 || kb <= 0) return HFPX_OK;
  if (kb <= LB) return trsm_leaf<T, true>(ctx, L, ldl, B, ldb, kb, ncols);
  const int k1 = ((kb / 2 + LB - 1) / LB) * LB;
  int rc;
  if ((rc = trsm_lunit<T>(ctx, L, ldl, B, ldb, k1, ncols))) return rc;
  if ((rc = gemm_sub<T>(ctx, B + k1, ldb, L + k1, ldl, B, ldb, kb - k1, ncols, k1))) return rc;
  return trsm_lunit<T>(ctx, L + k1 + (int64_t)k1 * ldl, ldl, B + k1, ldb, kb - k1, ncols);
}

// B (kb x ncols) := U^{-1} B, U upper (non-unit) kb x kb, recursive halving
// (bottom half first).
template <typename T>
static int trsm_upper(hfpx_ctx* ctx, const T* U, int64_t ldu, T* B, int64_t ldb, int kb, int64_t ncols) {
  if (ncols <= 0 || kb <= 0) return HFPX_OK;
  if (kb <= LB) return trsm_leaf<T, false>(ctx, U, ldu, B, ldb, kb, ncols);
  const int k1 = ((kb / 2 + LB - 1) / LB) * LB;
  int rc;
  if ((rc = trsm_upper<T>(ctx, U + k1 + (int64_t)k1 * ldu, ldu, B + k1, ldb, kb - k1, ncols))) return rc;
  if ((rc = gemm_sub<T>(ctx, B, ldb, U + (int64_t)k1 * ldu, ldu, B + k1, ldb, k1, ncols, kb - k1))) return rc;
  return trsm_upper<T>(ctx, U, ldu, B, ldb, k1, ncols);
}

// Grow-only workspace of the LU routines (separate from ctx->work, which
// the mixed-precision solve holds while it calls getrs).
static int ensure_lu_work(hfpx_ctx* ctx, size_t bytes) {
  if (ctx->lu_work_bytes >= bytes) return HFPX_OK;
  if (ctx->lu_work) {
    HFPX_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    HFPX_TRY_HIP(ctx, hipFree(ctx->lu_work));
    ctx->lu_work = nullptr;
    ctx->lu_work_bytes = 0;
  }
  HFPX_TRY_HIP(ctx, hipMalloc(&ctx->lu_work, bytes));
  ctx->lu_work_bytes = bytes;
  return HFPX_OK;
}

template <typename T, int IB>
static int panel(hfpx_ctx* ctx, T* A, int64_t lda, int64_t n, int64_t j0, int jb, int32_t* ipiv, int32_t* dflags,
                 PanelScratch<T>* ps, T* rows, uint64_t* xw, int max_wg, bool coop_ok) {
  const int64_t m = n - j0;
  const int64_t nwg64 = (m + PWG - 1) / PWG;
  if (nwg64 > max_wg)
    return set_error(ctx, HFPX_ERR_INVALID, "getrf: n = %lld too large for the panel kernel (at most %lld rows)",
                     (long long)n, (long long)max_wg * PWG);
  int nwg = (int)nwg64;
  int32_t* info = dflags;
  int32_t* err = dflags + 1;
  uint64_t* ctr = ctx->sync;
  uint64_t base = ctx->sync_base;
  // the grid-barrier panel unless the exchange panel is asked for (and the
  // grid fits its LDS staging); on MI355X the barrier panel is the faster
  // (72,000 rows FP32: 396 against 559 ms per factorisation)
  const bool xchg = ctx->knob[HFPX_KNOB_PANEL] == 2 && nwg <= PX_MAX_WG;
  void* args[] = {&A, &lda, &n, &j0, &jb, &ipiv, &info, &err, &ps, &rows, &ctr, &base};
  void* xargs[] = {&A, &lda, &n, &j0, &jb, &ipiv, &info, &err, &xw};
  const void* kern = xchg ? (const void*)k_panel_x<T, IB> : (const void*)k_panel<T, IB>;
  const int tok = prof_begin(ctx, HFPX_KC_PANEL);
  // A cooperative launch guarantees that the workgroups are co-resident.
  // The runtime refuses it while another process holds the device's
  // cooperative capability; the grid (at most one workgroup per CU) is then
  // launched normally, and the barrier's bounded wait turns a workgroup that
  // never becomes resident into an error instead of a hang.
  hipError_t e = hipErrorIllegalState;
  const bool coop = coop_ok && !ctx->no_coop;
  if (coop) e = hipLaunchCooperativeKernel(kern, dim3(nwg), dim3(PWG), xchg ? xargs : args, 0, ctx->stream);
  if (!coop || e == hipErrorIllegalState) {
    if (coop) {
      (void)hipGetLastError();
      ctx->no_coop = true;
    }
    if (xchg)
      hipLaunchKernelGGL((k_panel_x<T, IB>), dim3(nwg), dim3(PWG), 0, ctx->stream, A, lda, n, j0, jb, ipiv, info,
                         err, xw);
    else
      hipLaunchKernelGGL((k_panel<T, IB>), dim3(nwg), dim3(PWG), 0, ctx->stream, A, lda, n, j0, jb, ipiv, info,
                         err, ps, rows, ctr, base);
    e = hipGetLastError();
  }
  prof_end(ctx, tok, HFPX_KC_PANEL, (double)m * jb * jb, 0.0);
  HFPX_TRY_HIP(ctx, e);
  if (!xchg) ctx->sync_base += (uint64_t)jb * nwg;
  return HFPX_OK;
}

// Blocking: outer blocks of NB columns are factorised recursively (halves,
// down to IB-column panels); the trailing matrix is updated once per outer
// block with K = NB, so it is read and written n / NB times in all.
constexpr int LU_IB = 32;
constexpr int LU_NB = 1024;      // default outer block (HFPX_KNOB_LU_NB)
constexpr int LU_NB_MAX = 4096;
constexpr int LU_MAX_PAIR_BLOCKS = LU_NB_MAX / SWAP_BLK;

template <typename T>
struct LuWork {
  int32_t* dflags;  // [0] info, [1] barrier timeout
  PanelScratch<T>* ps;
  T* rows;
  uint64_t* xw;  // exchange words of k_panel_x, cleared per factorisation
  PermPairs* pp;  // LU_MAX_PAIR_BLOCKS blocks
  int max_wg;
  bool coop;      // launch panels cooperatively (co-residency guaranteed by the runtime)
};

// Interchanges ipiv[k0, k1) on columns [c0, c1) and [c2, c3).
template <typename T>
static int swap_cols(hfpx_ctx* ctx, const LuWork<T>& W, T* A, int64_t lda, const int32_t* ipiv, int64_t n,
                     int64_t k0, int64_t k1, int64_t c0, int64_t c1, int64_t c2, int64_t c3) {
  if (k1 <= k0 || (c1 <= c0 && c3 <= c2)) return HFPX_OK;
  int rc;
  if (k1 - k0 <= SWAP_SMALL) {
    const int64_t ncol = std::max<int64_t>(0, c1 - c0) + std::max<int64_t>(0, c3 - c2);
    if (c1 < c0) c1 = c0;
    if (c3 < c2) c3 = c2;
    const int tok = prof_begin(ctx, HFPX_KC_LASWP);
    hipLaunchKernelGGL(k_swap_small<T>, dim3((unsigned)((ncol + 3) / 4)), dim3(256), 0, ctx->stream, A, lda, ipiv,
                       k0, k1, n, c0, c1, c2, c3);
    prof_end(ctx, tok, HFPX_KC_LASWP, 0.0, 2.0 * sizeof(T) * ncol * 2 * (k1 - k0));
    HFPX_TRY_HIP(ctx, hipGetLastError());
    return HFPX_OK;
  }
  if ((rc = perm_pairs(ctx, ipiv, k0, k1, n, W.pp))) return rc;
  const int64_t nblk = (k1 - k0 + SWAP_BLK - 1) / SWAP_BLK;
  for (int64_t b = 0; b < nblk; ++b)
    if ((rc = apply_pairs<T>(ctx, A, lda, W.pp + b, c0, c1, c2, c3))) return rc;
  return HFPX_OK;
}

// Factorise columns [j, j + w) over rows [j, n): interchanges are applied to
// these columns only; the caller applies them elsewhere.
template <typename T>
static int factor_cols(hfpx_ctx* ctx, const LuWork<T>& W, T* A, int64_t lda, int64_t n, int64_t j, int64_t w,
                       int32_t* ipiv) {
  if (w <= LU_IB)
    return panel<T, LU_IB>(ctx, A, lda, n, j, (int)w, ipiv, W.dflags, W.ps, W.rows, W.xw, W.max_wg, W.coop);
  const int64_t w1 = ((w / 2 + LU_IB - 1) / LU_IB) * LU_IB;
  int rc;
  if ((rc = factor_cols<T>(ctx, W, A, lda, n, j, w1, ipiv))) return rc;
  if ((rc = swap_cols<T>(ctx, W, A, lda, ipiv, n, j, j + w1, j + w1, j + w, 0, 0))) return rc;
  T* a12 = A + j + (j + w1) * lda;
  if ((rc = trsm_lunit<T>(ctx, A + j + j * lda, lda, a12, lda, (int)w1, w - w1))) return rc;
  if ((rc = gemm_sub<T>(ctx, A + (j + w1) + (j + w1) * lda, lda, A + (j + w1) + j * lda, lda, a12, lda,
                        n - (j + w1), w - w1, w1)))
    return rc;
  if ((rc = factor_cols<T>(ctx, W, A, lda, n, j + w1, w - w1, ipiv))) return rc;
  return swap_cols<T>(ctx, W, A, lda, ipiv, n, j + w1, j + w, j, j + w1, 0, 0);
}

// Launches of gemm_sub on the look-ahead stream, with its own split-operand
// workspace (the main stream's factorisation uses the context's).
struct OnSecondStream {
  hfpx_ctx* c;
  hipStream_t s0;
  void* w0;
  size_t b0;
  explicit OnSecondStream(hfpx_ctx* ctx) : c(ctx), s0(ctx->stream), w0(ctx->x3_work), b0(ctx->x3_work_bytes) {
    c->stream = c->stream2;
    c->x3_work = c->x3_work2;
    c->x3_work_bytes = c->x3_work2_bytes;
  }
  ~OnSecondStream() {
    c->x3_work2 = c->x3_work;
    c->x3_work2_bytes = c->x3_work_bytes;
    c->stream = s0;
    c->x3_work = w0;
    c->x3_work_bytes = b0;
  }
};

static int ensure_lookahead(hfpx_ctx* ctx) {
  if (!ctx->stream2) HFPX_TRY_HIP(ctx, hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking));
  for (hipEvent_t& e : ctx->la_ev)
    if (!e) HFPX_TRY_HIP(ctx, hipEventCreateWithFlags(&e, hipEventDisableTiming));
  return HFPX_OK;
}

// Right-looking blocked LU with a look-ahead of one outer block.  After
// block k (columns [k, k + kb)) is factorised, its interchanges and its
// U12 rows are applied to every column right of it; the next block's
// columns [k2, k2 + kb2) are updated first (main stream), then
//   main stream:   factorise the next block;
//   second stream: update the remaining columns [k2 + kb2, n) with block k;
// both read block k's L21 and U12, neither writes what the other reads.
// The next block's interchanges on the columns left of it permute rows of
// block k's L21, which the second stream reads, so they wait for it.  The
// operations and their order on every matrix entry are those of the plain
// loop, so the factors are bitwise the same.
template <typename T>
static int getrf_lookahead(hfpx_ctx* ctx, const LuWork<T>& W, T* A, int64_t lda, int64_t n, int64_t NB,
                           int32_t* ipiv) {
  int rc;
  if ((rc = ensure_lookahead(ctx))) return rc;
  if constexpr (sizeof(T) == 4) {
    // both streams' split operands, sized for the largest trailing update
    const size_t b = gemm_x3_work_bytes(n - NB, n - NB, NB);
    if ((rc = ensure_x3_work(ctx, b))) return rc;
    {
      OnSecondStream on(ctx);
      rc = ensure_x3_work(ctx, b);
    }
    if (rc) return rc;
  }
  int64_t k = 0, kb = std::min<int64_t>(NB, n);
  if ((rc = factor_cols<T>(ctx, W, A, lda, n, k, kb, ipiv))) return rc;
  while (true) {
    const int64_t rest = n - (k + kb);
    // block k's interchanges on the columns right of it (those left of it
    // were applied when it was factorised)
    if ((rc = swap_cols<T>(ctx, W, A, lda, ipiv, n, k, k + kb, k + kb, n, 0, 0))) return rc;
    if (rest <= 0) break;
    T* a12 = A + k + (k + kb) * lda;
    const T* l21 = A + (k + kb) + k * lda;
    if ((rc = trsm_lunit<T>(ctx, A + k + k * lda, lda, a12, lda, (int)kb, rest))) return rc;
    const int64_t k2 = k + kb, kb2 = std::min<int64_t>(NB, n - k2), rest2 = rest - kb2;
    // the next block's columns
    if ((rc = gemm_sub<T>(ctx, A + k2 + k2 * lda, lda, l21, lda, a12, lda, rest, kb2, kb))) return rc;
    if (rest2 > 0) {
      HFPX_TRY_HIP(ctx, hipEventRecord(ctx->la_ev[0], ctx->stream));
      HFPX_TRY_HIP(ctx, hipStreamWaitEvent(ctx->stream2, ctx->la_ev[0], 0));
      {
        OnSecondStream on(ctx);
        rc = gemm_sub<T>(ctx, A + k2 + (k2 + kb2) * lda, lda, l21, lda, a12 + kb2 * lda, lda, rest, rest2, kb);
      }
      if (rc) return rc;
      HFPX_TRY_HIP(ctx, hipEventRecord(ctx->la_ev[1], ctx->stream2));
    }
    if ((rc = factor_cols<T>(ctx, W, A, lda, n, k2, kb2, ipiv))) return rc;
    if (rest2 > 0) HFPX_TRY_HIP(ctx, hipStreamWaitEvent(ctx->stream, ctx->la_ev[1], 0));
    // the next block's interchanges on every column left of it
    if ((rc = swap_cols<T>(ctx, W, A, lda, ipiv, n, k2, k2 + kb2, 0, k2, 0, 0))) return rc;
    k = k2;
    kb = kb2;
  }
  return HFPX_OK;
}

template <typename T>
static int getrf_impl(hfpx_ctx* ctx, T* A, int64_t lda, int64_t n, int32_t* ipiv, int64_t* info_out) {
  static_assert(LU_NB_MAX % SWAP_BLK == 0, "pair blocks per outer block");
  const int64_t NB = ctx->knob[HFPX_KNOB_LU_NB] > 0 ? ctx->knob[HFPX_KNOB_LU_NB] : LU_NB;
  // co-resident panel workgroups
  int occ = 0, occx = 0;
  HFPX_TRY_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_panel<T, LU_IB>, PWG, 0));
  HFPX_TRY_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&occx, k_panel_x<T, LU_IB>, PWG, 0));
  occ = std::min(occ, occx);
  if (occ < 1) return set_error(ctx, HFPX_ERR_HIP, "getrf: panel kernel cannot be resident");
  LuWork<T> W;
  W.max_wg = std::min(PANEL_MAX_WG, occ * std::max(1, ctx->num_cus));
  const size_t rows_bytes = sizeof(T) * 2 * (size_t)(W.max_wg + 1) * LU_IB;
  const size_t off_ps = 256, off_rows = off_ps + ((sizeof(PanelScratch<T>) + 255) & ~(size_t)255);
  const size_t off_pp = off_rows + ((rows_bytes + 255) & ~(size_t)255);
  const size_t xw_bytes = sizeof(uint64_t) * 2 * (PX_MAX_WG + 1) * PxFmt<T, LU_IB>::WORDS;
  const size_t off_xw = off_pp + ((LU_MAX_PAIR_BLOCKS * sizeof(PermPairs) + 255) & ~(size_t)255);
  int rc = ensure_lu_work(ctx, off_xw + xw_bytes);
  if (rc) return rc;
  char* w = (char*)ctx->lu_work;
  W.dflags = (int32_t*)w;
  W.ps = (PanelScratch<T>*)(w + off_ps);
  W.rows = (T*)(w + off_rows);
  W.pp = (PermPairs*)(w + off_pp);
  W.xw = (uint64_t*)(w + off_xw);
  HFPX_TRY_HIP(ctx, hipMemsetAsync(W.dflags, 0, 2 * sizeof(int32_t), ctx->stream));
  // host-forced test of the time-out path: the flag starts raised
  if (ctx->knob[HFPX_KNOB_TEST_PANEL_TIMEOUT])
    HFPX_TRY_HIP(ctx, hipMemsetD32Async((hipDeviceptr_t)(W.dflags + 1), 1, 1, ctx->stream));
  HFPX_TRY_HIP(ctx, hipMemsetAsync(W.xw, 0, xw_bytes, ctx->stream));
  HFPX_TRY_HIP(ctx, hipMemsetAsync(ctx->sync, 0, sizeof(uint64_t), ctx->stream));
  ctx->sync_base = 0;

  // look-ahead by default in FP32 only (72,000 DoF: FP32 2.14 -> 1.81 s,
  // FP64 5.24 -> 6.28 s with the overlapped FP64 updates), with the panels
  // launched on the context's stream rather than through the cooperative
  // launch (mode 3): the cooperative launches go to a queue of their own,
  // and each panel then paid two cross-queue waits of about 12 us (72,000
  // DoF: 14 to 25 ms per factorisation).  Co-residency still holds: the
  // grid is at most one workgroup per CU, and the only other kernels that
  // can hold CUs meanwhile are this factorisation's own trailing updates on
  // the second stream, which never wait on the panel, so every panel
  // workgroup is placed once they drain; the barrier's bounded wait turns
  // anything else into an error.
  const int la = ctx->knob[HFPX_KNOB_LU_LOOKAHEAD] ? ctx->knob[HFPX_KNOB_LU_LOOKAHEAD] : (sizeof(T) == 4 ? 3 : 1);
  const bool lookahead = (la == 2 || la == 3) && n > NB;
  W.coop = la != 3;
  if (!lookahead) {
    for (int64_t k = 0; k < n; k += NB) {
      const int64_t kb = std::min<int64_t>(NB, n - k);
      if ((rc = factor_cols<T>(ctx, W, A, lda, n, k, kb, ipiv))) return rc;
      // interchanges on the columns left and right of the outer block
      if ((rc = swap_cols<T>(ctx, W, A, lda, ipiv, n, k, k + kb, 0, k, k + kb, n))) return rc;
      const int64_t rest = n - (k + kb);
      if (rest > 0) {
        T* a12 = A + k + (k + kb) * lda;
        if ((rc = trsm_lunit<T>(ctx, A + k + k * lda, lda, a12, lda, (int)kb, rest))) return rc;
        if ((rc = gemm_sub<T>(ctx, A + (k + kb) + (k + kb) * lda, lda, A + (k + kb) + k * lda, lda, a12, lda,
                              rest, rest, kb)))
          return rc;
      }
    }
  } else if ((rc = getrf_lookahead<T>(ctx, W, A, lda, n, NB, ipiv))) {
    return rc;
  }
  int32_t hflags[2] = {0, 0};
  HFPX_TRY_HIP(ctx, hipMemcpyAsync(hflags, W.dflags, sizeof(hflags), hipMemcpyDeviceToHost, ctx->stream));
  HFPX_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (hflags[1]) return set_error(ctx, HFPX_ERR_HIP, "getrf: a panel wait timed out (factors are invalid)");
  if (info_out) *info_out = hflags[0];
  return hflags[0] ? HFPX_ERR_SINGULAR : HFPX_OK;
}

// B (n x nrhs) := A^{-1} B using the factors (dgetrs 'N').
template <typename T>
static int getrs_impl(hfpx_ctx* ctx, const T* LU, int64_t lda, const int32_t* ipiv, int64_t n,
                      T* B, int64_t ldb, int64_t nrhs) {
  constexpr int NB = 256;
  int rc;
  // row interchanges on B: pair blocks for all of ipiv, then one workgroup
  // per right-hand side applies them in order
  const int64_t npb = (n + SWAP_BLK - 1) / SWAP_BLK;
  const size_t pairs_bytes = (sizeof(PermPairs) * (size_t)npb + 255) & ~(size_t)255;
  const bool sweep = std::is_same<T, float>::value;  // FP32: block-inverse sweeps
  if ((rc = ensure_lu_work(ctx, pairs_bytes + (sweep ? sweep_f32_work_bytes(n, nrhs) : 0)))) return rc;
  PermPairs* pp = (PermPairs*)ctx->lu_work;
  if ((rc = perm_pairs(ctx, ipiv, 0, n, n, pp))) return rc;
  {
    const int tok = prof_begin(ctx, HFPX_KC_LASWP);
    hipLaunchKernelGGL(k_apply_pairs_seq<T>, dim3((unsigned)nrhs), dim3(256), 0, ctx->stream, B, ldb, pp,
                       (int)npb);
    prof_end(ctx, tok, HFPX_KC_LASWP, 0.0, 0.0);
    HFPX_TRY_HIP(ctx, hipGetLastError());
  }
  if constexpr (std::is_same<T, float>::value)
    return sweep_f32(ctx, LU, lda, n, B, ldb, nrhs, (char*)ctx->lu_work + pairs_bytes);
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

}  // namespace hfpx
