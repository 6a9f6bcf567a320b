// hfpx_sweep.hip -- FP32 triangular sweeps of the LU solve for many right-hand
// sides (the preconditioner applications of the mixed-precision solve).
//
// Replaces, for the FP32 factors, the forward/back substitution of
// LAPACKE_dgetrs (il/linear_algebra/dense/factorization/LU.h:106-138).
//
// Blocks of SB = 128 rows.  One launch inverts every diagonal block of L
// (unit lower) and of U.  Each sweep is then one launch in dependency-flow
// form: workgroups take row blocks in sweep order from an atomic ticket
// counter; the workgroup of block i accumulates T(i, c) X_c over the blocks
// c before it in sweep order as each one's flag appears, then multiplies
// B_i - sum by block i's inverse,
// which makes X_i final in place, and raises block i's flag.  The last
// predecessor p of block i (its neighbour in sweep order) is taken through
// W_i = inverse(i) T(i, p), formed once per factorisation next to the
// inverses: X_i = inverse(i) (B_i - sum over the other predecessors) - W_i X_p,
// where the first term is complete before X_p is, so the critical path is
// one 128 x 128 product per block instead of two, with no dependent load of
// the inverse; the factor is read once per sweep.
#include "hfpx_internal.h"

namespace hfpx {

typedef float floatx16 __attribute__((ext_vector_type(16)));

constexpr int SB = 128;  // rows per block
// right-hand sides per sweep workgroup: 32 (four waves) up to 64 columns,
// which halves the products on a block's critical path (72,000 DoF, 64
// columns: 23.0 -> 19.2 ms per forward + backward pair, the same bits), 64
// (eight waves) above, where every column block reads the factor once more
constexpr int SWN_MAX = 64, SWN_MIN = 32;

// Inverse of diagonal block b (rows/columns [128 b, 128 b + kb)) of the
// n x n factor LU (lda), padded with the identity to 128 x 128, column-major
// into out (SB^2 floats).  128 threads: thread t forms column t of the
// inverse by substitution in 32-row chunks.  A chunk's partial sums stay in
// registers (fully unrolled over the chunk, factor entries read from LDS as
// broadcasts); the finished chunks of every column sit in sX for the later
// chunks and the final coalesced store.  The terms are added in the order
// of plain column-oriented substitution (unit lower: columns of L ascending;
// upper: columns of U descending, each entry divided by its diagonal when it
// is reached; entries below the diagonal of column t of an upper inverse are
// never divided), so the result is that of the one-column-at-a-time form.
template <bool LOWER>
__device__ __forceinline__ void block_inv_tile(const float* __restrict__ LU, int64_t lda, int64_t n, int64_t b,
                                               float* __restrict__ out) {
  constexpr int BC = 32;       // rows per chunk
  constexpr int NC = SB / BC;  // chunks
  constexpr int PT = SB + 4;   // sT pitch: 16-byte aligned column segments
  constexpr int PX = SB + 1;   // sX pitch
  __shared__ __align__(16) float sT[SB * PT];  // sT[i + j PT] = T(i, j)
  __shared__ float sX[SB * PX];                // sX[i PX + c] = inverse(i, c)
  const int64_t k = b * SB;
  const int kb = (int)std::min<int64_t>(SB, n - k);
  const int t = threadIdx.x;
  // stage T: thread t holds row t of every column; 16 loads in flight
  for (int j0 = 0; j0 < SB; j0 += 16) {
    float v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int j = j0 + u;
      v[u] = (t < kb && j < kb) ? LU[(k + t) + (k + j) * lda] : (t == j ? 1.f : 0.f);
    }
#pragma unroll
    for (int u = 0; u < 16; ++u) sT[t + (j0 + u) * PT] = v[u];
  }
  __syncthreads();
  for (int q = 0; q < NC; ++q) {
    const int I = LOWER ? q : NC - 1 - q;  // chunk in substitution order
    float acc[BC];
#pragma unroll
    for (int i = 0; i < BC; ++i) acc[i] = (I * BC + i == t) ? 1.f : 0.f;
    // finished chunks: acc -= T(chunk I, chunk K) x(chunk K)
    for (int r = 0; r < q; ++r) {
      const int K = LOWER ? r : NC - 1 - r;
#pragma unroll
      for (int u = 0; u < BC; ++u) {
        const int kk = LOWER ? u : BC - 1 - u;
        const float xk = sX[(K * BC + kk) * PX + t];
        const float* col = sT + I * BC + (K * BC + kk) * PT;
#pragma unroll
        for (int i = 0; i < BC; i += 4) {
          const float4 tv = *reinterpret_cast<const float4*>(col + i);
          acc[i] = fmaf(-tv.x, xk, acc[i]);
          acc[i + 1] = fmaf(-tv.y, xk, acc[i + 1]);
          acc[i + 2] = fmaf(-tv.z, xk, acc[i + 2]);
          acc[i + 3] = fmaf(-tv.w, xk, acc[i + 3]);
        }
      }
    }
    // the chunk's own triangle
    if (LOWER) {
#pragma unroll
      for (int kk = 0; kk < BC; ++kk) {
        const float xk = acc[kk];
        const float* col = sT + I * BC + (I * BC + kk) * PT;
#pragma unroll
        for (int i = kk + 1; i < BC; ++i) acc[i] = fmaf(-col[i], xk, acc[i]);
      }
    } else {
#pragma unroll
      for (int kk = BC - 1; kk >= 0; --kk) {
        const int j = I * BC + kk;
        const float* col = sT + I * BC + j * PT;
        const float xk = (j > t) ? 0.f : acc[kk] / col[kk];
        acc[kk] = xk;
#pragma unroll
        for (int i = 0; i < kk; ++i) acc[i] = fmaf(-col[i], xk, acc[i]);
      }
    }
#pragma unroll
    for (int i = 0; i < BC; ++i) sX[(I * BC + i) * PX + t] = acc[i];
  }
  __syncthreads();
  for (int e = t; e < SB * SB; e += SB) out[e] = sX[(e % SB) * PX + e / SB];
}

// Inverse of diagonal block blockIdx.x into inv + blockIdx.x SB^2.
template <bool LOWER>
__global__ __launch_bounds__(SB) void k_block_inv(const float* __restrict__ LU, int64_t lda, int64_t n,
                                                  float* __restrict__ inv) {
  block_inv_tile<LOWER>(LU, lda, n, blockIdx.x, inv + (int64_t)blockIdx.x * SB * SB);
}

// W_b = inverse(b) T(b, p) for the block p before b in sweep order
// (LOWER: p = b - 1, unit L; UPPER: p = b + 1, U), 128 x 128 column-major
// at W + b SB^2; nothing for the first block in sweep order.  T(b, p) is
// staged in LDS (zero outside the matrix; the inverse is identity-padded, so
// W is zero outside the blocks' rows and columns); 4 waves, wave w forms
// rows 32 w .. 32 w + 31 in the transposed MFMA form of the sweep.
template <bool LOWER>
__global__ __launch_bounds__(256) void k_block_w(const float* __restrict__ LU, int64_t lda, int64_t n, int64_t nblk,
                                                 const float* __restrict__ inv, float* __restrict__ W) {
  const int64_t b = blockIdx.x;
  const int64_t p = LOWER ? b - 1 : b + 1;
  if (p < 0 || p >= nblk) return;
  __shared__ float sT[SB][SB + 1];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, l32 = lane & 31, hk = lane >> 5;
  const int64_t r0 = b * SB, k0 = p * SB;
  const int rows = (int)std::min<int64_t>(SB, n - r0), kc = (int)std::min<int64_t>(SB, n - k0);
  {
    // 64 loads per thread in flight, then the LDS stores: sT[k][j] = T(k, j)
    float v[SB * SB / 256];
#pragma unroll
    for (int q = 0; q < SB * SB / 256; ++q) {
      const int e = tid + 256 * q, k = e % SB, j = e / SB;
      v[q] = (k < rows && j < kc) ? LU[(r0 + k) + (k0 + j) * lda] : 0.f;
    }
#pragma unroll
    for (int q = 0; q < SB * SB / 256; ++q) {
      const int e = tid + 256 * q;
      sT[e % SB][e / SB] = v[q];
    }
  }
  const int i = 32 * wave + l32;
  const float* iv = inv + b * SB * SB + i;
  float ivv[SB / 2];
#pragma unroll
  for (int s2 = 0; s2 < SB / 2; ++s2) ivv[s2] = iv[(2 * s2 + hk) * SB];
  __syncthreads();
  float* wb = W + b * SB * SB;
#pragma unroll 1
  for (int jb = 0; jb < SB / 32; ++jb) {
    floatx16 acc;
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[v] = 0.f;
#pragma unroll
    for (int s2 = 0; s2 < SB / 2; ++s2)
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(sT[2 * s2 + hk][32 * jb + l32], ivv[s2], acc, 0, 0, 0);
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const int j = 32 * jb + 8 * (v >> 2) + 4 * hk + (v & 3);
      wb[i + j * SB] = acc[v];
    }
  }
}

// One sweep (see the file comment).  LOWER: forward with unit L, blocks in
// increasing order; UPPER: backward with U, decreasing order.  A ticket t is
// the pair (q = t / gy, right-hand-side block t % gy) with q the position of
// the row block in sweep order; flags[rb nblk + block] = want once X_block
// of right-hand-side block rb is final.  Tickets are taken in order and a
// workgroup only waits on lower tickets, which were handed to workgroups
// already running: any grid size makes progress.
// Memory ordering: the solution rows that cross workgroups are written and
// read only as agent-scope relaxed atomics (device-coherent stores and loads
// that do not stop in an XCD's L2), so no fence has to write back or
// invalidate a whole L2: an agent-scope fence costs a write-back of the
// producer's L2 and an invalidation of the consumer's, on every chain step
// of the sweep and under every other workgroup's streaming loads.  The
// producer waits until its own stores are complete (vmcnt 0) before the
// workgroup barrier, then thread 0 stores the flag; the consumer's thread 0
// polls relaxed (bounded: past ~2^20 polls *err is raised, after which every
// other wait of the sweep gives up too and the sweep completes with garbage,
// reported by the host) and the workgroup barrier orders the other threads'
// loads of the rows after the flag was seen.
// 4 SWN / 32 waves, one 32 x 32 block of the 128 x SWN tile each.  Products use
// the transposed MFMA form (v_mfma_f32_32x32x2f32): lane l holds row
// 32 (wave % 4) + l % 32 of the tile and column 8 (v/4) + 4 (l/32) + v%4 of
// the wave's 32-column block, so tile loads and stores are column segments.
template <bool LOWER, int SWN>
__global__ __launch_bounds__(8 * SWN) void k_sweep_flow(const float* __restrict__ LU, int64_t lda, int64_t n,
                                                    int64_t nblk, const float* __restrict__ inv,
                                                    const float* __restrict__ wpred,
                                                    float* __restrict__ B, int64_t ldb, int64_t nrhs, int64_t gy,
                                                    uint64_t* __restrict__ flags, uint64_t want,
                                                    uint64_t* __restrict__ ticket, int32_t* __restrict__ err) {
  constexpr int SWT = 8 * SWN;  // threads
  __shared__ float sX[SB][SWN + 1];  // solution rows of block c: sX[k][j]
  __shared__ float sC[SB][SWN + 1];  // B_i - sum, then X_i: sC[i][j]
  __shared__ uint64_t s_tk;
  const int tid = threadIdx.x, wave = tid >> 6, jb = wave >> 2;  // jb: 32-column block of this wave
  const int64_t total = nblk * gy;
  // Thread 0 takes the tickets; every thread reads them as wave-uniform
  // (scalar) values.  The ticket is taken at the end of the loop body, not at
  // its head: with the divergent "thread 0 takes a ticket" block at the loop
  // head, the compiler's structurisation split the loop so that lanes 1..63
  // of wave 0 kept iterating on the stale ticket while lane 0 waited to take
  // the next one (a hang).
  auto next_ticket = [&]() -> int64_t {
    if (tid == 0) s_tk = __hip_atomic_fetch_add(ticket, (uint64_t)1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const uint64_t u = s_tk;
    const uint64_t t = ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(u >> 32)) << 32) |
                       (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)u);
    __syncthreads();  // every thread has s_tk before thread 0 may take the next ticket
    return (int64_t)t;
  };
  for (int64_t tk = next_ticket(); tk < total; tk = next_ticket()) {
    // The lane's coordinates are made opaque per ticket: otherwise the
    // compiler hoists every value derived from them (the 64 column indices
    // and their comparisons, the LDS addresses of each product) out of the
    // ticket loop and holds them in registers for the whole kernel, which
    // leaves too few for the double-buffered factor tiles below.
    int lane = tid & 63;
    asm volatile("" : "+v"(lane));
    const int l32 = lane & 31, hk = lane >> 5;
    const int i = 32 * (wave & 3) + l32;  // tile row of this lane
    const int64_t q = tk / gy, rb = tk % gy;
    const int64_t blk = LOWER ? q : nblk - 1 - q;
    const int64_t r0 = blk * SB;
    const int rows = (int)std::min<int64_t>(SB, n - r0);
    const int64_t j0 = rb * SWN;
    const int ncol = (int)std::min<int64_t>(SWN, nrhs - j0);
    const bool rok = i < rows;
    uint64_t* fl = flags + rb * nblk;

    // wait until predecessor p (in sweep order) is final, then take every
    // later one already final, so that one acquire fence covers the run
    int64_t ready = 0;  // predecessors 0 .. ready-1 (in sweep order) are known final
    auto wait_for = [&](int64_t p) {
      if (tid == 0) {
        const int64_t c = LOWER ? p : nblk - 1 - p;
        uint32_t polls = 0;
        while (__hip_atomic_load(fl + c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != want) {
          __builtin_amdgcn_s_sleep(1);
          // give up past ~2^20 polls, or at once when another wait already has
          if (++polls > (1u << 20) ||
              ((polls & 255) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
            __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
          }
        }
        int64_t r = p + 1;
        while (r < q && __hip_atomic_load(fl + (LOWER ? r : nblk - 1 - r), __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT) == want)
          ++r;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        s_tk = (uint64_t)r;
      }
      __syncthreads();
      const uint64_t u = s_tk;
      ready = (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(u >> 32)) << 32) |
                        (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)u));
    };
    // the solution rows of block c (kc rows) into sX: every load in flight
    // before the first LDS store, one memory latency per predecessor
    auto stage_x = [&](int64_t k0, int kc) {
      constexpr int XQ = SB * SWN / SWT;
      float xq[XQ];
#pragma unroll
      for (int q2 = 0; q2 < XQ; ++q2) {
        const int e = tid + SWT * q2, k = e % SB, j = e / SB;
        xq[q2] = (k < kc && j < ncol)
                     ? __hip_atomic_load(B + (k0 + k) + (j0 + j) * ldb, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                     : 0.f;
      }
#pragma unroll
      for (int q2 = 0; q2 < XQ; ++q2) {
        const int e = tid + SWT * q2;
        sX[e % SB][e / SB] = xq[q2];
      }
      __syncthreads();
    };

    // predecessors through the factor tiles T(i, c); the last one through
    // W_i when it is given.  The loop is software-pipelined over two static
    // register sets (unrolled by two): the next predecessor's factor tile,
    // and its solution rows when it is already known final, are loaded while
    // the current one's products run, so an iteration costs its products
    // rather than two memory latencies (the sweep is bound by this loop:
    // every workgroup walks all of its block's predecessors).
    const bool usew = wpred != nullptr && q > 0;
    const int64_t pend = usew ? q - 1 : q;
    floatx16 acc;
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[v] = 0.f;
    // factor tile T(i, c) of predecessor p into lv: a wave-uniform column
    // base stepped by two columns per load plus one 32-bit lane offset, so
    // the addresses live in scalar registers (a 64-bit pointer per load would
    // cost two vector registers each)
    // (the step is made opaque per call so that the 64 column offsets are
    // not hoisted out of the ticket loop into registers).  A tile inside the
    // matrix loads without masks, back to back; only tiles of the last row
    // or column block take the masked path.
    const uint32_t loff = (uint32_t)i + (uint32_t)hk * (uint32_t)lda;
    auto load_t = [&](int64_t p, float* lv) {
      const int64_t c = LOWER ? p : nblk - 1 - p;
      const int64_t k0 = c * SB;
      const int kc = (int)std::min<int64_t>(SB, n - k0);
      const float* col = LU + r0 + k0 * lda;
      int64_t step = 2 * lda;
      asm volatile("" : "+s"(step));
      if (kc == SB && rows == SB) {
#pragma unroll
        for (int s = 0; s < SB / 2; ++s) {
          lv[s] = col[loff];
          col += step;
        }
      } else {
        const float* lp = LU + (r0 + i) + (k0 + hk) * lda;
#pragma unroll
        for (int s = 0; s < SB / 2; ++s) {
          lv[s] = (rok && 2 * s + hk < kc) ? *lp : 0.f;
          lp += step;
        }
      }
    };
    auto stage_p = [&](int64_t p) {
      if (p >= ready) wait_for(p);
      const int64_t c = LOWER ? p : nblk - 1 - p;
      stage_x(c * SB, (int)std::min<int64_t>(SB, n - c * SB));
    };
    auto mma = [&](const float* lv) {
#pragma unroll
      for (int s = 0; s < SB / 2; ++s)
        acc = __builtin_amdgcn_mfma_f32_32x32x2f32(sX[2 * s + hk][32 * jb + l32], lv[s], acc, 0, 0, 0);
      __syncthreads();  // sX is rewritten by the next predecessor
    };
    if constexpr (SWN == SWN_MIN) {
      float lvA[SB / 2], lvB[SB / 2];
      if (pend > 0) load_t(0, lvA);
      for (int64_t p = 0; p < pend; p += 2) {
        stage_p(p);
        if (p + 1 < pend) load_t(p + 1, lvB);
        mma(lvA);
        if (p + 1 >= pend) break;
        stage_p(p + 1);
        if (p + 2 < pend) load_t(p + 2, lvA);
        mma(lvB);
      }
    } else {
      // two waves per SIMD leave no registers for a second tile
      for (int64_t p = 0; p < pend; ++p) {
        float lv[SB / 2];
        load_t(p, lv);
        stage_p(p);
        mma(lv);
      }
    }
    // inverse(i) (B_i - sum).  The block's inverse is fetched here,
    // together with B_i, so both share one memory latency (held across the
    // predecessor loop it would cost the registers that loop's staging uses)
    const float* iv = inv + blk * SB * SB + i;  // inverse(i, k) at iv[k SB]
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
    if (usew) {
      // X_i = y - W_i X_p for the last predecessor p: W_i is fetched before
      // the wait, so after X_p appears one product remains.  (The scheduling
      // barrier keeps the W_i loads after the inverse product: hoisted above
      // it they would hold the inverse and W_i in registers together.)
      __builtin_amdgcn_sched_barrier(0);
      const int64_t p = q - 1;
      const int64_t c = LOWER ? p : nblk - 1 - p;
      const int64_t k0 = c * SB;
      const int kc = (int)std::min<int64_t>(SB, n - k0);
      const float* wp = wpred + blk * SB * SB + i;  // W_i(i, k) at wp[k SB]
      float wv[SB / 2];
#pragma unroll
      for (int s = 0; s < SB / 2; ++s) wv[s] = wp[(2 * s + hk) * SB];
      if (p >= ready) wait_for(p);
      stage_x(k0, kc);
      floatx16 a2;
#pragma unroll
      for (int v = 0; v < 16; ++v) a2[v] = 0.f;
#pragma unroll
      for (int s = 0; s < SB / 2; ++s)
        a2 = __builtin_amdgcn_mfma_f32_32x32x2f32(sX[2 * s + hk][32 * jb + l32], wv[s], a2, 0, 0, 0);
#pragma unroll
      for (int v = 0; v < 16; ++v) y[v] -= a2[v];
    }
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const int j = 32 * jb + 8 * (v >> 2) + 4 * hk + (v & 3);
      if (rok && j < ncol)
        __hip_atomic_store(B + (r0 + i) + (j0 + j) * ldb, y[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // this thread's stores of X_i are complete
    __syncthreads();  // every store of X_i, and every read of sC, before the flag
    if (tid == 0) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __hip_atomic_store(fl + blk, want, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

__global__ void k_fill_u64(uint64_t* __restrict__ p, int64_t cnt) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e < cnt) p[e] = 0;
}

// work: the block inverses of L and U, then the flags (one per row block
// and right-hand-side block, per sweep), two ticket counters and two error
// flags
static int sweep_width(int64_t nrhs) { return nrhs <= SWN_MAX ? SWN_MIN : SWN_MAX; }

template <int SWN>
static int launch_sweeps(hfpx_ctx* ctx, const float* LU, int64_t lda, int64_t n, int64_t nblk, const float* invL,
                         const float* invU, const float* wL, const float* wU, float* B, int64_t ldb, int64_t nrhs,
                         uint64_t* flags, uint64_t* tickets, int32_t* errs) {
  const int64_t gy = (nrhs + SWN - 1) / SWN;
  // as many workgroups as are resident at once (any grid size would make
  // progress: a workgroup takes tickets only once it runs)
  int occ = 0;
  HFPX_TRY_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_sweep_flow<true, SWN>, 8 * SWN, 0));
  const unsigned nwg = (unsigned)std::min<int64_t>(nblk * gy, std::max(1, occ) * std::max(1, ctx->num_cus));
  hipLaunchKernelGGL((k_sweep_flow<true, SWN>), dim3(nwg), dim3(8 * SWN), 0, ctx->stream, LU, lda, n, nblk, invL, wL,
                     B, ldb, nrhs, gy, flags, (uint64_t)1, tickets, errs);
  hipLaunchKernelGGL((k_sweep_flow<false, SWN>), dim3(nwg), dim3(8 * SWN), 0, ctx->stream, LU, lda, n, nblk, invU,
                     wU, B, ldb, nrhs, gy, flags + nblk * gy, (uint64_t)1, tickets + 1, errs + 1);
  HFPX_TRY_HIP(ctx, hipGetLastError());
  return HFPX_OK;
}

// ---------------------------------------------------------------------
// Batched form: one right-hand side per crack, many cracks of uneven size.
// One launch per sweep for all the cracks of a list, persistent: workgroups
// take tickets from an atomic counter in a table ordered by level (the
// position of a row block in sweep order) and, within a level, by crack
// size, largest first -- a ticket only waits on tickets of lower levels of
// its own crack, so any number of resident workgroups makes progress, the
// first tickets start every crack's chain at once and the largest cracks'
// long chains start first.  A ticket (crack c, level q) accumulates
// T(i, p) X_p over the crack's earlier blocks p as each one's flag appears
// (256 threads: row r = t % 128, half t / 128 of the 128 columns), then
// X_i = inverse(i) (b_i - sum) in place and raises block i's flag.  The
// solved rows cross workgroups as device-coherent atomics, with the same
// ordering as the single-crack sweep (no L2 write-back or invalidation).
// ---------------------------------------------------------------------
// block inverses of every (crack, block) of the ticket table (the level of
// a ticket is its block in forward order)
template <bool LOWER>
__global__ __launch_bounds__(SB) void k_bblock_inv(const float* __restrict__ A32, const BCrack* __restrict__ cr,
                                                   const int32_t* __restrict__ tk, const int64_t* __restrict__ bofs,
                                                   float* __restrict__ inv) {
  const int c = tk[2 * blockIdx.x], b = tk[2 * blockIdx.x + 1];
  const BCrack& C = cr[c];
  block_inv_tile<LOWER>(A32 + C.a32_off, C.n, C.n, b, inv + (bofs[c] + b) * SB * SB);
}

template <bool LOWER>
__global__ __launch_bounds__(2 * SB) __attribute__((amdgpu_waves_per_eu(4))) void k_bsweep(
    const float* __restrict__ A32, const BCrack* __restrict__ cr,
                                                   const int32_t* __restrict__ tk, int64_t ntk,
                                                   const int64_t* __restrict__ bofs, const float* __restrict__ inv,
                                                   float* __restrict__ vec, uint64_t* __restrict__ flags,
                                                   uint64_t* __restrict__ ticket, int32_t* __restrict__ err) {
  constexpr int KH = SB / 2;  // columns per half
  __shared__ float sX[SB];
  __shared__ float sP[2][SB];
  __shared__ uint64_t s_tk;
  const int tid = threadIdx.x, r = tid & (SB - 1), kh = tid >> 7;
  auto uniform = [](uint64_t u) -> int64_t {
    return (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(u >> 32)) << 32) |
                     (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)u));
  };
  // the ticket is taken at the end of the loop body (see k_sweep_flow)
  auto next_ticket = [&]() -> int64_t {
    if (tid == 0) s_tk = __hip_atomic_fetch_add(ticket, (uint64_t)1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const int64_t t = uniform(s_tk);
    __syncthreads();
    return t;
  };
  for (int64_t tkn = next_ticket(); tkn < ntk; tkn = next_ticket()) {
    const int c = tk[2 * tkn], q = tk[2 * tkn + 1];
    const BCrack& C = cr[c];
    const int64_t n = C.n, nb = (n + SB - 1) / SB;
    const int64_t blk = LOWER ? q : nb - 1 - q;
    const int64_t r0 = blk * SB;
    const int rows = (int)std::min<int64_t>(SB, n - r0);
    const float* LU = A32 + C.a32_off;
    float* x = vec + C.x_off;
    uint64_t* fl = flags + bofs[c];
    const bool rok = r < rows;
    float acc = 0.f;
    int64_t ready = 0;
    for (int64_t p = 0; p < q; ++p) {
      const int64_t cb = LOWER ? p : nb - 1 - p;
      const int64_t k0 = cb * SB;
      const int kc = (int)std::min<int64_t>(SB, n - k0);
      // the factor's tile does not depend on the flag: loaded before the
      // wait (the column pointer steps, so no table of the column offsets
      // is held in registers)
      const float* lp = LU + (r0 + r) + (k0 + KH * kh) * n;
      float lv[KH];
#pragma unroll
      for (int k = 0; k < KH; ++k) {
        lv[k] = (rok && KH * kh + k < kc) ? *lp : 0.f;
        lp += n;
      }
      if (p >= ready) {
        if (tid == 0) {
          uint32_t polls = 0;
          while (__hip_atomic_load(fl + cb, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 1) {
            __builtin_amdgcn_s_sleep(1);
            if (++polls > (1u << 20) ||
                ((polls & 255) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
              __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              break;
            }
          }
          int64_t rr = p + 1;
          while (rr < q && __hip_atomic_load(fl + (LOWER ? rr : nb - 1 - rr), __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT) == 1)
            ++rr;
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
          s_tk = (uint64_t)rr;
        }
        __syncthreads();
        ready = uniform(s_tk);
      }
      if (tid < SB)
        sX[tid] = tid < kc ? __hip_atomic_load(x + k0 + tid, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0.f;
      __syncthreads();
#pragma unroll
      for (int k = 0; k < KH; ++k) acc += lv[k] * sX[KH * kh + k];
      __syncthreads();  // sX is rewritten by the next predecessor
    }
    sP[kh][r] = acc;
    __syncthreads();
    if (tid < SB) sX[tid] = tid < rows ? x[r0 + tid] - (sP[0][tid] + sP[1][tid]) : 0.f;
    __syncthreads();
    // X_i = inverse(i) (b_i - sum): inverse(r, k) at iv[k SB]
    const float* iv = inv + (bofs[c] + blk) * SB * SB + r + KH * kh * SB;
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < KH; ++k) {
      s += *iv * sX[KH * kh + k];
      iv += SB;
    }
    sP[kh][r] = s;
    __syncthreads();
    if (tid < rows)
      __hip_atomic_store(x + r0 + tid, sP[0][tid] + sP[1][tid], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // this thread's store of X_i is complete
    __syncthreads();  // every store of X_i before the flag
    if (tid == 0) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __hip_atomic_store(fl + blk, (uint64_t)1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

int launch_bblock_inv(hfpx_ctx* ctx, const float* A32, const BCrack* cr, const int32_t* tk, int64_t ntk,
                      const int64_t* bofs, float* invL, float* invU) {
  if (ntk <= 0) return HFPX_OK;
  hipLaunchKernelGGL(k_bblock_inv<true>, dim3((unsigned)ntk), dim3(SB), 0, ctx->stream, A32, cr, tk, bofs, invL);
  hipLaunchKernelGGL(k_bblock_inv<false>, dim3((unsigned)ntk), dim3(SB), 0, ctx->stream, A32, cr, tk, bofs, invU);
  HFPX_TRY_HIP(ctx, hipGetLastError());
  return HFPX_OK;
}

// flags: 2 ntk, then 2 tickets (uint64), then 2 error flags (int32); cleared
// here by a kernel (graph-safe)
int launch_bsweeps(hfpx_ctx* ctx, const float* A32, const BCrack* cr, const int32_t* tk, int64_t ntk,
                   const int64_t* bofs, const float* invL, const float* invU, float* vec, uint64_t* flags) {
  if (ntk <= 0) return HFPX_OK;
  uint64_t* tickets = flags + 2 * ntk;
  int32_t* errs = (int32_t*)(tickets + 2);
  const int64_t nclear = 2 * ntk + 3;
  hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)((nclear + 255) / 256)), dim3(256), 0, ctx->stream, flags, nclear);
  int occ = 0;
  HFPX_TRY_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_bsweep<true>, 2 * SB, 0));
  const unsigned nwg = (unsigned)std::min<int64_t>(ntk, std::max(1, occ) * std::max(1, ctx->num_cus));
  hipLaunchKernelGGL(k_bsweep<true>, dim3(nwg), dim3(2 * SB), 0, ctx->stream, A32, cr, tk, ntk, bofs, invL, vec,
                     flags, tickets, errs);
  hipLaunchKernelGGL(k_bsweep<false>, dim3(nwg), dim3(2 * SB), 0, ctx->stream, A32, cr, tk, ntk, bofs, invU, vec,
                     flags + ntk, tickets + 1, errs + 1);
  HFPX_TRY_HIP(ctx, hipGetLastError());
  return HFPX_OK;
}

size_t sweep_f32_work_bytes(int64_t n, int64_t nrhs) {
  const int swn = sweep_width(nrhs);
  const int64_t nblk = (n + SB - 1) / SB, gy = (nrhs + swn - 1) / swn;
  return sizeof(float) * 4 * (size_t)nblk * SB * SB + sizeof(uint64_t) * (2 * (size_t)nblk * gy + 4);
}

// B := U^{-1} L^{-1} B for the FP32 factors (row interchanges already
// applied to B); work holds sweep_f32_work_bytes(n, nrhs).
int sweep_f32(hfpx_ctx* ctx, const float* LU, int64_t lda, int64_t n, float* B, int64_t ldb, int64_t nrhs,
              void* work) {
  if (n <= 0 || nrhs <= 0) return HFPX_OK;
  const int swn = sweep_width(nrhs);
  const int64_t nblk = (n + SB - 1) / SB, gy = (nrhs + swn - 1) / swn;
  float* invL = (float*)work;
  float* invU = invL + (size_t)nblk * SB * SB;
  float* wL = invU + (size_t)nblk * SB * SB;  // W blocks of the forward sweep
  float* wU = wL + (size_t)nblk * SB * SB;    // and of the backward sweep
  uint64_t* flags = (uint64_t*)(wU + (size_t)nblk * SB * SB);  // [2][gy][nblk]
  uint64_t* tickets = flags + 2 * nblk * gy;                      // [2]
  int32_t* errs = (int32_t*)(tickets + 2);                       // [2]
  const int64_t nclear = 2 * nblk * gy + 4;
  hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)((nclear + 255) / 256)), dim3(256), 0, ctx->stream, flags, nclear);
  int tok = prof_begin(ctx, HFPX_KC_TRSM);
  hipLaunchKernelGGL(k_block_inv<true>, dim3((unsigned)nblk), dim3(SB), 0, ctx->stream, LU, lda, n, invL);
  hipLaunchKernelGGL(k_block_inv<false>, dim3((unsigned)nblk), dim3(SB), 0, ctx->stream, LU, lda, n, invU);
  hipLaunchKernelGGL(k_block_w<true>, dim3((unsigned)nblk), dim3(256), 0, ctx->stream, LU, lda, n, nblk, invL, wL);
  hipLaunchKernelGGL(k_block_w<false>, dim3((unsigned)nblk), dim3(256), 0, ctx->stream, LU, lda, n, nblk, invU, wU);
  prof_end(ctx, tok, HFPX_KC_TRSM, 2.0 * nblk * SB * SB * SB / 3.0 + 4.0 * nblk * SB * SB * SB, 0.0);
  HFPX_TRY_HIP(ctx, hipGetLastError());
  tok = prof_begin(ctx, HFPX_KC_TRSM);
  const int rc = swn == SWN_MIN
                     ? launch_sweeps<SWN_MIN>(ctx, LU, lda, n, nblk, invL, invU, wL, wU, B, ldb, nrhs, flags,
                                              tickets, errs)
                     : launch_sweeps<SWN_MAX>(ctx, LU, lda, n, nblk, invL, invU, wL, wU, B, ldb, nrhs, flags,
                                              tickets, errs);
  // algorithmic: 2 n^2 nrhs flops, the factor read once (4 n^2 bytes)
  prof_end(ctx, tok, HFPX_KC_TRSM, 2.0 * n * n * nrhs, 4.0 * n * n);
  if (rc) return rc;
  int32_t herr[2] = {0, 0};
  HFPX_TRY_HIP(ctx, hipMemcpyAsync(herr, errs, sizeof herr, hipMemcpyDeviceToHost, ctx->stream));
  HFPX_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (herr[0] || herr[1]) return set_error(ctx, HFPX_ERR_HIP, "getrs: sweep dependency wait timed out");
  return HFPX_OK;
}

}  // namespace hfpx
