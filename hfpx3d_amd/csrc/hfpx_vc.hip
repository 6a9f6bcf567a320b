// hfpx_vc.hip -- volume-control systems on gfx950: the border row and column
// that extend the influence matrix to the volume-control matrix, and the
// gather of a truncated volume-control system from a resident one.
//
// Replaces, on the device:
//   hfp3d::make_3dbem_matrix_vc  src/Solvers/system_assembly.cpp:458-557
//     (with hfp3d::el_p2_sf_integral, src/Core/element_utilities.cpp:573-595)
//   hfp3d::mod_3dbem_system_vc   src/Solvers/system_assembly.cpp:560-629
//
// Layout in HBM: the volume-control matrix is the (n+1) x (n+1) column-major
// matrix [A, c; v^T, 0] with the n x n influence matrix A (as written by the
// assembly), row n = v (opening volume per unit DD) and column n = c
// (traction per unit pressure).  The truncated system is written the same
// way, FP64 and optionally its FP32 rounding in the same pass.
#include <algorithm>
#include <climits>
#include <cstring>
#include <vector>

#include "hfpx_internal.h"

namespace hfpx {

// Border: one thread per element.  For the free slot (node ns, component j)
// of element e with DoF d:
//   A(n, d) = (int N_ns dA) e3_j,   A(d, n) = -e3_j,
// e3 the element normal (row 2 of its rotation tensor).  The area integrals
// of the 6-node triangle's shape functions are 0 for the vertex nodes and
// area / 3 for the mid-edge nodes of the uniform element (closed forms for
// vertex weights below); the reference evaluates the same integrals as
// contour integrals of the monomials.  Thread 0 also writes the corner
// A(n, n) = 0 (the reference leaves it unset).  A DoF shared between
// elements keeps one of their values, as in the reference's OpenMP loop.
// With DD in element frames (dd_local) the volume row holds the area
// integral on the opening component only (system_assembly.cpp:529-536).
__global__ void k_vc_border(const SrcElem<double>* __restrict__ elem, int64_t n_el,
                            const int32_t* __restrict__ dof, int64_t n, bool dd_local,
                            double* __restrict__ a64, int64_t lda64, float* __restrict__ a32, int64_t lda32) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e == 0) {
    if (a64) a64[n + n * lda64] = 0.0;
    if (a32) a32[n + n * lda32] = 0.0f;
  }
  if (e >= n_el) return;
  const SrcElem<double>& E = elem[e];
  const double area = 0.5 * ((E.t[2] - E.t[0]) * (E.t[5] - E.t[1]) - (E.t[3] - E.t[1]) * (E.t[4] - E.t[0]));
  for (int ns = 0; ns < 6; ++ns) {
    // int N dA of the quadratic basis: int L_a L_b dA = A (1 + delta_ab) / 12,
    // so N_v gives A (2 - sw[2v] - sw[2v+1]) / 12 (0 for the uniform
    // element) and N_{3+v} gives sw[6+v] A / 12 (A / 3 uniform)
    const double sf = ns < 3 ? (2.0 - E.sw[2 * ns] - E.sw[2 * ns + 1]) * area / 12.0 : E.sw[3 + ns] * area / 12.0;
    for (int j = 0; j < 3; ++j) {
      const int l = 3 * ns + j;
      const int64_t d = dof ? (int64_t)dof[18 * e + l] : 18 * e + l;
      if (d < 0) continue;
      const double nj = E.R[6 + j];
      // local DD: only the opening component (2) carries volume
      const double vol = dd_local ? (j == 2 ? sf : 0.0) : sf * nj;
      if (a64) {
        a64[n + d * lda64] = vol;
        a64[d + n * lda64] = -nj;
      }
      if (a32) {
        a32[n + d * lda32] = (float)vol;
        a32[d + n * lda32] = (float)(-nj);
      }
    }
  }
}

// Truncated system: M(t, s) = sum over b in list(s), then a in list(t), of
// A(a, b) -- the reference's summation order -- where list(u) holds the
// original DoF of the used slots mapped to u in slot order and list(nu - 1)
// the border index; M(nu - 1, nu - 1) = 0.  A thread owns row t of a strided
// set of columns; a wavefront writes 64 consecutive rows of a column.  F64 /
// F32 select the outputs written (M and / or M32).
template <bool F64, bool F32>
__global__ __launch_bounds__(256) void k_vc_gather(const double* __restrict__ A, int64_t lda,
                                                   const int32_t* __restrict__ off,
                                                   const int32_t* __restrict__ idx, int64_t nu,
                                                   double* __restrict__ M, int64_t ldm,
                                                   float* __restrict__ M32, int64_t ldm32) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= nu) return;
  const int32_t t0 = off[t], t1 = off[t + 1];
  const bool single = t1 == t0 + 1;
  const int64_t a0 = single ? (int64_t)idx[t0] : 0;
  for (int64_t s = blockIdx.y; s < nu; s += gridDim.y) {
    const int32_t s0 = off[s], s1 = off[s + 1];
    double v = 0.0;
    if (t != nu - 1 || s != nu - 1) {
      for (int32_t b = s0; b < s1; ++b) {
        const double* col = A + (int64_t)idx[b] * lda;
        if (single) {
          v += col[a0];
        } else {
          for (int32_t a = t0; a < t1; ++a) v += col[idx[a]];
        }
      }
    }
    if (F64) M[t + s * ldm] = v;
    if (F32) M32[t + s * ldm32] = (float)v;
  }
}

// The common case, every used DoF from exactly one slot: M(t, s) =
// A(idx[t], idx[s]) with idx[nu - 1] the border index and M(nu - 1, nu - 1)
// = 0.  Four columns per trip keep four independent loads in flight per lane;
// the outputs are written with non-temporal stores (streamed, not re-read by
// this kernel).
template <bool F64, bool F32>
__global__ __launch_bounds__(256) void k_vc_gather1(const double* __restrict__ A, int64_t lda,
                                                    const int32_t* __restrict__ idx, int64_t nu,
                                                    double* __restrict__ M, int64_t ldm,
                                                    float* __restrict__ M32, int64_t ldm32) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= nu) return;
  const double* At = A + idx[t];
  const int64_t G = gridDim.y;
  int64_t s = blockIdx.y;
  for (; s + 3 * G < nu; s += 4 * G) {
    double v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = At[(int64_t)idx[s + k * G] * lda];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int64_t c = s + k * G;
      const double x = (t == nu - 1 && c == nu - 1) ? 0.0 : v[k];
      if (F64) __builtin_nontemporal_store(x, M + t + c * ldm);
      if (F32) __builtin_nontemporal_store((float)x, M32 + t + c * ldm32);
    }
  }
  for (; s < nu; s += G) {
    const double x = (t == nu - 1 && s == nu - 1) ? 0.0 : At[(int64_t)idx[s] * lda];
    if (F64) __builtin_nontemporal_store(x, M + t + s * ldm);
    if (F32) __builtin_nontemporal_store((float)x, M32 + t + s * ldm32);
  }
}

// M64 / M32 (either may be null) from A by the device list idx (nu
// entries, the last the border index), one slot per DoF
int launch_vc_gather_idx(hfpx_ctx* ctx, const double* A, int64_t lda, const int32_t* idx, int64_t nu, double* M64,
                         int64_t ldm64, float* M32, int64_t ldm32) {
  const dim3 grid((unsigned)((nu + 255) / 256), (unsigned)std::min<int64_t>(nu, 32768));
  if (M64 && M32)
    hipLaunchKernelGGL((k_vc_gather1<true, true>), grid, dim3(256), 0, ctx->stream, A, lda, idx, nu, M64, ldm64, M32,
                       ldm32);
  else if (M64)
    hipLaunchKernelGGL((k_vc_gather1<true, false>), grid, dim3(256), 0, ctx->stream, A, lda, idx, nu, M64, ldm64,
                       M32, ldm32);
  else
    hipLaunchKernelGGL((k_vc_gather1<false, true>), grid, dim3(256), 0, ctx->stream, A, lda, idx, nu, M64, ldm64,
                       M32, ldm32);
  HFPX_TRY_HIP(ctx, hipGetLastError());
  return HFPX_OK;
}

int launch_vc_border(hfpx_ctx* ctx, const hfpx_mesh* m, double* a64, int64_t lda64, float* a32,
                     int64_t lda32) {
  const int bs = 128;
  hipLaunchKernelGGL(k_vc_border, dim3((unsigned)((m->n_el + bs - 1) / bs)), dim3(bs), 0, ctx->stream, m->elem,
                     m->n_el, m->identity_dof ? nullptr : m->dof, m->n_dof, m->dd_local, a64, lda64, a32, lda32);
  HFPX_TRY_HIP(ctx, hipGetLastError());
  return HFPX_OK;
}

}  // namespace hfpx

using namespace hfpx;

extern "C" {

int hfpx_assemble_vc(hfpx_ctx* ctx, hfpx_mesh* m, double shear_m, double poiss_r, double* A64, int64_t lda64,
                     float* A32, int64_t lda32) {
  if (!ctx || !m) return HFPX_ERR_INVALID;
  if (!A64 && !A32) return set_error(ctx, HFPX_ERR_INVALID, "no output matrix");
  const int64_t n = m->n_dof;
  if ((A64 && lda64 < n + 1) || (A32 && lda32 < n + 1))
    return set_error(ctx, HFPX_ERR_INVALID, "lda smaller than n_dof + 1 = %lld", (long long)(n + 1));
  if (!(poiss_r < 1.0) || !(shear_m > 0.0))
    return set_error(ctx, HFPX_ERR_INVALID, "need shear modulus > 0 and Poisson ratio < 1");
  HFPX_TRY_HIP(ctx, hipSetDevice(ctx->device));
  int rc = launch_geometry(ctx, m);
  if (rc) return rc;
  rc = launch_assemble(ctx, m, shear_m, poiss_r, A64, lda64, A32, lda32);
  if (rc) return rc;
  return launch_vc_border(ctx, m, A64, lda64, A32, lda32);
}

int hfpx_vc_truncate(hfpx_ctx* ctx, const double* A, int64_t lda, int64_t orig_ndof,
                     const int64_t* orig_dof_h_host, const int64_t* dof_h_host, int64_t n_el, int64_t ndpe,
                     int64_t used_ndof, const double* delta_t_host, int64_t tsize, double delta_v, double* M64,
                     int64_t ldm64, float* M32, int64_t ldm32, double* rhs) {
  if (!ctx) return HFPX_ERR_INVALID;
  if (!A || !orig_dof_h_host || !dof_h_host || !rhs || (!M64 && !M32))
    return set_error(ctx, HFPX_ERR_INVALID, "vc_truncate: null argument");
  if (n_el < 1 || ndpe < 1 || orig_ndof < 1 || used_ndof < 1 || used_ndof > orig_ndof)
    return set_error(ctx, HFPX_ERR_INVALID, "vc_truncate: need n_el, ndpe >= 1 and 1 <= used_ndof <= orig_ndof");
  const int64_t nu = used_ndof + 1, full = n_el * ndpe;
  if (lda < orig_ndof + 1 || (M64 && ldm64 < nu) || (M32 && ldm32 < nu))
    return set_error(ctx, HFPX_ERR_INVALID, "vc_truncate: leading dimension too small");
  if (orig_ndof >= INT32_MAX || full >= INT32_MAX - 1)
    return set_error(ctx, HFPX_ERR_INVALID, "vc_truncate: more than 2^31 DoF");
  if (tsize != full && tsize != orig_ndof && tsize != used_ndof)
    return set_error(ctx, HFPX_ERR_INVALID, "vc_truncate: delta_t has %lld entries, not n_el * ndpe, orig_ndof or "
                     "used_ndof", (long long)tsize);
  if (!delta_t_host) return set_error(ctx, HFPX_ERR_INVALID, "vc_truncate: null delta_t");
  // slot lists per used DoF (slot order: element-major, as the reference's loops)
  std::vector<int32_t> off((size_t)nu + 1, 0);
  for (int64_t k = 0; k < full; ++k) {
    const int64_t u = dof_h_host[k], o = orig_dof_h_host[k];
    if (u < -1 || u >= used_ndof || o < -1 || o >= orig_ndof)
      return set_error(ctx, HFPX_ERR_INVALID, "vc_truncate: DoF index out of range in slot %lld", (long long)k);
    if (u >= 0 && o < 0)
      return set_error(ctx, HFPX_ERR_INVALID, "vc_truncate: slot %lld is used but fixed in the original handle",
                       (long long)k);
    if (u >= 0) ++off[(size_t)u + 1];
  }
  off[(size_t)nu] = 1;  // the border
  for (int64_t u = 0; u < nu; ++u) off[(size_t)u + 1] += off[(size_t)u];
  const int64_t total = off[(size_t)nu];
  bool one_slot_each = true;
  for (int64_t u = 0; u < nu; ++u) one_slot_each = one_slot_each && off[(size_t)u + 1] - off[(size_t)u] == 1;
  std::vector<int32_t> idx((size_t)total), fill(off.begin(), off.end() - 1);
  std::vector<double> r((size_t)nu, 0.0);
  for (int64_t e = 0; e < n_el; ++e)
    for (int64_t j = 0; j < ndpe; ++j) {
      const int64_t k = e + n_el * j, u = dof_h_host[k];
      if (u < 0) continue;
      const int64_t o = orig_dof_h_host[k];
      idx[(size_t)fill[(size_t)u]++] = (int32_t)o;
      r[(size_t)u] = delta_t_host[tsize == full ? e * ndpe + j : tsize == orig_ndof ? o : u];
    }
  idx[(size_t)total - 1] = (int32_t)orig_ndof;
  r[(size_t)used_ndof] = delta_v;

  HFPX_TRY_HIP(ctx, hipSetDevice(ctx->device));
  const size_t rb = sizeof(double) * (size_t)nu, ob = sizeof(int32_t) * off.size(),
               ib = sizeof(int32_t) * idx.size();
  const size_t bytes = rb + ob + ib;
  int rc = ensure_hwork(ctx, bytes);
  if (rc) return rc;
  rc = ensure_work(ctx, bytes);
  if (rc) return rc;
  char* h = (char*)ctx->hwork;
  memcpy(h, r.data(), rb);
  memcpy(h + rb, off.data(), ob);
  memcpy(h + rb + ob, idx.data(), ib);
  HFPX_TRY_HIP(ctx, hipMemcpyAsync(ctx->work, ctx->hwork, bytes, hipMemcpyHostToDevice, ctx->stream));
  HFPX_TRY_HIP(ctx, hipMemcpyAsync(rhs, ctx->work, rb, hipMemcpyDeviceToDevice, ctx->stream));
  const int32_t* d_off = (const int32_t*)((char*)ctx->work + rb);
  const int32_t* d_idx = (const int32_t*)((char*)ctx->work + rb + ob);
  const dim3 grid((unsigned)((nu + 255) / 256), (unsigned)std::min<int64_t>(nu, 32768));
  const int tok = prof_begin(ctx, HFPX_KC_VC);
  if (one_slot_each) {
    if (M64 && M32)
      hipLaunchKernelGGL((k_vc_gather1<true, true>), grid, dim3(256), 0, ctx->stream, A, lda, d_idx, nu, M64, ldm64,
                         M32, ldm32);
    else if (M64)
      hipLaunchKernelGGL((k_vc_gather1<true, false>), grid, dim3(256), 0, ctx->stream, A, lda, d_idx, nu, M64, ldm64,
                         M32, ldm32);
    else
      hipLaunchKernelGGL((k_vc_gather1<false, true>), grid, dim3(256), 0, ctx->stream, A, lda, d_idx, nu, M64, ldm64,
                         M32, ldm32);
  } else if (M64 && M32)
    hipLaunchKernelGGL((k_vc_gather<true, true>), grid, dim3(256), 0, ctx->stream, A, lda, d_off, d_idx, nu, M64,
                       ldm64, M32, ldm32);
  else if (M64)
    hipLaunchKernelGGL((k_vc_gather<true, false>), grid, dim3(256), 0, ctx->stream, A, lda, d_off, d_idx, nu, M64,
                       ldm64, M32, ldm32);
  else
    hipLaunchKernelGGL((k_vc_gather<false, true>), grid, dim3(256), 0, ctx->stream, A, lda, d_off, d_idx, nu, M64,
                       ldm64, M32, ldm32);
  HFPX_TRY_HIP(ctx, hipGetLastError());
  // algorithmic HBM bytes: every entry of the truncated system read once from
  // A and written once (FP64 and/or FP32)
  const double entries = (double)nu * (double)nu;
  prof_end(ctx, tok, HFPX_KC_VC, 0.0, entries * (8.0 + (M64 ? 8.0 : 0.0) + (M32 ? 4.0 : 0.0)));
  // hwork/work are reused by the next call: finish before returning
  HFPX_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return HFPX_OK;
}

}  // extern "C"
