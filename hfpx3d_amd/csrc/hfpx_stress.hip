// hfpx_stress.hip -- stress postprocessing at observation points on gfx950.
//
// Replaces hfp3d::make_3dbem_stress_f_s (src/Solvers/system_assembly.cpp:
// 279-455): sigma(x) = s_inf + sum over elements of R^T S(x) R applied to
// the element's nodal DD, with S the 6x18 local influence.  Unlike the
// reference (which accumulates from several OpenMP threads into the same
// rows without synchronisation, system_assembly.cpp:415-417), the sum is
// formed in a fixed order, so the result is deterministic.
//
// Work decomposition: a wavefront = 64 observation points x one chunk of
// elements; all lanes visit the same element at the same time (element data
// and DD are wave-uniform).  Partial sums per chunk go to a workspace and a
// second kernel adds the chunks in order plus s_inf.
#include <mutex>

#include "hfpx_internal.h"

namespace hfpx {

// DD of every node in its element's frame: dloc[3*(6e+j)+k] = (R_e d_glob)_k,
// or the given DD unchanged when they are already local (is_dd_local,
// system_assembly.cpp:358-387)
__global__ void k_dd_local(const SrcElem<double>* __restrict__ elem, const double* __restrict__ dd,
                           int64_t ldd, int64_t n_el, bool already_local, double* __restrict__ dloc) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 6 * n_el) return;
  const int64_t e = t / 6;
  const double g[3] = {dd[t], dd[t + ldd], dd[t + 2 * ldd]};
  const double* R = elem[e].R;
  for (int k = 0; k < 3; ++k)
    dloc[3 * t + k] = already_local ? g[k] : R[3 * k] * g[0] + R[3 * k + 1] * g[1] + R[3 * k + 2] * g[2];
}

template <bool NU>
__global__ __launch_bounds__(256) void k_stress(const SrcElem<double>* __restrict__ elem,
                                                const double* __restrict__ dloc, int64_t n_el,
                                                const double* __restrict__ pts, int64_t n_pts,
                                                int64_t chunk, double shear_m, double nu,
                                                double* __restrict__ part) {
  const int lane = threadIdx.x & 63;
  const int64_t p = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * 64 + lane;
  const int64_t c = blockIdx.y;
  const int64_t e0 = c * chunk, e1 = min(n_el, e0 + chunk);
  const bool active = p < n_pts;
  const int64_t pp = active ? p : 0;
  const double x[3] = {pts[3 * pp], pts[3 * pp + 1], pts[3 * pp + 2]};
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int64_t e = e0; e < e1; ++e) {
    const SrcElem<double>& E = elem[e];
    const double d[3] = {x[0] - E.v0[0], x[1] - E.v0[1], x[2] - E.v0[2]};
    const double y0 = E.R[0] * d[0] + E.R[1] * d[1] + E.R[2] * d[2];
    const double y1 = E.R[3] * d[0] + E.R[4] * d[1] + E.R[5] * d[2];
    const double y2 = E.R[6] * d[0] + E.R[7] * d[1] + E.R[8] * d[2];
    double sl[6] = {0, 0, 0, 0, 0, 0};
    const double* de = dloc + 18 * e;
    pair_influence<double, NU>(E, y0, y1, y2, shear_m, nu, [&](int j, const double (&S)[6][3]) {
      const double a = de[3 * j], b = de[3 * j + 1], cc = de[3 * j + 2];
#pragma unroll
      for (int q = 0; q < 6; ++q) sl[q] += S[q][0] * a + S[q][1] * b + S[q][2] * cc;
    });
    // R^T s R with s = [[s11 s12 s13] [s12 s22 s23] [s13 s23 s33]]
    const double m[3][3] = {{sl[0], sl[3], sl[4]}, {sl[3], sl[1], sl[5]}, {sl[4], sl[5], sl[2]}};
    double mr[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int jj = 0; jj < 3; ++jj) mr[a][jj] = m[a][0] * E.R[jj] + m[a][1] * E.R[3 + jj] + m[a][2] * E.R[6 + jj];
    double g[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int jj = 0; jj < 3; ++jj) g[i][jj] = E.R[i] * mr[0][jj] + E.R[3 + i] * mr[1][jj] + E.R[6 + i] * mr[2][jj];
    acc[0] += g[0][0];
    acc[1] += g[1][1];
    acc[2] += g[2][2];
    acc[3] += g[0][1];
    acc[4] += g[0][2];
    acc[5] += g[1][2];
  }
  if (active)
    for (int q = 0; q < 6; ++q) part[(c * 6 + q) * n_pts + p] = acc[q];
}

// the far-field stress, passed by value (no device copy to stage)
struct Sinf6 {
  double v[6];
};

__global__ void k_stress_sum(const double* __restrict__ part, int64_t nchunk, int64_t n_pts, Sinf6 sinf,
                             double* __restrict__ out, int64_t ldo) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_pts * 6) return;
  const int64_t p = t % n_pts, q = t / n_pts;
  double s = 0.0;
  for (int64_t c = 0; c < nchunk; ++c) s += part[(c * 6 + q) * n_pts + p];
  out[p + q * ldo] = s + sinf.v[q];
}

// Chunks of elements: enough wavefronts to fill the device.
static int64_t stress_chunks(int64_t n_el, int64_t n_pts) {
  const int64_t pblocks = (n_pts + 255) / 256;
  int64_t nchunk = (8 * 1024 + pblocks - 1) / pblocks;
  nchunk = std::max<int64_t>(1, std::min<int64_t>(nchunk, (n_el + 7) / 8));
  return std::min<int64_t>(nchunk, 65535);
}

// Doubles of workspace for nchunk chunks: the DD in element frames
// (18 n_el) and the per-chunk partial sums (6 nchunk n_pts).
static size_t stress_work_doubles(int64_t n_el, int64_t n_pts, int64_t nchunk) {
  return 18 * (size_t)n_el + 6 * (size_t)nchunk * (size_t)n_pts;
}

// Enqueue the stress evaluation with its workspace `work` (stress_work_doubles
// of nchunk); no allocation, no synchronisation.
static int stress_enqueue(hfpx_ctx* ctx, hfpx_mesh* m, double shear_m, double nu, const double* dd, int64_t ldd,
                          const double* pts, int64_t n_pts, const double* sinf_host, double* out, int64_t ldo,
                          double* work, int64_t nchunk, bool profile) {
  int rc = launch_geometry(ctx, m);
  if (rc) return rc;
  const int64_t pblocks = (n_pts + 255) / 256;
  const int64_t chunk = (m->n_el + nchunk - 1) / nchunk;
  nchunk = (m->n_el + chunk - 1) / chunk;
  double* dloc = work;
  double* part = dloc + 18 * m->n_el;
  Sinf6 sv;
  for (int q = 0; q < 6; ++q) sv.v[q] = sinf_host[q];
  hipLaunchKernelGGL(k_dd_local, dim3((unsigned)((6 * m->n_el + 255) / 256)), dim3(256), 0, ctx->stream,
                     m->elem, dd, ldd, m->n_el, m->dd_local, dloc);
  const int tok = profile ? prof_begin(ctx, HFPX_KC_STRESS) : -1;
  if (m->wts)
    hipLaunchKernelGGL(k_stress<true>, dim3((unsigned)pblocks, (unsigned)nchunk), dim3(256), 0, ctx->stream,
                       m->elem, dloc, m->n_el, pts, n_pts, chunk, shear_m, nu, part);
  else
    hipLaunchKernelGGL(k_stress<false>, dim3((unsigned)pblocks, (unsigned)nchunk), dim3(256), 0, ctx->stream,
                       m->elem, dloc, m->n_el, pts, n_pts, chunk, shear_m, nu, part);
  if (profile) prof_end(ctx, tok, HFPX_KC_STRESS, 0.0, 0.0);
  hipLaunchKernelGGL(k_stress_sum, dim3((unsigned)((6 * n_pts + 255) / 256)), dim3(256), 0, ctx->stream, part,
                     nchunk, n_pts, sv, out, ldo);
  HFPX_TRY_HIP(ctx, hipGetLastError());
  return HFPX_OK;
}

int launch_stress(hfpx_ctx* ctx, hfpx_mesh* m, double shear_m, double nu, const double* dd, int64_t ldd,
                  const double* pts, int64_t n_pts, const double* sinf_host, double* out, int64_t ldo) {
  const int64_t nchunk = stress_chunks(m->n_el, n_pts);
  int rc = ensure_work(ctx, sizeof(double) * stress_work_doubles(m->n_el, n_pts, nchunk) + 512);
  if (rc) return rc;
  if ((rc = stress_enqueue(ctx, m, shear_m, nu, dd, ldd, pts, n_pts, sinf_host, out, ldo, (double*)ctx->work,
                           nchunk, true)))
    return rc;
  HFPX_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));  // the workspace is reused by the next call
  return HFPX_OK;
}

// ---------------------------------------------------------------------
// Allocation-free form (hfpx_stress_noalloc).  Its workspace is this
// file-scope device array: one copy per GPU, held until the process ends.
// Host code never uses the array's name as an address (on the host it is
// only a stand-in): the device address comes from hipGetSymbolAddress, taken
// once per device by stress_scratch_address with that device current.
// ---------------------------------------------------------------------
constexpr size_t STRESS_SCRATCH_DOUBLES = (size_t)8 << 20;  // 64 MiB
static __device__ __align__(256) double scratch[STRESS_SCRATCH_DOUBLES];

size_t stress_scratch_doubles() { return STRESS_SCRATCH_DOUBLES; }

int stress_scratch_address(hfpx_ctx* ctx, double** out) {
  constexpr int MAX_DEV = 64;
  static std::mutex mu;
  static double* addr[MAX_DEV] = {nullptr};
  *out = nullptr;
  if (ctx->device < 0 || ctx->device >= MAX_DEV)
    return set_error(ctx, HFPX_ERR_INVALID, "stress scratch: device %d out of range", ctx->device);
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[ctx->device]) {
    HFPX_TRY_HIP(ctx, hipSetDevice(ctx->device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return check_hip(ctx, e != hipSuccess ? e : hipErrorInvalidSymbol, "stress scratch");
    addr[ctx->device] = (double*)p;
  }
  *out = addr[ctx->device];
  return HFPX_OK;
}

int launch_stress_noalloc(hfpx_ctx* ctx, hfpx_mesh* m, double shear_m, double nu, const double* dd, int64_t ldd,
                          const double* pts, int64_t n_pts, const double* sinf_host, double* out, int64_t ldo) {
  double* work = ctx->stress_scratch;
  if (!work) return set_error(ctx, HFPX_ERR_HIP, "stress_noalloc: the scratch array has no device address");
  // as many chunks as the allocating form, fewer if the partial sums would
  // not fit the scratch array; refuse what does not fit with one chunk
  int64_t nchunk = stress_chunks(m->n_el, n_pts);
  const size_t fixed = 18 * (size_t)m->n_el;
  if (fixed + 6 * (size_t)n_pts > STRESS_SCRATCH_DOUBLES)
    return set_error(ctx, HFPX_ERR_INVALID,
                     "stress_noalloc: 18 n_el + 6 n_pts = %zu doubles exceed the %zu-double scratch array",
                     fixed + 6 * (size_t)n_pts, STRESS_SCRATCH_DOUBLES);
  nchunk = std::min<int64_t>(nchunk, (int64_t)((STRESS_SCRATCH_DOUBLES - fixed) / (6 * (size_t)n_pts)));
  return stress_enqueue(ctx, m, shear_m, nu, dd, ldd, pts, n_pts, sinf_host, out, ldo, work, nchunk, false);
}

}  // namespace hfpx
