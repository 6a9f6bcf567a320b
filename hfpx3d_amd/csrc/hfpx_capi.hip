// hfpx_capi.hip -- extern "C" entry points: context, mesh, assembly, RHS.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "hfpx_internal.h"

namespace hfpx {

int set_error(hfpx_ctx* ctx, int code, const char* fmt, ...) {
  if (ctx) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    ctx->last_error = buf;
  }
  return code;
}

int check_hip(hfpx_ctx* ctx, hipError_t e, const char* what) {
  if (e == hipSuccess) return HFPX_OK;
  const int code = (e == hipErrorOutOfMemory) ? HFPX_ERR_NOMEM : HFPX_ERR_HIP;
  return set_error(ctx, code, "%s: %s", what, hipGetErrorString(e));
}

int ensure_work(hfpx_ctx* ctx, size_t bytes) {
  if (ctx->work_bytes >= bytes) return HFPX_OK;
  if (ctx->work) {
    HFPX_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    HFPX_TRY_HIP(ctx, hipFree(ctx->work));
    ctx->work = nullptr;
    ctx->work_bytes = 0;
  }
  HFPX_TRY_HIP(ctx, hipMalloc(&ctx->work, bytes));
  ctx->work_bytes = bytes;
  return HFPX_OK;
}

int ensure_hwork(hfpx_ctx* ctx, size_t bytes) {
  if (ctx->hwork_bytes >= bytes) return HFPX_OK;
  if (ctx->hwork) {
    HFPX_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    HFPX_TRY_HIP(ctx, hipHostFree(ctx->hwork));
    ctx->hwork = nullptr;
    ctx->hwork_bytes = 0;
  }
  HFPX_TRY_HIP(ctx, hipHostMalloc(&ctx->hwork, bytes, hipHostMallocDefault));
  ctx->hwork_bytes = bytes;
  return HFPX_OK;
}

int prof_begin(hfpx_ctx* ctx, int cls) {
  if (!ctx->prof || !(ctx->prof_mask & (1u << cls))) return -1;
  hipEvent_t a, b;
  if (ctx->prof_pool.size() >= 2) {
    a = ctx->prof_pool.back();
    ctx->prof_pool.pop_back();
    b = ctx->prof_pool.back();
    ctx->prof_pool.pop_back();
  } else {
    if (hipEventCreate(&a) != hipSuccess) return -1;
    if (hipEventCreate(&b) != hipSuccess) {
      (void)hipEventDestroy(a);
      return -1;
    }
  }
  (void)hipEventRecord(a, ctx->stream);
  ctx->prof_recs.push_back({-1, a, b, 0.0, 0.0});
  return (int)ctx->prof_recs.size() - 1;
}

void prof_end(hfpx_ctx* ctx, int token, int cls, double flops, double bytes) {
  if (token < 0 || token >= (int)ctx->prof_recs.size()) return;
  auto& r = ctx->prof_recs[token];
  (void)hipEventRecord(r.b, ctx->stream);
  r.cls = cls;
  r.flops = flops;
  r.bytes = bytes;
}

static void prof_clear(hfpx_ctx* ctx) {
  for (auto& r : ctx->prof_recs) {
    ctx->prof_pool.push_back(r.a);
    ctx->prof_pool.push_back(r.b);
  }
  ctx->prof_recs.clear();
}

}  // namespace hfpx

using namespace hfpx;

extern "C" {

const char* hfpx_status_string(int status) {
  switch (status) {
    case HFPX_OK: return "ok";
    case HFPX_ERR_INVALID: return "invalid argument";
    case HFPX_ERR_NOMEM: return "out of memory";
    case HFPX_ERR_SINGULAR: return "matrix is singular";
    case HFPX_ERR_NOT_CONVERGED: return "iterative refinement did not converge";
    case HFPX_ERR_HIP: return "HIP runtime error";
    default: return "unknown status";
  }
}

const char* hfpx_last_error(const hfpx_ctx* ctx) { return ctx ? ctx->last_error.c_str() : ""; }

int hfpx_create(hfpx_ctx** out, int device, void* stream) {
  if (!out) return HFPX_ERR_INVALID;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return HFPX_ERR_HIP;
  if (device < 0 || device >= ndev) return HFPX_ERR_INVALID;
  hfpx_ctx* ctx = new hfpx_ctx();
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess) {
    delete ctx;
    return HFPX_ERR_HIP;
  }
  if (stream) {
    ctx->stream = (hipStream_t)stream;
  } else {
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
      delete ctx;
      return HFPX_ERR_HIP;
    }
    ctx->own_stream = true;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) ctx->num_cus = prop.multiProcessorCount;
  if (hipMalloc((void**)&ctx->sync, 256) != hipSuccess || hipMemset(ctx->sync, 0, 256) != hipSuccess) {
    if (ctx->sync) (void)hipFree(ctx->sync);
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return HFPX_ERR_HIP;
  }
  // the allocation-free stress workspace's device address (once per device);
  // a failure here only disables hfpx_stress_noalloc
  if (stress_scratch_address(ctx, &ctx->stress_scratch) != HFPX_OK) ctx->stress_scratch = nullptr;
  *out = ctx;
  return HFPX_OK;
}

int hfpx_destroy(hfpx_ctx* ctx) {
  if (!ctx) return HFPX_OK;
  (void)hipSetDevice(ctx->device);
  // a borrowed stream belongs to the caller, who finishes its work before
  // destroying the context; hipFree below waits for the device either way
  if (ctx->own_stream && ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->work) (void)hipFree(ctx->work);
  if (ctx->lu_work) (void)hipFree(ctx->lu_work);
  if (ctx->gemm_work) (void)hipFree(ctx->gemm_work);
  if (ctx->x3_work) (void)hipFree(ctx->x3_work);
  if (ctx->x3_work2) (void)hipFree(ctx->x3_work2);
  if (ctx->stream2) {
    (void)hipStreamSynchronize(ctx->stream2);
    (void)hipStreamDestroy(ctx->stream2);
  }
  for (hipEvent_t e : ctx->la_ev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->sync) (void)hipFree(ctx->sync);
  if (ctx->hwork) (void)hipHostFree(ctx->hwork);
  prof_clear(ctx);
  for (auto e : ctx->prof_pool) (void)hipEventDestroy(e);
  if (ctx->own_stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return HFPX_OK;
}

void* hfpx_stream(const hfpx_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int hfpx_synchronize(hfpx_ctx* ctx) {
  if (!ctx) return HFPX_ERR_INVALID;
  HFPX_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return HFPX_OK;
}

int hfpx_mesh_create(hfpx_ctx* ctx, const double* nods_host, int64_t nods_rows, int64_t n_nodes,
                     const int64_t* conn_host, int64_t conn_rows, int64_t n_el,
                     const int64_t* dof_h_host, int64_t n_dof, double beta, hfpx_mesh** out) {
  if (!ctx || !out) return HFPX_ERR_INVALID;
  *out = nullptr;
  if (!nods_host || !conn_host || nods_rows < 3 || conn_rows < 3 || n_nodes < 3 || n_el < 1)
    return set_error(ctx, HFPX_ERR_INVALID, "mesh needs >= 3 coordinate rows, >= 3 nodes, >= 1 element");
  if (n_el * 6 > (int64_t)1 << 30)
    return set_error(ctx, HFPX_ERR_INVALID, "n_el = %lld too large", (long long)n_el);
  std::vector<double> xyz(3 * n_nodes);
  for (int64_t j = 0; j < n_nodes; ++j)
    for (int k = 0; k < 3; ++k) {
      const double v = nods_host[k + nods_rows * j];
      if (!std::isfinite(v))
        return set_error(ctx, HFPX_ERR_INVALID, "node %lld has a non-finite coordinate", (long long)j);
      xyz[3 * j + k] = v;
    }
  std::vector<int32_t> cn(3 * n_el);
  for (int64_t e = 0; e < n_el; ++e)
    for (int k = 0; k < 3; ++k) {
      const int64_t v = conn_host[k + conn_rows * e];
      if (v < 0 || v >= n_nodes)
        return set_error(ctx, HFPX_ERR_INVALID, "element %lld references node %lld out of range",
                         (long long)e, (long long)v);
      cn[3 * e + k] = (int32_t)v;
    }
  // degenerate elements have no frame: the geometry would fill the matrix
  // with inf / NaN
  for (int64_t e = 0; e < n_el; ++e)
    if (degenerate_triangle(&xyz[3 * cn[3 * e]], &xyz[3 * cn[3 * e + 1]], &xyz[3 * cn[3 * e + 2]]))
      return set_error(ctx, HFPX_ERR_INVALID, "element %lld is degenerate (zero area or a needle)", (long long)e);
  std::vector<int32_t> dh;
  bool identity = (dof_h_host == nullptr);
  if (identity) {
    n_dof = 18 * n_el;
  } else {
    dh.resize(18 * n_el);
    std::vector<char> seen(n_dof > 0 ? n_dof : 1, 0);
    identity = (n_dof == 18 * n_el);
    for (int64_t e = 0; e < n_el; ++e)
      for (int c = 0; c < 18; ++c) {
        const int64_t d = dof_h_host[e + n_el * c];
        if (d < -1 || d >= n_dof)
          return set_error(ctx, HFPX_ERR_INVALID, "dof_h(%lld, %d) = %lld outside [-1, n_dof)",
                           (long long)e, c, (long long)d);
        if (d >= 0) {
          if (seen[d]) return set_error(ctx, HFPX_ERR_INVALID, "dof %lld used twice", (long long)d);
          seen[d] = 1;
        }
        if (d != 18 * e + c) identity = false;
        dh[18 * e + c] = (int32_t)d;
      }
    for (int64_t d = 0; d < n_dof; ++d)
      if (!seen[d]) return set_error(ctx, HFPX_ERR_INVALID, "dof %lld never used", (long long)d);
  }
  HFPX_TRY_HIP(ctx, hipSetDevice(ctx->device));
  hfpx_mesh* m = new hfpx_mesh();
  m->n_el = n_el;
  m->n_dof = n_dof;
  m->beta = beta;
  m->identity_dof = identity;
  m->n_nodes = n_nodes;
  hipError_t e = hipMalloc(&m->nods, sizeof(double) * 3 * n_nodes);
  if (e == hipSuccess) e = hipMalloc(&m->conn, sizeof(int32_t) * 3 * n_el);
  if (e == hipSuccess) e = hipMalloc(&m->dof, sizeof(int32_t) * 18 * n_el);
  if (e == hipSuccess) e = hipMalloc(&m->elem, sizeof(SrcElem<double>) * n_el);
  if (e == hipSuccess) e = hipMalloc(&m->cp, sizeof(double) * 36 * n_el);
  if (e == hipSuccess)
    e = hipMemcpyAsync(m->nods, xyz.data(), sizeof(double) * 3 * n_nodes, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(m->conn, cn.data(), sizeof(int32_t) * 3 * n_el, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess && !dh.empty())
    e = hipMemcpyAsync(m->dof, dh.data(), sizeof(int32_t) * 18 * n_el, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    hfpx_mesh_destroy(m);
    return check_hip(ctx, e, "mesh upload");
  }
  *out = m;
  return HFPX_OK;
}

int hfpx_mesh_destroy(hfpx_mesh* m) {
  if (!m) return HFPX_OK;
  if (m->nods) (void)hipFree(m->nods);
  if (m->conn) (void)hipFree(m->conn);
  if (m->dof) (void)hipFree(m->dof);
  if (m->elem) (void)hipFree(m->elem);
  if (m->cp) (void)hipFree(m->cp);
  if (m->wts) (void)hipFree(m->wts);
  delete m;
  return HFPX_OK;
}

int64_t hfpx_mesh_n_dof(const hfpx_mesh* m) { return m ? m->n_dof : -1; }

int hfpx_mesh_set_dd_local(hfpx_mesh* m, int is_dd_local) {
  if (!m || (is_dd_local != 0 && is_dd_local != 1)) return HFPX_ERR_INVALID;
  m->dd_local = is_dd_local != 0;
  return HFPX_OK;
}

int hfpx_mesh_set_vertex_weights(hfpx_ctx* ctx, hfpx_mesh* m, const double* wts_host, int64_t n_nodes) {
  if (!ctx || !m) return HFPX_ERR_INVALID;
  HFPX_TRY_HIP(ctx, hipSetDevice(ctx->device));
  if (!wts_host) {
    if (m->wts) {
      HFPX_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
      HFPX_TRY_HIP(ctx, hipFree(m->wts));
      m->wts = nullptr;
    }
    return HFPX_OK;
  }
  if (n_nodes != m->n_nodes)
    return set_error(ctx, HFPX_ERR_INVALID, "%lld weights for a mesh of %lld nodes", (long long)n_nodes,
                     (long long)m->n_nodes);
  for (int64_t j = 0; j < n_nodes; ++j)
    if (!(wts_host[j] > 0.0) || !std::isfinite(wts_host[j]))
      return set_error(ctx, HFPX_ERR_INVALID, "vertex weight %lld = %g (need finite and > 0)", (long long)j,
                       wts_host[j]);
  if (!m->wts) HFPX_TRY_HIP(ctx, hipMalloc(&m->wts, sizeof(double) * n_nodes));
  // the copy is synchronous: the caller's buffer may go once this returns
  HFPX_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  HFPX_TRY_HIP(ctx, hipMemcpy(m->wts, wts_host, sizeof(double) * n_nodes, hipMemcpyHostToDevice));
  return HFPX_OK;
}

int hfpx_assemble(hfpx_ctx* ctx, hfpx_mesh* m, double shear_m, double poiss_r, double* A,
                  int64_t lda) {
  return hfpx_assemble_mixed(ctx, m, shear_m, poiss_r, A, lda, nullptr, 0);
}

int hfpx_assemble_mixed(hfpx_ctx* ctx, hfpx_mesh* m, double shear_m, double poiss_r, double* A64,
                        int64_t lda64, float* A32, int64_t lda32) {
  if (!ctx || !m) return HFPX_ERR_INVALID;
  if (!A64 && !A32) return set_error(ctx, HFPX_ERR_INVALID, "no output matrix");
  if ((A64 && lda64 < m->n_dof) || (A32 && lda32 < m->n_dof))
    return set_error(ctx, HFPX_ERR_INVALID, "lda smaller than n_dof = %lld", (long long)m->n_dof);
  if (!(poiss_r < 1.0) || !(shear_m > 0.0))
    return set_error(ctx, HFPX_ERR_INVALID, "need shear modulus > 0 and Poisson ratio < 1");
  HFPX_TRY_HIP(ctx, hipSetDevice(ctx->device));
  int rc = launch_geometry(ctx, m);
  if (rc) return rc;
  return launch_assemble(ctx, m, shear_m, poiss_r, A64, lda64, A32, lda32);
}

int hfpx_local_submatrix(hfpx_ctx* ctx, int kernel_id, double shear_m, double poiss_r, int64_t n_cases,
                         const double* h, const double* z, const double* tau, const double* sfm, double* out) {
  if (!ctx) return HFPX_ERR_INVALID;
  if (kernel_id != 1)
    return set_error(ctx, HFPX_ERR_INVALID, "local submatrix: kernel_id %d (only 1, the hypersingular kernel)",
                     kernel_id);
  if (n_cases < 1 || !h || !z || !tau || !sfm || !out)
    return set_error(ctx, HFPX_ERR_INVALID, "local submatrix: need n_cases >= 1 and device arrays");
  if (!(poiss_r < 1.0) || !(shear_m > 0.0))
    return set_error(ctx, HFPX_ERR_INVALID, "need shear modulus > 0 and Poisson ratio < 1");
  HFPX_TRY_HIP(ctx, hipSetDevice(ctx->device));
  return launch_local_submatrix(ctx, shear_m, poiss_r, n_cases, h, z, tau, sfm, out);
}

int hfpx_rhs_sinf(hfpx_ctx* ctx, hfpx_mesh* m, const double* sinf_host, int nrhs, double* B,
                  int64_t ldb) {
  if (!ctx || !m || !sinf_host || !B || nrhs < 1) return HFPX_ERR_INVALID;
  if (ldb < m->n_dof) return set_error(ctx, HFPX_ERR_INVALID, "ldb smaller than n_dof");
  HFPX_TRY_HIP(ctx, hipSetDevice(ctx->device));
  int rc = ensure_hwork(ctx, sizeof(double) * 6 * nrhs);
  if (rc) return rc;
  rc = ensure_work(ctx, sizeof(double) * 6 * nrhs);
  if (rc) return rc;
  memcpy(ctx->hwork, sinf_host, sizeof(double) * 6 * nrhs);
  HFPX_TRY_HIP(ctx, hipMemcpyAsync(ctx->work, ctx->hwork, sizeof(double) * 6 * nrhs,
                                   hipMemcpyHostToDevice, ctx->stream));
  rc = launch_geometry(ctx, m);
  if (rc) return rc;
  rc = launch_rhs(ctx, m, (const double*)ctx->work, nrhs, B, ldb);
  if (rc) return rc;
  // hwork/work are reused by the next call: finish before returning
  HFPX_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return HFPX_OK;
}

int hfpx_getrf(hfpx_ctx* ctx, int dtype, void* A, int64_t lda, int64_t n, int32_t* ipiv,
               int64_t* info_host) {
  if (!ctx) return HFPX_ERR_INVALID;
  if (info_host) *info_host = 0;
  if (!A || !ipiv || n < 1 || lda < n || (dtype != HFPX_F64 && dtype != HFPX_F32))
    return set_error(ctx, HFPX_ERR_INVALID, "getrf: need A, ipiv, n >= 1, lda >= n, dtype F64/F32");
  if (n > INT32_MAX - 1) return set_error(ctx, HFPX_ERR_INVALID, "getrf: n too large for int32 pivots");
  HFPX_TRY_HIP(ctx, hipSetDevice(ctx->device));
  int64_t info = 0;
  const int rc = lu_getrf(ctx, dtype, A, lda, n, ipiv, &info);
  if (info_host) *info_host = info;
  if (rc == HFPX_ERR_SINGULAR)
    set_error(ctx, rc, "U(%lld,%lld) is exactly zero", (long long)info, (long long)info);
  return rc;
}

int hfpx_getrs(hfpx_ctx* ctx, int dtype, const void* LU, int64_t lda, const int32_t* ipiv, int64_t n,
               void* B, int64_t ldb, int64_t nrhs) {
  if (!ctx) return HFPX_ERR_INVALID;
  if (!LU || !ipiv || !B || n < 1 || lda < n || ldb < n || nrhs < 1 ||
      (dtype != HFPX_F64 && dtype != HFPX_F32))
    return set_error(ctx, HFPX_ERR_INVALID, "getrs: invalid arguments");
  HFPX_TRY_HIP(ctx, hipSetDevice(ctx->device));
  return lu_getrs(ctx, dtype, LU, lda, ipiv, n, B, ldb, nrhs);
}

int hfpx_gemm_sub(hfpx_ctx* ctx, int dtype, void* C, int64_t ldc, const void* A, int64_t lda,
                  const void* B, int64_t ldb, int64_t M, int64_t N, int64_t K) {
  if (!ctx) return HFPX_ERR_INVALID;
  if (!C || !A || !B || M < 0 || N < 0 || K < 0 || ldc < M || lda < M || ldb < K ||
      (dtype != HFPX_F64 && dtype != HFPX_F32))
    return set_error(ctx, HFPX_ERR_INVALID, "gemm_sub: invalid arguments");
  HFPX_TRY_HIP(ctx, hipSetDevice(ctx->device));
  return gemm_sub_any(ctx, dtype, C, ldc, A, lda, B, ldb, M, N, K);
}

int hfpx_solve_ir(hfpx_ctx* ctx, const double* A64, int64_t lda64, const float* LU32, int64_t ldlu32,
                  const int32_t* ipiv, int64_t n, const double* B, int64_t ldb, double* X, int64_t ldx,
                  int64_t nrhs, double tol, int max_outer, int gmres_m, hfpx_ir_report* rep) {
  if (!ctx) return HFPX_ERR_INVALID;
  if (rep) *rep = hfpx_ir_report{0, 0, 0.0, 0, 0};
  if (!A64 || !LU32 || !ipiv || !B || !X || n < 1 || lda64 < n || ldlu32 < n || ldb < n || ldx < n ||
      nrhs < 1 || max_outer < 0 || !(tol > 0))
    return set_error(ctx, HFPX_ERR_INVALID, "solve_ir: invalid arguments");
  HFPX_TRY_HIP(ctx, hipSetDevice(ctx->device));
  return solve_ir(ctx, A64, lda64, LU32, ldlu32, ipiv, n, B, ldb, X, ldx, nrhs, tol, max_outer, gmres_m, rep);
}

int hfpx_profile(hfpx_ctx* ctx, int enable) {
  if (!ctx) return HFPX_ERR_INVALID;
  HFPX_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (enable) prof_clear(ctx);
  ctx->prof = enable != 0;
  return HFPX_OK;
}

int hfpx_profile_classes(hfpx_ctx* ctx, uint32_t mask) {
  if (!ctx) return HFPX_ERR_INVALID;
  ctx->prof_mask = mask;
  return HFPX_OK;
}

int hfpx_profile_read(hfpx_ctx* ctx, int kc, int64_t* launches, double* ms, double* flops, double* bytes) {
  if (!ctx || kc < 0 || kc >= HFPX_KC_COUNT) return HFPX_ERR_INVALID;
  HFPX_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  int64_t n = 0;
  double t = 0, f = 0, b = 0;
  for (auto& r : ctx->prof_recs) {
    if (r.cls != kc) continue;
    float e = 0.f;
    HFPX_TRY_HIP(ctx, hipEventElapsedTime(&e, r.a, r.b));
    ++n;
    t += e;
    f += r.flops;
    b += r.bytes;
  }
  if (launches) *launches = n;
  if (ms) *ms = t;
  if (flops) *flops = f;
  if (bytes) *bytes = b;
  return HFPX_OK;
}

int hfpx_tune_set(hfpx_ctx* ctx, int knob, int value) {
  if (!ctx) return HFPX_ERR_INVALID;
  if (knob == HFPX_KNOB_GEMM_F32) {
    if (value < 0 || value >= gemm_f32_variants())
      return set_error(ctx, HFPX_ERR_INVALID, "tune: GEMM_F32 variant %d out of range [0, %d)", value,
                       gemm_f32_variants());
    ctx->knob[knob] = value;
    return HFPX_OK;
  }
  if (knob == HFPX_KNOB_LU_NB) {
    if (value != 0 && value != 256 && value != 512 && value != 1024 && value != 2048 && value != 4096)
      return set_error(ctx, HFPX_ERR_INVALID, "tune: LU_NB must be 0, 256, 512, 1024, 2048 or 4096");
    ctx->knob[knob] = value;
    return HFPX_OK;
  }
  if (knob == HFPX_KNOB_X3_BAND) {
    if (value < 0 || value > 64) return set_error(ctx, HFPX_ERR_INVALID, "tune: X3_BAND must be in [0, 64]");
    ctx->knob[knob] = value;
    return HFPX_OK;
  }
  if (knob == HFPX_KNOB_PANEL) {
    if (value < 0 || value > 2) return set_error(ctx, HFPX_ERR_INVALID, "tune: PANEL must be 0, 1 or 2");
    ctx->knob[knob] = value;
    return HFPX_OK;
  }
  if (knob == HFPX_KNOB_LU_LOOKAHEAD) {
    if (value < 0 || value > 3) return set_error(ctx, HFPX_ERR_INVALID, "tune: LU_LOOKAHEAD must be 0 .. 3");
    ctx->knob[knob] = value;
    return HFPX_OK;
  }
  if (knob == HFPX_KNOB_GEMM_F64) {
    if (value < 0 || value > 5) return set_error(ctx, HFPX_ERR_INVALID, "tune: GEMM_F64 must be 0 .. 5");
    ctx->knob[knob] = value;
    return HFPX_OK;
  }
  if (knob == HFPX_KNOB_IR_FALLBACK) {
    if (value < 0 || value > 2) return set_error(ctx, HFPX_ERR_INVALID, "tune: IR_FALLBACK must be 0, 1 or 2");
    ctx->knob[knob] = value;
    return HFPX_OK;
  }
  if (knob == HFPX_KNOB_IR_MODE) {
    if (value < 0 || value > 2) return set_error(ctx, HFPX_ERR_INVALID, "tune: IR_MODE must be 0, 1 or 2");
    ctx->knob[knob] = value;
    return HFPX_OK;
  }
  if (knob == HFPX_KNOB_TEST_PANEL_TIMEOUT) {
    if (value < 0 || value > 1) return set_error(ctx, HFPX_ERR_INVALID, "tune: TEST_PANEL_TIMEOUT must be 0 or 1");
    ctx->knob[knob] = value;
    return HFPX_OK;
  }
  if (knob == HFPX_KNOB_BGEMM_F32) {
    if (value < 0 || value > 3) return set_error(ctx, HFPX_ERR_INVALID, "tune: BGEMM_F32 must be 0 .. 3");
    ctx->knob[knob] = value;
    return HFPX_OK;
  }
  return set_error(ctx, HFPX_ERR_INVALID, "tune: unknown knob %d", knob);
}

int hfpx_tune_get(const hfpx_ctx* ctx, int knob, int* value) {
  if (!ctx || !value || knob < 0 || knob >= HFPX_KNOB_COUNT) return HFPX_ERR_INVALID;
  *value = ctx->knob[knob];
  return HFPX_OK;
}

int hfpx_stress_noalloc(hfpx_ctx* ctx, hfpx_mesh* m, double shear_m, double poiss_r, const double* dd, int64_t ldd,
                        const double* pts, int64_t n_pts, const double* sinf_host, double* S, int64_t ldS) {
  if (!ctx || !m) return HFPX_ERR_INVALID;
  if (!dd || !pts || !sinf_host || !S || n_pts < 1 || ldd < 6 * m->n_el || ldS < n_pts)
    return set_error(ctx, HFPX_ERR_INVALID, "stress_noalloc: invalid arguments");
  if (!(poiss_r < 1.0) || !(shear_m > 0.0))
    return set_error(ctx, HFPX_ERR_INVALID, "need shear modulus > 0 and Poisson ratio < 1");
  HFPX_TRY_HIP(ctx, hipSetDevice(ctx->device));
  return launch_stress_noalloc(ctx, m, shear_m, poiss_r, dd, ldd, pts, n_pts, sinf_host, S, ldS);
}

int hfpx_stress(hfpx_ctx* ctx, hfpx_mesh* m, double shear_m, double poiss_r, const double* dd, int64_t ldd,
                const double* pts, int64_t n_pts, const double* sinf_host, double* S, int64_t ldS) {
  if (!ctx || !m) return HFPX_ERR_INVALID;
  if (!dd || !pts || !sinf_host || !S || n_pts < 1 || ldd < 6 * m->n_el || ldS < n_pts)
    return set_error(ctx, HFPX_ERR_INVALID, "stress: invalid arguments");
  if (!(poiss_r < 1.0) || !(shear_m > 0.0))
    return set_error(ctx, HFPX_ERR_INVALID, "need shear modulus > 0 and Poisson ratio < 1");
  HFPX_TRY_HIP(ctx, hipSetDevice(ctx->device));
  return launch_stress(ctx, m, shear_m, poiss_r, dd, ldd, pts, n_pts, sinf_host, S, ldS);
}

}  // extern "C"
