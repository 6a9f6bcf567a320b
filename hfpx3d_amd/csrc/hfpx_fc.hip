// hfpx_fc.hip -- cohesion / friction state of the crack surface, node by node,
// on gfx950.
//
// Replaces, on the device:
//   hfp3d::match_f_c   src/Development/cohesion_friction.cpp:22-127
//   hfp3d::update_mrd  src/Development/cohesion_friction.cpp:131-189
//
// One thread per node (collocation node of the DD array, n_nod = 6 n_el).
// Both are streaming, HBM-bound passes: 24 B of DD, 4 B of material id and
// 16 B of state read, 24 (match_f_c) or 16 (update_mrd) B written per node.
// The per-material tables (a few doubles per surface type) travel with an
// error counter in one host-to-device copy; a material id outside the table
// counts there and the call returns an error after the pass.
#include "hfpx_internal.h"

namespace hfpx {

namespace {

// work layout: [0] error counter (u64), then par (6 doubles per material),
// then the model ids (2 int32 per material: cohesion, friction)
struct FcTables {
  unsigned long long* bad;
  const double* par;
  const int32_t* cmid;
};

FcTables fc_tables(void* work, int64_t n_mat) {
  char* w = (char*)work;
  return {(unsigned long long*)w, (const double*)(w + 8), (const int32_t*)(w + 8 + 48 * n_mat)};
}

size_t fc_table_bytes(int64_t n_mat) { return 8 + 48 * (size_t)n_mat + 8 * (size_t)n_mat; }

}  // namespace

// Cohesion factor: Dugdale box (model 0) or linear softening (model 1) while
// the node is partly open (0 < opening / cr_open < 1), 0 otherwise (model 2
// and unknown ids); below the largest opening reached it unloads linearly.
// Friction: none once fully open; otherwise the residual coefficient,
// the peak one without weakening (model 0), or linear slip weakening from
// the largest of the current and reached slip (model 1).
__global__ __launch_bounds__(256) void k_match_f_c(int64_t n, const double* __restrict__ dd, int64_t ldd,
                                                   const int32_t* __restrict__ mat_id, FcTables t, int64_t n_mat,
                                                   const double* __restrict__ mr_open,
                                                   const double* __restrict__ mr_slip, double* __restrict__ fc,
                                                   double* __restrict__ sc, double* __restrict__ oc) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int m = mat_id[i];
  if (m < 0 || m >= n_mat) {
    atomicAdd(t.bad, 1ull);
    return;
  }
  const double* p = t.par + 6 * m;
  const int c_model = t.cmid[2 * m], f_model = t.cmid[2 * m + 1];
  const double d0 = dd[i], d1 = dd[i + ldd], d2 = dd[i + 2 * ldd];
  const double open = d2 / p[0];
  const double slip = sqrt(d0 * d0 + d1 * d1) / p[2];
  const double pr_open = mr_open[i];
  double f = 0.0;
  if (open > 0.0 && open < 1.0) f = c_model == 0 ? 1.0 : (c_model == 1 ? 1.0 - open : 0.0);
  if (open > 0.0 && open < pr_open) f *= open / pr_open;
  oc[i] = p[1] * f;
  sc[i] = p[3] * f;
  double mu = 0.0;
  if (open < 1.0) {
    mu = p[5];
    if (f_model == 0) {
      mu = p[4];
    } else if (f_model == 1 && slip < 1.0) {
      const double pr_slip = mr_slip[i];
      const double s = slip < pr_slip ? pr_slip : slip;
      mu = p[5] + (p[4] - p[5]) * (1.0 - s);
    }
  }
  fc[i] = mu;
}

// Largest relative opening and slip reached.  A full opening sets 1 unless
// the opening itself exceeds the previous maximum, which then wins.
__global__ __launch_bounds__(256) void k_update_mrd(int64_t n, const double* __restrict__ dd, int64_t ldd,
                                                    const int32_t* __restrict__ mat_id, FcTables t, int64_t n_mat,
                                                    double* __restrict__ mr_open, double* __restrict__ mr_slip) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int m = mat_id[i];
  if (m < 0 || m >= n_mat) {
    atomicAdd(t.bad, 1ull);
    return;
  }
  const double* p = t.par + 6 * m;
  const double d0 = dd[i], d1 = dd[i + ldd], d2 = dd[i + 2 * ldd];
  const double open = d2 / p[0];
  const double slip = sqrt(d0 * d0 + d1 * d1) / p[2];
  const double pr_open = mr_open[i], pr_slip = mr_slip[i];
  double mo = pr_open;
  if (open >= 1.0) mo = 1.0;
  if (open > pr_open) mo = open;
  mr_open[i] = mo;
  if (slip > pr_slip) mr_slip[i] = slip;
}

// Uploads the tables (and a zeroed error counter) through the pinned host
// workspace; returns the device tables.
static int fc_upload(hfpx_ctx* ctx, const int32_t* cmid_host, const hfpx_fc_param* par_host, int64_t n_mat,
                     FcTables* out) {
  const size_t bytes = fc_table_bytes(n_mat);
  int rc;
  if ((rc = ensure_hwork(ctx, bytes)) || (rc = ensure_work(ctx, bytes))) return rc;
  // the pinned buffer may still feed an earlier call's copy
  HFPX_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  char* h = (char*)ctx->hwork;
  memset(h, 0, bytes);
  double* par = (double*)(h + 8);
  int32_t* cm = (int32_t*)(h + 8 + 48 * n_mat);
  for (int64_t m = 0; m < n_mat; ++m) {
    const hfpx_fc_param& q = par_host[m];
    const double v[6] = {q.cr_open, q.peak_ts, q.cr_slip, q.peak_sc, q.peak_sf, q.res_sf};
    for (int k = 0; k < 6; ++k) par[6 * m + k] = v[k];
    if (cmid_host) {
      cm[2 * m] = cmid_host[m];
      cm[2 * m + 1] = cmid_host[m + n_mat];
    }
  }
  HFPX_TRY_HIP(ctx, hipMemcpyAsync(ctx->work, h, bytes, hipMemcpyHostToDevice, ctx->stream));
  *out = fc_tables(ctx->work, n_mat);
  return HFPX_OK;
}

// After the pass: the error counter decides the status.
static int fc_finish(hfpx_ctx* ctx, const char* what) {
  HFPX_TRY_HIP(ctx, hipGetLastError());
  unsigned long long bad = 0;
  HFPX_TRY_HIP(ctx, hipMemcpyAsync(ctx->hwork, ctx->work, 8, hipMemcpyDeviceToHost, ctx->stream));
  HFPX_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  memcpy(&bad, ctx->hwork, 8);
  if (bad) return set_error(ctx, HFPX_ERR_INVALID, "%s: %llu nodes with a material id outside the table", what, bad);
  return HFPX_OK;
}

static int fc_check(hfpx_ctx* ctx, int64_t n_nod, const double* dd, int64_t ldd, const int32_t* mat_id,
                    const hfpx_fc_param* par, int64_t n_mat) {
  if (!ctx) return HFPX_ERR_INVALID;
  if (n_nod < 0 || (n_nod > 0 && (!dd || !mat_id)) || !par || n_mat < 1 || ldd < n_nod)
    return set_error(ctx, HFPX_ERR_INVALID, "friction/cohesion: bad sizes or null arrays");
  for (int64_t m = 0; m < n_mat; ++m)
    if (!(par[m].cr_open != 0.0) || !(par[m].cr_slip != 0.0))
      return set_error(ctx, HFPX_ERR_INVALID, "friction/cohesion: material %lld has a zero critical opening or slip",
                       (long long)m);
  return HFPX_OK;
}

}  // namespace hfpx

using namespace hfpx;

extern "C" int hfpx_match_f_c(hfpx_ctx* ctx, int64_t n_nod, const double* dd, int64_t ldd, const int32_t* mat_id,
                              const int32_t* c_model_id_host, const hfpx_fc_param* f_c_param_host, int64_t n_mat,
                              const double* mr_open, const double* mr_slip, double* friction_coef,
                              double* slip_cohesion, double* open_cohesion) {
  int rc = fc_check(ctx, n_nod, dd, ldd, mat_id, f_c_param_host, n_mat);
  if (rc) return rc;
  if (!c_model_id_host || (n_nod > 0 && (!mr_open || !mr_slip || !friction_coef || !slip_cohesion || !open_cohesion)))
    return set_error(ctx, HFPX_ERR_INVALID, "match_f_c: null arrays");
  if (n_nod == 0) return HFPX_OK;
  HFPX_TRY_HIP(ctx, hipSetDevice(ctx->device));
  FcTables t;
  if ((rc = fc_upload(ctx, c_model_id_host, f_c_param_host, n_mat, &t))) return rc;
  hipLaunchKernelGGL(k_match_f_c, dim3((unsigned)((n_nod + 255) / 256)), dim3(256), 0, ctx->stream, n_nod, dd, ldd,
                     mat_id, t, n_mat, mr_open, mr_slip, friction_coef, slip_cohesion, open_cohesion);
  return fc_finish(ctx, "match_f_c");
}

extern "C" int hfpx_update_mrd(hfpx_ctx* ctx, int64_t n_nod, const double* dd, int64_t ldd, const int32_t* mat_id,
                               const hfpx_fc_param* f_c_param_host, int64_t n_mat, double* mr_open, double* mr_slip) {
  int rc = fc_check(ctx, n_nod, dd, ldd, mat_id, f_c_param_host, n_mat);
  if (rc) return rc;
  if (n_nod > 0 && (!mr_open || !mr_slip)) return set_error(ctx, HFPX_ERR_INVALID, "update_mrd: null arrays");
  if (n_nod == 0) return HFPX_OK;
  HFPX_TRY_HIP(ctx, hipSetDevice(ctx->device));
  FcTables t;
  if ((rc = fc_upload(ctx, nullptr, f_c_param_host, n_mat, &t))) return rc;
  hipLaunchKernelGGL(k_update_mrd, dim3((unsigned)((n_nod + 255) / 256)), dim3(256), 0, ctx->stream, n_nod, dd, ldd,
                     mat_id, t, n_mat, mr_open, mr_slip);
  return fc_finish(ctx, "update_mrd");
}
