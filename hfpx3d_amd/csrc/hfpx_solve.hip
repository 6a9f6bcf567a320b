// hfpx_solve.hip -- mixed-precision solve: FP32 LU factors + FP64 GMRES
// iterative refinement (GMRES-IR), batched over many right-hand sides.
//
// No reference counterpart for the method; it replaces the FP64
// il::LU::solve path (il/linear_algebra/dense/factorization/LU.h:106-138) and
// is pinned by the FP64 oracle solution.
//
// Outer loop (FP64):  r = b - A x  with the FP64 matrix, normwise backward
//   error w = |r|_inf / (|A|_inf |x|_inf + |b|_inf) per column; stop when
//   w <= tol for every column.
// Correction: classic refinement  x += M^-1 r  (one preconditioner
//   application per outer step) while every such step lowers the backward
//   error at least 100-fold, i.e. while the FP32 factors contract the error
//   strongly (72,000-DoF ellipse: 2e-6, 1e-9, 1e-12, 3e-15); from the first
//   step that does not, GMRES(m) on  M^-1 A d = M^-1 r  with M = (P^T L U)
//   from the FP32 factorisation (vectors FP64, the preconditioner applied in
//   FP32), modified Gram-Schmidt, Givens rotations on the host; x += d.
//   Both cost one FP64 product with A and one preconditioner application
//   per step; the classic step saves the restart's extra application and
//   the Gram-Schmidt work.  HFPX_KNOB_IR_MODE selects GMRES or classic
//   steps only.
// All right-hand sides advance together: every matrix product is one
// n x nrhs block product (the FP64 matrix and the FP32 factors are streamed
// once per sweep for all columns).
// Failure: an outer step that does not lower the backward error by 10 %, a
// non-finite iterate (an FP32 zero pivot, FP32 overflow) or max_outer steps
// without convergence end the refinement; the solution is then recomputed
// by an FP64 LU of a copy of A (report fell_back = 1) unless
// HFPX_KNOB_IR_FALLBACK is 2.  Non-finite A or B is reported as is.
#include <algorithm>
#include <cmath>
#include <vector>

#include "hfpx_internal.h"

namespace hfpx {

constexpr int RB = 256;  // rows per reduction block

__global__ void k_cast_d2f(const double* __restrict__ x, int64_t ldx, float* __restrict__ y,
                           int64_t ldy, int64_t n, int64_t ncol) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t j = blockIdx.y;
  if (i < n && j < ncol) y[i + j * ldy] = (float)x[i + j * ldx];
}
__global__ void k_cast_f2d(const float* __restrict__ x, int64_t ldx, double* __restrict__ y,
                           int64_t ldy, int64_t n, int64_t ncol) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t j = blockIdx.y;
  if (i < n && j < ncol) y[i + j * ldy] = (double)x[i + j * ldx];
}
__global__ void k_copy(const double* __restrict__ x, int64_t ldx, double* __restrict__ y, int64_t ldy,
                       int64_t n, int64_t ncol) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t j = blockIdx.y;
  if (i < n && j < ncol) y[i + j * ldy] = x[i + j * ldx];
}

// part[b + nb*j] = sum (op) over rows of block b of column j; op 0: x*y, 1: max|x|
__global__ __launch_bounds__(RB) void k_colreduce(const double* __restrict__ x, int64_t ldx,
                                                  const double* __restrict__ y, int64_t ldy, int64_t n,
                                                  int op, double* __restrict__ part) {
  __shared__ double s[RB];
  const int64_t j = blockIdx.y;
  const int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x;
  double v = 0.0;
  if (i < n) v = (op == 0) ? x[i + j * ldx] * y[i + j * ldy] : fabs(x[i + j * ldx]);
  s[threadIdx.x] = v;
  __syncthreads();
  for (int w = RB / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) s[threadIdx.x] = (op == 0) ? s[threadIdx.x] + s[threadIdx.x + w]
                                                    : nan_max(s[threadIdx.x], s[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x + (int64_t)gridDim.x * j] = s[0];
}
// out[j] = reduce over nb partials (fixed order: deterministic)
__global__ void k_colreduce2(const double* __restrict__ part, int64_t nb, int64_t ncol, int op,
                             double* __restrict__ out) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ncol) return;
  double v = 0.0;
  for (int64_t b = 0; b < nb; ++b) v = (op == 0) ? v + part[b + nb * j] : nan_max(v, part[b + nb * j]);
  out[j] = v;
}
// y[:, j] += a[j] * x[:, j]   (a on device)
__global__ void k_axpy(const double* __restrict__ a, const double* __restrict__ x, int64_t ldx,
                       double* __restrict__ y, int64_t ldy, int64_t n, double sign) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t j = blockIdx.y;
  if (i < n) y[i + j * ldy] += sign * a[j] * x[i + j * ldx];
}
// x[:, j] *= a[j]
// x += d (n x ncol)
__global__ void k_add(const double* __restrict__ d, int64_t ldd, double* __restrict__ x, int64_t ldx, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t j = blockIdx.y;
  if (i < n) x[i + j * ldx] += d[i + j * ldd];
}
__global__ void k_scale(const double* __restrict__ a, double* __restrict__ x, int64_t ldx, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t j = blockIdx.y;
  if (i < n) x[i + j * ldx] *= a[j];
}
// row sums of |A|: column slice blockIdx.y of every row -> part[y n + i]
// (consecutive threads read consecutive rows of a column), then the slices
// added in order -> rs[i]
__global__ void k_rowabs(const double* __restrict__ A, int64_t lda, int64_t n, int64_t cw,
                         double* __restrict__ part) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t c0 = (int64_t)blockIdx.y * cw, c1 = (c0 + cw < n) ? c0 + cw : n;
  double s = 0.0;
  for (int64_t c = c0; c < c1; ++c) s += fabs(A[i + c * lda]);
  part[blockIdx.y * n + i] = s;
}

__global__ void k_rowabs_sum(const double* __restrict__ part, int64_t n, int ns, double* __restrict__ rs) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int z = 0; z < ns; ++z) s += part[z * n + i];
  rs[i] = s;
}

namespace {

struct Work {
  double* R;      // n x k residual / scratch (FP64)
  double* V;      // (m+1) blocks of n x k (Krylov basis)
  double* Wv;     // n x k
  float* W32;     // n x k FP32 staging
  double* part;   // reduction partials
  double* col;    // k column scalars (device)
  double* col2;   // k
  int64_t m1;     // blocks in V (m + 1)
};

inline dim3 grid2(int64_t n, int64_t k, int bs) { return dim3((unsigned)((n + bs - 1) / bs), (unsigned)k); }

}  // namespace

// column reductions to host
static int colred(hfpx_ctx* ctx, const Work& w, const double* x, int64_t ldx, const double* y,
                  int64_t ldy, int64_t n, int64_t k, int op, double* host) {
  const int64_t nb = (n + RB - 1) / RB;
  hipLaunchKernelGGL(k_colreduce, dim3((unsigned)nb, (unsigned)k), dim3(RB), 0, ctx->stream, x, ldx,
                     y ? y : x, y ? ldy : ldx, n, op, w.part);
  hipLaunchKernelGGL(k_colreduce2, dim3((unsigned)((k + 63) / 64)), dim3(64), 0, ctx->stream, w.part, nb,
                     k, op, w.col);
  HFPX_TRY_HIP(ctx, hipGetLastError());
  HFPX_TRY_HIP(ctx, hipMemcpyAsync(host, w.col, sizeof(double) * k, hipMemcpyDeviceToHost, ctx->stream));
  HFPX_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return HFPX_OK;
}

// host column scalars -> device; synchronous, so the host buffer can be reused
static int put_col(hfpx_ctx* ctx, double* dev, const double* host, int64_t k) {
  HFPX_TRY_HIP(ctx, hipMemcpyAsync(dev, host, sizeof(double) * k, hipMemcpyHostToDevice, ctx->stream));
  HFPX_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return HFPX_OK;
}

// y := M^-1 x  (FP32 factors), x, y FP64 n x k
static int precond(hfpx_ctx* ctx, const Work& w, const float* LU, int64_t ldlu, const int32_t* ipiv,
                   int64_t n, const double* x, int64_t ldx, double* y, int64_t ldy, int64_t k) {
  hipLaunchKernelGGL(k_cast_d2f, grid2(n, k, 256), dim3(256), 0, ctx->stream, x, ldx, w.W32, n, n, k);
  HFPX_TRY_HIP(ctx, hipGetLastError());
  int rc = lu_getrs(ctx, HFPX_F32, LU, ldlu, ipiv, n, w.W32, n, k);
  if (rc) return rc;
  hipLaunchKernelGGL(k_cast_f2d, grid2(n, k, 256), dim3(256), 0, ctx->stream, w.W32, n, y, ldy, n, k);
  HFPX_TRY_HIP(ctx, hipGetLastError());
  return HFPX_OK;
}

// R := B - A X; with rowabs (the first residual of a solve), also the row
// sums of |A| from the same pass over A
static int residual(hfpx_ctx* ctx, const double* A, int64_t lda, int64_t n, const double* B,
                    int64_t ldb, const double* X, int64_t ldx, double* R, int64_t k, double* rowabs = nullptr) {
  hipLaunchKernelGGL(k_copy, grid2(n, k, 256), dim3(256), 0, ctx->stream, B, ldb, R, n, n, k);
  HFPX_TRY_HIP(ctx, hipGetLastError());
  if (rowabs) {
    const int tok = prof_begin(ctx, HFPX_KC_GEMM_F64);
    const int rc = dgemm_stream_sub(ctx, R, n, A, lda, X, ldx, n, k, n, rowabs);
    prof_end(ctx, tok, HFPX_KC_GEMM_F64, 2.0 * n * k * n, 8.0 * ((double)n * n + 2.0 * n * k));
    return rc;
  }
  return gemm_sub_any(ctx, HFPX_F64, R, n, A, lda, X, ldx, n, k, n);
}

// Normwise backward error of every column of X: omega_j =
// |b - A x|_inf / (|A|_inf |x|_inf + |b|_inf); *worst = max over columns,
// NaN if any is NaN.  Leaves the residual in w.R.
static int backward_error(hfpx_ctx* ctx, const Work& w, const double* A, int64_t lda, int64_t n, const double* B,
                          int64_t ldb, const double* X, int64_t ldx, int64_t k, double& anorm,
                          const std::vector<double>& bnorm, std::vector<double>& rnorm, std::vector<double>& xnorm,
                          std::vector<double>& omega, double* worst) {
  int rc;
  // anorm < 0: not known yet.  Where the residual streams A (many rows,
  // few right-hand sides) that pass also gives the row sums of |A|;
  // otherwise they take a pass of their own.  V is free until the first
  // correction.
  const bool first = anorm < 0;
  const bool fused = first && dgemm_stream_applies(k, n);
  if (first && !fused) {
    const int ns = (int)std::max<int64_t>(1, std::min<int64_t>(32, (int64_t)w.m1 * k - 1));
    const int64_t cw = (n + ns - 1) / ns;
    hipLaunchKernelGGL(k_rowabs, dim3((unsigned)((n + 255) / 256), (unsigned)ns), dim3(256), 0, ctx->stream, A,
                       lda, n, cw, w.V + n);
    hipLaunchKernelGGL(k_rowabs_sum, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, w.V + n, n, ns,
                       w.V);
    HFPX_TRY_HIP(ctx, hipGetLastError());
  }
  if ((rc = residual(ctx, A, lda, n, B, ldb, X, ldx, w.R, k, fused ? w.V : nullptr))) return rc;
  if (first) {
    double h;
    if ((rc = colred(ctx, w, w.V, n, nullptr, 0, n, 1, 1, &h))) return rc;
    anorm = h;
  }
  if ((rc = colred(ctx, w, w.R, n, nullptr, 0, n, k, 1, rnorm.data()))) return rc;
  if ((rc = colred(ctx, w, X, ldx, nullptr, 0, n, k, 1, xnorm.data()))) return rc;
  double wv = 0.0;
  for (int64_t j = 0; j < k; ++j) {
    const double den = anorm * xnorm[j] + bnorm[j];
    // 0 / 0 (a zero column of B solved exactly) is 0; a NaN anywhere stays NaN
    omega[j] = (den > 0 || den != den) ? rnorm[j] / den : (rnorm[j] == 0.0 ? 0.0 : rnorm[j] / den);
    wv = nan_max(wv, omega[j]);
  }
  *worst = wv;
  return HFPX_OK;
}

// X := A^-1 B by an FP64 LU of a copy of A (A itself stays untouched: it
// is the caller's FP64 operator).  The copy is allocated here and freed
// before returning.
static int fallback_fp64(hfpx_ctx* ctx, const double* A, int64_t lda, int64_t n, const double* B, int64_t ldb,
                         double* X, int64_t ldx, int64_t k) {
  double* F = nullptr;
  int32_t* piv = nullptr;
  hipError_t e = hipMalloc((void**)&F, sizeof(double) * (size_t)n * (size_t)n);
  if (e == hipSuccess) e = hipMalloc((void**)&piv, sizeof(int32_t) * (size_t)n);
  if (e != hipSuccess) {
    if (F) (void)hipFree(F);
    return check_hip(ctx, e, "GMRES-IR fallback: FP64 copy of the matrix");
  }
  int rc = HFPX_OK;
  e = hipMemcpy2DAsync(F, sizeof(double) * n, A, sizeof(double) * lda, sizeof(double) * n, n,
                       hipMemcpyDeviceToDevice, ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpy2DAsync(X, sizeof(double) * ldx, B, sizeof(double) * ldb, sizeof(double) * n, k,
                         hipMemcpyDeviceToDevice, ctx->stream);
  if (e != hipSuccess) rc = check_hip(ctx, e, "GMRES-IR fallback: copies");
  int64_t info = 0;
  if (!rc) {
    rc = lu_getrf(ctx, HFPX_F64, F, n, n, piv, &info);
    if (rc == HFPX_ERR_SINGULAR)
      set_error(ctx, rc, "GMRES-IR fallback: FP64 U(%lld,%lld) is exactly zero", (long long)info, (long long)info);
  }
  if (!rc) rc = lu_getrs(ctx, HFPX_F64, F, n, piv, n, X, ldx, k);
  const hipError_t es = hipStreamSynchronize(ctx->stream);
  (void)hipFree(F);
  (void)hipFree(piv);
  if (!rc && es != hipSuccess) rc = check_hip(ctx, es, "GMRES-IR fallback");
  return rc;
}

int solve_ir(hfpx_ctx* ctx, const double* A, int64_t lda, const float* LU, int64_t ldlu,
             const int32_t* ipiv, int64_t n, const double* B, int64_t ldb, double* X, int64_t ldx,
             int64_t k, double tol, int max_outer, int m, hfpx_ir_report* rep) {
  if (m < 1) m = 1;
  if (m > 64) m = 64;
  const size_t nk = (size_t)n * k;
  const int64_t nb = (n + RB - 1) / RB;
  const size_t bytes = sizeof(double) * nk * (size_t)(m + 3) + sizeof(float) * nk +
                       sizeof(double) * ((size_t)nb * k + 2 * k) + 4096;
  int rc = ensure_work(ctx, bytes);
  if (rc) return rc;
  Work w;
  char* p = (char*)ctx->work;
  auto take = [&](size_t b) {
    char* q = p;
    p += (b + 255) & ~(size_t)255;
    return q;
  };
  w.R = (double*)take(sizeof(double) * nk);
  w.Wv = (double*)take(sizeof(double) * nk);
  w.V = (double*)take(sizeof(double) * nk * (m + 1));
  w.m1 = m + 1;
  w.W32 = (float*)take(sizeof(float) * nk);
  w.part = (double*)take(sizeof(double) * nb * k);
  w.col = (double*)take(sizeof(double) * k);
  w.col2 = (double*)take(sizeof(double) * k);

  // |A|_inf comes with the first residual (backward_error, anorm < 0)
  double anorm = -1.0;
  std::vector<double> bnorm(k), xnorm(k), rnorm(k), omega(k), beta(k), tmp(k);
  if ((rc = colred(ctx, w, B, ldb, nullptr, 0, n, k, 1, bnorm.data()))) return rc;

  // x0 = M^-1 b
  if ((rc = precond(ctx, w, LU, ldlu, ipiv, n, B, ldb, X, ldx, k))) return rc;

  std::vector<double> H((size_t)(m + 1) * m * k), cs((size_t)m * k), sn((size_t)m * k), g((size_t)(m + 1) * k);
  int outer = 0, inner_total = 0;
  double worst = 0.0, prev_worst = HUGE_VAL;
  bool converged = false, stalled = false;
  // non-finite data: nothing to refine, and an FP64 LU of the same data
  // cannot do better
  bool finite_data = true;  // set once |A|_inf is known (first backward error)
  // 0: classic steps while they contract 100-fold, then GMRES; 1: GMRES
  // only; 2: classic only
  const int ir_mode = ctx->knob[HFPX_KNOB_IR_MODE];
  bool classic = ir_mode != 1, last_classic = false;
  for (outer = 0; outer <= max_outer; ++outer) {
    if ((rc = backward_error(ctx, w, A, lda, n, B, ldb, X, ldx, k, anorm, bnorm, rnorm, xnorm, omega, &worst)))
      return rc;
    if (outer == 0) {
      double data_norm = anorm;
      for (int64_t j = 0; j < k; ++j) data_norm = nan_max(data_norm, bnorm[j]);
      finite_data = std::isfinite(data_norm);
    }
    if (finite_data && worst <= tol) {
      converged = true;
      break;
    }
    // a classic step that contracted less than 100-fold (or not at all):
    // GMRES from here on, from the current iterate
    bool to_gmres = false;
    if (last_classic && ir_mode == 0 && std::isfinite(worst) && !(worst <= 0.01 * prev_worst)) {
      classic = false;
      to_gmres = true;
    }
    // NaN / inf in the iterate, or no progress: the FP32 factors do not
    // precondition this matrix well enough (or are singular)
    if (!finite_data || !std::isfinite(worst) || (outer >= 1 && !to_gmres && !(worst < 0.9 * prev_worst))) {
      stalled = true;
      break;
    }
    prev_worst = worst;
    if (outer == max_outer) break;
    last_classic = classic;
    if (classic) {
      // x += M^-1 r (backward_error left r in w.R)
      if ((rc = precond(ctx, w, LU, ldlu, ipiv, n, w.R, n, w.Wv, n, k))) return rc;
      hipLaunchKernelGGL(k_add, grid2(n, k, 256), dim3(256), 0, ctx->stream, w.Wv, n, X, ldx, n);
      HFPX_TRY_HIP(ctx, hipGetLastError());
      continue;
    }
    // the correction has to take the backward error from `worst` below tol:
    // the inner iteration stops at a preconditioned-residual reduction of
    // 0.01 tol / worst (clamped to [1e-12, 1e-6]); the outer loop still runs
    // until the backward error itself is below tol
    const double inner_tol = std::max(1e-12, std::min(1e-6, 0.01 * tol / worst));
    // GMRES on M^-1 A d = M^-1 r, zero initial guess
    double* V0 = w.V;
    if ((rc = precond(ctx, w, LU, ldlu, ipiv, n, w.R, n, V0, n, k))) return rc;
    if ((rc = colred(ctx, w, V0, n, V0, n, n, k, 0, beta.data()))) return rc;
    for (int64_t j = 0; j < k; ++j) {
      beta[j] = std::sqrt(beta[j]);
      tmp[j] = beta[j] > 0 ? 1.0 / beta[j] : 0.0;
      g[(size_t)j * (m + 1)] = beta[j];
      for (int q = 1; q <= m; ++q) g[(size_t)j * (m + 1) + q] = 0.0;
    }
    if ((rc = put_col(ctx, w.col2, tmp.data(), k))) return rc;
    hipLaunchKernelGGL(k_scale, grid2(n, k, 256), dim3(256), 0, ctx->stream, w.col2, V0, n, n);
    int steps = 0;
    for (int s = 0; s < m; ++s) {
      double* Vs = w.V + nk * s;
      double* Vn = w.V + nk * (s + 1);
      // Wv = A Vs  (Wv := 0 - A Vs, sign folded below)
      HFPX_TRY_HIP(ctx, hipMemsetAsync(w.Wv, 0, sizeof(double) * nk, ctx->stream));
      if ((rc = gemm_sub_any(ctx, HFPX_F64, w.Wv, n, A, lda, Vs, n, n, k, n))) return rc;
      // Vn = M^-1 (A Vs) = -M^-1 Wv
      if ((rc = precond(ctx, w, LU, ldlu, ipiv, n, w.Wv, n, Vn, n, k))) return rc;
      for (int64_t j = 0; j < k; ++j) tmp[j] = -1.0;
      if ((rc = put_col(ctx, w.col2, tmp.data(), k))) return rc;
      hipLaunchKernelGGL(k_scale, grid2(n, k, 256), dim3(256), 0, ctx->stream, w.col2, Vn, n, n);
      // modified Gram-Schmidt
      for (int q = 0; q <= s; ++q) {
        double* Vq = w.V + nk * q;
        if ((rc = colred(ctx, w, Vn, n, Vq, n, n, k, 0, tmp.data()))) return rc;
        for (int64_t j = 0; j < k; ++j) H[((size_t)j * m + s) * (m + 1) + q] = tmp[j];
        if ((rc = put_col(ctx, w.col2, tmp.data(), k))) return rc;
        hipLaunchKernelGGL(k_axpy, grid2(n, k, 256), dim3(256), 0, ctx->stream, w.col2, Vq, n, Vn, n, n, -1.0);
      }
      if ((rc = colred(ctx, w, Vn, n, Vn, n, n, k, 0, tmp.data()))) return rc;
      double worst_res = 0.0;
      for (int64_t j = 0; j < k; ++j) {
        const double hn = std::sqrt(tmp[j]);
        double* Hj = &H[((size_t)j * m + s) * (m + 1)];
        Hj[s + 1] = hn;
        tmp[j] = hn > 0 ? 1.0 / hn : 0.0;
        // apply previous rotations, then a new one
        for (int q = 0; q < s; ++q) {
          const double c = cs[(size_t)j * m + q], sv = sn[(size_t)j * m + q];
          const double t0 = c * Hj[q] + sv * Hj[q + 1];
          Hj[q + 1] = -sv * Hj[q] + c * Hj[q + 1];
          Hj[q] = t0;
        }
        const double a = Hj[s], b = Hj[s + 1];
        const double r = std::hypot(a, b);
        const double c = r > 0 ? a / r : 1.0, sv = r > 0 ? b / r : 0.0;
        cs[(size_t)j * m + s] = c;
        sn[(size_t)j * m + s] = sv;
        Hj[s] = r;
        Hj[s + 1] = 0.0;
        double* gj = &g[(size_t)j * (m + 1)];
        gj[s + 1] = -sv * gj[s];
        gj[s] = c * gj[s];
        const double rel = beta[j] > 0 ? std::fabs(gj[s + 1]) / beta[j] : 0.0;
        worst_res = std::max(worst_res, rel);
      }
      if ((rc = put_col(ctx, w.col2, tmp.data(), k))) return rc;
      hipLaunchKernelGGL(k_scale, grid2(n, k, 256), dim3(256), 0, ctx->stream, w.col2, Vn, n, n);
      steps = s + 1;
      if (worst_res < inner_tol) break;
    }
    inner_total += steps;
    // y = H^-1 g (upper triangular, per column); x += V y
    std::vector<double> y((size_t)steps * k);
    for (int64_t j = 0; j < k; ++j) {
      for (int q = steps - 1; q >= 0; --q) {
        double v = g[(size_t)j * (m + 1) + q];
        for (int qq = q + 1; qq < steps; ++qq) v -= H[((size_t)j * m + qq) * (m + 1) + q] * y[(size_t)qq * k + j];
        const double d = H[((size_t)j * m + q) * (m + 1) + q];
        y[(size_t)q * k + j] = d != 0 ? v / d : 0.0;
      }
    }
    for (int q = 0; q < steps; ++q) {
      if ((rc = put_col(ctx, w.col2, &y[(size_t)q * k], k))) return rc;
      hipLaunchKernelGGL(k_axpy, grid2(n, k, 256), dim3(256), 0, ctx->stream, w.col2, w.V + nk * q, n, X,
                         ldx, n, 1.0);
      HFPX_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));  // host y is reused
    }
  }
  // FP64 LU of a copy of A when refinement failed (stagnation, non-finite
  // iterate or no convergence in max_outer steps) and the data are finite
  bool fell_back = false;
  if (!converged && finite_data && ctx->knob[HFPX_KNOB_IR_FALLBACK] != 2) {
    if ((rc = fallback_fp64(ctx, A, lda, n, B, ldb, X, ldx, k))) return rc;
    fell_back = true;
    if ((rc = backward_error(ctx, w, A, lda, n, B, ldb, X, ldx, k, anorm, bnorm, rnorm, xnorm, omega, &worst)))
      return rc;
    converged = worst <= tol;
  }
  if (rep) {
    rep->iterations = outer;
    rep->inner_iterations = inner_total;
    rep->backward_error = worst;
    rep->converged = converged ? 1 : 0;
    rep->fell_back = fell_back ? 1 : 0;
  }
  HFPX_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (converged) return HFPX_OK;
  return set_error(ctx, HFPX_ERR_NOT_CONVERGED, "iterative refinement %s at backward error %.3e%s",
                   stalled ? "stagnated" : "stopped", worst, fell_back ? " after the FP64 LU fallback" : "");
}

}  // namespace hfpx
