// hfpx_internal.h -- shared host-side definitions of the HIP library.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/hfpx3d.h"
#include "hfpx_pair.h"

// Per-process context: one device, one stream, reusable device workspace.
struct hfpx_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  void* work = nullptr;      // grow-only device workspace
  size_t work_bytes = 0;
  void* hwork = nullptr;     // pinned host workspace
  size_t hwork_bytes = 0;
  void* lu_work = nullptr;   // grow-only device workspace of getrf / getrs
  size_t lu_work_bytes = 0;
  void* gemm_work = nullptr;  // grow-only device workspace of split-K products
  size_t gemm_work_bytes = 0;
  uint64_t* sync = nullptr;  // grid-barrier arrival counter (device)
  uint64_t sync_base = 0;    // its value once every launched barrier completes
  int num_cus = 0;
  int knob[HFPX_KNOB_COUNT] = {0};  // kernel-variant choices (hfpx_tune_set)
  std::string last_error;
  // opt-in per-launch timing (HIP events on the context's stream)
  bool prof = false;
  struct ProfRec {
    int cls;
    hipEvent_t a, b;
    double flops, bytes;
  };
  std::vector<ProfRec> prof_recs;
  std::vector<hipEvent_t> prof_pool;
};

// Device-resident mesh: raw nodes/connectivity/DoF map plus the element
// table that the geometry kernel derives from them.
struct hfpx_mesh {
  int64_t n_el = 0;
  int64_t n_dof = 0;
  double beta = 0.125;
  bool identity_dof = true;
  double* nods = nullptr;      // 3 x n_nodes (x, y, z rows only)
  int64_t n_nodes = 0;
  int32_t* conn = nullptr;     // 3 x n_el (vertex ids, 0-based)
  int32_t* dof = nullptr;      // n_el x 18 (row = element), -1 = fixed
  hfpx::SrcElem<double>* elem = nullptr;  // n_el
  double* cp = nullptr;        // 6 arrays of 6*n_el: cx cy cz nx ny nz
};

namespace hfpx {

// Profiling bracket around one launch: prof_begin before the launch,
// prof_end after it (no-ops unless profiling is enabled on the context).
int prof_begin(hfpx_ctx* ctx);
void prof_end(hfpx_ctx* ctx, int token, int cls, double flops, double bytes);

int set_error(hfpx_ctx* ctx, int code, const char* fmt, ...);
int check_hip(hfpx_ctx* ctx, hipError_t e, const char* what);
int ensure_work(hfpx_ctx* ctx, size_t bytes);
int ensure_hwork(hfpx_ctx* ctx, size_t bytes);

// launchers (hfpx_assemble.hip)
int launch_geometry(hfpx_ctx* ctx, hfpx_mesh* m);
int launch_assemble(hfpx_ctx* ctx, const hfpx_mesh* m, double shear_m, double poiss_r,
                    double* a64, int64_t lda64, float* a32, int64_t lda32);
int launch_rhs(hfpx_ctx* ctx, const hfpx_mesh* m, const double* sinf_dev, int nrhs,
               double* b, int64_t ldb);

// dense linear algebra (hfpx_lu.hip)
int lu_getrf(hfpx_ctx* ctx, int dtype, void* A, int64_t lda, int64_t n, int32_t* ipiv, int64_t* info);
int lu_getrs(hfpx_ctx* ctx, int dtype, const void* LU, int64_t lda, const int32_t* ipiv, int64_t n,
             void* B, int64_t ldb, int64_t nrhs);
int gemm_f32_variants();
// C (M x N) -= A (M x K) B (K x N), column-major, on the context stream
// (hfpx_gemm.hip; instantiated for float and double).
template <typename T>
int gemm_sub(hfpx_ctx* ctx, T* C, int64_t ldc, const T* A, int64_t lda, const T* B, int64_t ldb,
             int64_t M, int64_t N, int64_t K);
int gemm_sub_any(hfpx_ctx* ctx, int dtype, void* C, int64_t ldc, const void* A, int64_t lda,
                 const void* B, int64_t ldb, int64_t M, int64_t N, int64_t K);

// FP32 triangular sweeps with block inverses (hfpx_sweep.hip)
size_t sweep_f32_work_bytes(int64_t n);
int sweep_f32(hfpx_ctx* ctx, const float* LU, int64_t lda, int64_t n, float* B, int64_t ldb, int64_t nrhs,
              void* work);

// stress postprocessing (hfpx_stress.hip)
int launch_stress(hfpx_ctx* ctx, hfpx_mesh* m, double shear_m, double nu, const double* dd, int64_t ldd,
                  const double* pts, int64_t n_pts, const double* sinf_host, double* out, int64_t ldo);

// mixed-precision solve (hfpx_solve.hip)
int solve_ir(hfpx_ctx* ctx, const double* A, int64_t lda, const float* LU, int64_t ldlu,
             const int32_t* ipiv, int64_t n, const double* B, int64_t ldb, double* X, int64_t ldx,
             int64_t k, double tol, int max_outer, int m, hfpx_ir_report* rep);

}  // namespace hfpx

#define HFPX_TRY_HIP(ctx, expr)                                   \
  do {                                                            \
    hipError_t e__ = (expr);                                      \
    if (e__ != hipSuccess) return hfpx::check_hip(ctx, e__, #expr); \
  } while (0)
