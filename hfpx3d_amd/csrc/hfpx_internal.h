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
  void* x3_work = nullptr;    // grow-only workspace of the bf16-split GEMM operands
  size_t x3_work_bytes = 0;
  // LU look-ahead: a second default-priority stream for the trailing
  // update, its own split-operand workspace, and the two events that order
  // it against the main stream (created at the first factorisation that
  // uses them)
  hipStream_t stream2 = nullptr;
  hipEvent_t la_ev[2] = {nullptr, nullptr};
  void* x3_work2 = nullptr;
  size_t x3_work2_bytes = 0;
  uint64_t* sync = nullptr;  // grid-barrier arrival counter (device)
  uint64_t sync_base = 0;    // its value once every launched barrier completes
  int num_cus = 0;
  bool no_coop = false;      // cooperative launches refused: launch panels normally
  int knob[HFPX_KNOB_COUNT] = {0};  // kernel-variant choices (hfpx_tune_set)
  std::string last_error;
  // opt-in per-launch timing (HIP events on the context's stream)
  bool prof = false;
  uint32_t prof_mask = ~0u;  // kernel classes recorded (bit HFPX_KC_*)
  struct ProfRec {
    int cls;
    hipEvent_t a, b;
    double flops, bytes;
  };
  std::vector<ProfRec> prof_recs;
  std::vector<hipEvent_t> prof_pool;
  // device address of the allocation-free stress workspace on this
  // context's device (hfpx_stress.hip), set at creation
  double* stress_scratch = nullptr;
};

// Device-resident mesh: raw nodes/connectivity/DoF map plus the element
// table that the geometry kernel derives from them.
struct hfpx_mesh {
  int64_t n_el = 0;
  int64_t n_dof = 0;
  double beta = 0.125;
  bool identity_dof = true;
  bool dd_local = false;       // DD in each element's own frame (Num_Param_T::is_dd_local)
  double* nods = nullptr;      // 3 x n_nodes (x, y, z rows only)
  int64_t n_nodes = 0;
  int32_t* conn = nullptr;     // 3 x n_el (vertex ids, 0-based)
  int32_t* dof = nullptr;      // n_el x 18 (row = element), -1 = fixed
  hfpx::SrcElem<double>* elem = nullptr;  // n_el
  double* cp = nullptr;        // 6 arrays of 6*n_el: cx cy cz nx ny nz
  double* wts = nullptr;       // n_nodes vertex weights of a non-uniform edge partition, or null
};

namespace hfpx {

// ---- batches of independent cracks --------------------------------------
// Blocking of the batched LU: outer blocks of BNB columns, panels of BIB.
constexpr int BNB = 128;
constexpr int BIB = 32;

// One crack of a batch (device table).  All of a crack's DoF are free
// (fix_tip = 0): DoF 18 e + 3 node + component of the crack's own elements.
struct BCrack {
  int64_t el_off;   // first element in the batch's element table
  int64_t n_el;
  int64_t n;        // DoF = 18 n_el
  int64_t a_off;    // first entry of the crack's FP64 n x n matrix (ld n)
  int64_t a32_off;  // first entry of its FP32 copy, or -1 (solved in FP64)
  int64_t f_off;    // first entry of its FP64 working copy (the FP64 LU), or -1
  int64_t x_off;    // first entry of its n-vectors (rhs, solution, pivots)
  int64_t wg_off;   // first workgroup of the crack in the batched assembly grid
  double G, nu;
  double sinf[6];   // far-field stress (xx yy zz xy xz yz), tension positive
};

// Column ranges of one step of the batched LU for a crack of n DoF: outer
// block [k, k + kb), panel [j, j + jb) inside it; false if the crack has no
// work in that step.
__host__ __device__ inline bool bstep(int64_t n, int64_t k, int64_t j, int64_t& kb, int64_t& jb) {
  if (k >= n) return false;
  kb = (n - k) < BNB ? (n - k) : BNB;
  if (j >= k + kb) return false;
  jb = (k + kb - j) < BIB ? (k + kb - j) : BIB;
  return true;
}

// Schur-complement product C -= A B of step (k, j) of the batched LU for a
// crack of n DoF (local column-major indices, ld n): outer = false is the
// update inside the outer block after panel [j, j + jb), outer = true the
// trailing update after the whole block.  C starts at (r0, c0), A at
// (r0, p0), B at (p0, c0); false if the product is empty.
__host__ __device__ inline bool bgemm_shape(int64_t n, int64_t k, int64_t j, bool outer, int64_t& r0, int64_t& c0,
                                            int64_t& p0, int64_t& M, int64_t& N, int64_t& K) {
  int64_t kb, jb;
  M = N = K = 0;
  r0 = c0 = p0 = 0;
  if (!bstep(n, k, outer ? k : j, kb, jb)) return false;
  if (outer) {
    r0 = k + kb; c0 = k + kb; p0 = k; M = n - r0; N = n - c0; K = kb;
  } else {
    r0 = j + jb; c0 = j + jb; p0 = j; M = n - r0; N = k + kb - c0; K = jb;
  }
  return M > 0 && N > 0 && K > 0;
}

// Walk of a GEMM's tile grid (shared by the FP32, bf16-split and batched
// kernels).
constexpr int SG_GROUP = 8;
// Tile (tile_r, tile_c) of a tm x tn tile grid for position nid of the
// grid's walk in bands of group row tiles (column by column inside a
// band), so that consecutive positions share A row blocks and B columns.
__device__ inline void sg_band_tile(int64_t nid, int64_t tm, int64_t tn, int64_t& tile_r, int64_t& tile_c,
                                    int64_t group = SG_GROUP) {
  const int64_t per = group * tn, first = (nid / per) * group;
  const int64_t gs = (tm - first) < group ? (tm - first) : group;
  const int64_t r = nid % per;
  tile_r = first + r % gs;
  tile_c = r / gs;
}

// Position of workgroup id in a walk of nt tiles where each of the 8 XCDs
// (workgroups are dealt to XCDs round-robin) owns one contiguous eighth.
__device__ inline int64_t sg_xcd_order(int64_t id, int64_t nt) {
  const int64_t xcd = id % 8, local = id / 8;
  const int64_t fl = nt / 8, rem = nt % 8;
  return xcd * fl + (xcd < rem ? xcd : rem) + local;
}

// Tile (rows and columns) of the batched Schur-complement updates, and the
// number of tiles of one crack's product.
constexpr int BGT = 128;
__host__ __device__ inline int64_t bgemm_tiles(int64_t n, int64_t k, int64_t j, bool outer) {
  int64_t r0, c0, p0, M, N, K;
  if (!bgemm_shape(n, k, j, outer, r0, c0, p0, M, N, K)) return 0;
  return ((M + BGT - 1) / BGT) * ((N + BGT - 1) / BGT);
}

// Which copy of a crack's matrix a batched kernel works on.
enum BWhich { B_A64 = 0, B_A32 = 1, B_F64 = 2 };
__host__ __device__ inline int64_t bcopy_off(const BCrack& c, int which) {
  return which == B_A32 ? c.a32_off : (which == B_F64 ? c.f_off : c.a_off);
}

// Profiling bracket around one launch: prof_begin before the launch,
// prof_end after it (no-ops unless profiling is enabled on the context).
int prof_begin(hfpx_ctx* ctx, int cls);
int launch_vc_border(hfpx_ctx* ctx, const hfpx_mesh* m, double* a64, int64_t lda64, float* a32, int64_t lda32);
void prof_end(hfpx_ctx* ctx, int token, int cls, double flops, double bytes);

// max that keeps NaN: a NaN entry makes the whole reduction NaN, so a
// non-finite residual or iterate cannot look converged (fmax drops NaN)
__host__ __device__ __forceinline__ double nan_max(double a, double b) {
  return (a != a) ? a : ((b != b || b > a) ? b : a);
}

// A triangle with no usable frame: twice its area at most 1e-12 of its
// longest edge squared (zero-area, needle or repeated vertices), or any
// non-finite value.  Such an element would divide by zero in the geometry.
inline bool degenerate_triangle(const double* p0, const double* p1, const double* p2) {
  double a[3], b[3], c[3];
  for (int k = 0; k < 3; ++k) {
    a[k] = p1[k] - p0[k];
    b[k] = p2[k] - p0[k];
    c[k] = p2[k] - p1[k];
  }
  const double cx = a[1] * b[2] - a[2] * b[1], cy = a[2] * b[0] - a[0] * b[2], cz = a[0] * b[1] - a[1] * b[0];
  const double area2 = __builtin_sqrt(cx * cx + cy * cy + cz * cz);
  const double l2 = std::max({a[0] * a[0] + a[1] * a[1] + a[2] * a[2], b[0] * b[0] + b[1] * b[1] + b[2] * b[2],
                              c[0] * c[0] + c[1] * c[1] + c[2] * c[2]});
  return !(area2 > 1e-12 * l2);
}

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
// the element integral on (h, z, tau, sfm) cases (device arrays)
int launch_local_submatrix(hfpx_ctx* ctx, double shear_m, double poiss_r, int64_t n_cases, const double* h,
                           const double* z, const double* tau, const double* sfm, double* out);

// dense linear algebra (hfpx_lu.hip)
int lu_getrf(hfpx_ctx* ctx, int dtype, void* A, int64_t lda, int64_t n, int32_t* ipiv, int64_t* info);
int lu_getrs(hfpx_ctx* ctx, int dtype, const void* LU, int64_t lda, const int32_t* ipiv, int64_t n,
             void* B, int64_t ldb, int64_t nrhs);
int gemm_f32_variants();
// FP32 C -= A B by three-way bf16 operand splitting on bf16 MFMA
// (hfpx_gemm_x3.hip); workspace bytes for an M x N x K product
int gemm_x3(hfpx_ctx* ctx, float* C, int64_t ldc, const float* A, int64_t lda, const float* B, int64_t ldb,
            int64_t M, int64_t N, int64_t K, int wide);
size_t gemm_x3_work_bytes(int64_t M, int64_t N, int64_t K);
int ensure_x3_work(hfpx_ctx* ctx, size_t bytes);
// C (M x N) -= A (M x K) B (K x N), column-major, on the context stream
// (hfpx_gemm.hip; instantiated for float and double).
template <typename T>
int gemm_sub(hfpx_ctx* ctx, T* C, int64_t ldc, const T* A, int64_t lda, const T* B, int64_t ldb,
             int64_t M, int64_t N, int64_t K);
// C -= A B (FP64, N <= 64, K >= 8192) streaming A once, and with rowabs
// also rowabs[i] = sum_k |A(i, k)|
int dgemm_stream_sub(hfpx_ctx* ctx, double* C, int64_t ldc, const double* A, int64_t lda, const double* B,
                     int64_t ldb, int64_t M, int64_t N, int64_t K, double* rowabs);
// the FP64 products that gemm_sub routes to dgemm_stream_sub
bool dgemm_stream_applies(int64_t N, int64_t K);
int gemm_sub_any(hfpx_ctx* ctx, int dtype, void* C, int64_t ldc, const void* A, int64_t lda,
                 const void* B, int64_t ldb, int64_t M, int64_t N, int64_t K);

// FP32 triangular sweeps with block inverses (hfpx_sweep.hip)
size_t sweep_f32_work_bytes(int64_t n, int64_t nrhs);
// batched FP32 sweeps, one right-hand side per crack: tk holds ntk tickets
// (crack table index, level), level-major and by size within a level; bofs
// (per table index) the offset of a crack's blocks in the block tables
// (inverses of 128 x 128 blocks, flags); flags holds 2 ntk + 3 words
int launch_bblock_inv(hfpx_ctx* ctx, const float* A32, const BCrack* cr, const int32_t* tk, int64_t ntk,
                      const int64_t* bofs, float* invL, float* invU);
int launch_bsweeps(hfpx_ctx* ctx, const float* A32, const BCrack* cr, const int32_t* tk, int64_t ntk,
                   const int64_t* bofs, const float* invL, const float* invU, float* vec, uint64_t* flags);
int sweep_f32(hfpx_ctx* ctx, const float* LU, int64_t lda, int64_t n, float* B, int64_t ldb, int64_t nrhs,
              void* work);

// batched assembly (hfpx_assemble.hip): every crack of the table; cp planes
// have stride cps; total_wg workgroups in one 1-D grid
int launch_bassemble(hfpx_ctx* ctx, const SrcElem<double>* elem, const double* cp, int64_t cps,
                     const BCrack* cr, int ncr, int64_t total_wg, double* a64, float* a32);
int launch_brhs(hfpx_ctx* ctx, const SrcElem<double>* elem, int64_t n_el_total, const BCrack* cr, int ncr,
                double* b);
// batched Schur-complement update (hfpx_gemm.hip) of step (k, j) for the
// cracks list[0 .. nact) (bgemm_shape); tstart (device, nact + 1 entries)
// is the prefix sum of bgemm_tiles along the list, tiles = tstart[nact]
template <typename T>
int launch_bgemm(hfpx_ctx* ctx, T* base, int which, const BCrack* cr, const int32_t* list, int nact,
                 const int32_t* tstart, int64_t tiles, int64_t k, int64_t j, bool outer);

// batched FP32 outer update on bf16 MFMA by three-way splitting
// (hfpx_gemm_x3.hip): step k of the FP32 group; rstart / cstart (nact + 1
// entries) are the prefix sums of the cracks' 128-row tiles of A and
// 128-column tiles of B, rtot / ctot their totals; split holds
// bgemm_x3_split_bytes(rtot, ctot)
int launch_bgemm_x3(hfpx_ctx* ctx, float* base, int which, const BCrack* cr, const int32_t* list, int nact,
                    const int32_t* tstart, int64_t tiles, const int32_t* rstart, int64_t rtot, const int32_t* cstart,
                    int64_t ctot, int64_t k, __bf16* split);
size_t bgemm_x3_split_bytes(int64_t rtot, int64_t ctot);

// stress postprocessing (hfpx_stress.hip)
int launch_stress(hfpx_ctx* ctx, hfpx_mesh* m, double shear_m, double nu, const double* dd, int64_t ldd,
                  const double* pts, int64_t n_pts, const double* sinf_host, double* out, int64_t ldo);
// the same on the library's file-scope device workspace: no allocation, no
// synchronisation; stress_scratch_address resolves that workspace's address
// on ctx's device (once per device), stress_scratch_doubles is its size
int launch_stress_noalloc(hfpx_ctx* ctx, hfpx_mesh* m, double shear_m, double nu, const double* dd, int64_t ldd,
                          const double* pts, int64_t n_pts, const double* sinf_host, double* out, int64_t ldo);
int stress_scratch_address(hfpx_ctx* ctx, double** out);
size_t stress_scratch_doubles();

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
