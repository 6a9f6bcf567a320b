// hfpx_cf.hip -- one step of the volume-control cohesion / friction
// iteration on gfx950, on a resident volume-control matrix.
//
// Replaces, on the device:
//   hfp3d::vc_cf_iteration  src/Development/c_f_iteration.cpp:27-436
//     (declared c_f_iteration.h:24-38), which chains match_f_c
//     (cohesion_friction.cpp:22-127), the tractions of the current DD
//     (:65-77), the traction admissibility of every collocation node
//     (:96-276), mod_3dbem_system_vc (system_assembly.cpp:560-629), a dense
//     solve (:295) and the DD / pressure update (:310-428).
//
// Step on the device (one stream, three host synchronisations: after
// match_f_c, after the renumbering, and at the end; the refinement
// synchronises on its own):
//   1. k_cf_ddv: the DD of the DoF active in the current handle, at their
//      original numbers; then e = A[:, :n0] d with the FP64 product of the
//      LU (the volume-control matrix stays resident; row n0 is the volume);
//   2. k_cf_nodes, one thread per collocation node: far-field and DD
//      tractions in the element frame, nodal pressure at the collocation
//      point, the reference's case analysis (full separation / cohesive zone
//      / slip / intact), the right-hand side -R^T dt by original DoF and the
//      new activity of the node's three DoF;
//   3. k_cf_count / k_cf_scan / k_cf_number: the active original DoF
//      renumbered in order (exclusive prefix sum), the gather list of the
//      truncated system, the new handle, and whether the activity differs
//      from the one the cached FP32 factors were built with;
//   4. when it differs (or on the first step): gather of the truncated
//      system (FP64 + FP32 in one pass) and FP32 LU; the FP32 factors are
//      cached for the next steps of the same active set;
//   5. FP64 GMRES-IR on the truncated system;
//   6. k_cf_update, one thread per element: DD increments, no overlap,
//      pressure at the collocation points that opened;
//   7. the residual |incr|_1 + |delta_trac|_1 by fixed-order partial sums.
//
// Bound: HBM.  The product in 1 streams the n0^2 FP64 matrix once (8 n0^2
// bytes); the gather, when the active set changed, reads and writes the
// truncated system (20 bytes per entry); the FP32 LU and the refinement
// are the LU's and the solve's own kernels.
#include <algorithm>
#include <climits>
#include <cmath>
#include <vector>

#include "hfpx_internal.h"

struct hfpx_cf_sys {
  hfpx_mesh* mesh = nullptr;
  const double* A = nullptr;  // (n0 + 1)^2 volume-control matrix (caller's, resident)
  int64_t lda = 0;
  int64_t n0 = 0, n_el = 0, nblk = 0;
  std::vector<int32_t> orig;  // host copy of the original handle, 18 e + l
  int32_t* slot_of = nullptr;  // n0: slot 18 e + l of each original DoF
  int32_t* cur = nullptr;      // 18 n_el: current handle (slot order 18 e + l)
  int32_t* nxt = nullptr;      // 18 n_el: new handle
  double* ddv = nullptr;       // n0
  double* eneg = nullptr;      // n0 + 1: -A d
  double* dtr = nullptr;       // n0: right-hand side by original DoF
  int32_t* act = nullptr;      // n0: new activity
  int32_t* act_f = nullptr;    // n0: activity of the cached factors
  int32_t* idx = nullptr;      // n0 + 1: gather list
  int32_t* bsum = nullptr;     // nblk
  int32_t* iscal = nullptr;    // [0] active count, [1] changed
  double* scal = nullptr;      // [0] pressure before, [1] pressure after, [2] residual
  double* rhs = nullptr;       // n0 + 1
  double* x = nullptr;         // n0 + 1
  double* part = nullptr;      // residual partial sums
  double* cr_open = nullptr;   // per material
  int64_t cr_cap = 0;
  double* M64 = nullptr;       // (n0 + 1)^2
  float* M32 = nullptr;        // (n0 + 1)^2, FP32 factors once factorised
  int32_t* ipiv = nullptr;     // n0 + 1
  bool cached = false;
  int64_t n_cached = -1;
};

namespace hfpx {

namespace {

constexpr int CF_BLK = 1024;  // original DoF per scan block
constexpr int CF_RED = 1024;  // entries per residual partial sum

struct CfLoad {
  double s[6];  // far-field stress xx yy zz xy xz yz
};

// Values of the element's six shape functions at its collocation point q,
// from the point's coordinates in the element frame and the barycentric
// form of the basis (uniform element, or the vertex weights').
__device__ __forceinline__ void sf_at_cp(const SrcElem<double>& E, const double* __restrict__ cp, int64_t ncp,
                                         int64_t e, int q, double N[6]) {
  const int64_t c = 6 * e + q;
  const double px = cp[c] - E.v0[0], py = cp[ncp + c] - E.v0[1], pz = cp[2 * ncp + c] - E.v0[2];
  const double lx = E.R[0] * px + E.R[1] * py + E.R[2] * pz;
  const double ly = E.R[3] * px + E.R[4] * py + E.R[5] * pz;
  double L[3];
  for (int k = 0; k < 3; ++k) L[k] = E.l0[k] + 2.0 * (E.l1[2 * k] * lx - E.l1[2 * k + 1] * ly);
  for (int v = 0; v < 3; ++v) {
    const int a = (v + 1) % 3, b = (v + 2) % 3;
    N[v] = L[v] * L[v] - E.sw[2 * v] * L[v] * L[a] - E.sw[2 * v + 1] * L[v] * L[b];
    N[3 + v] = E.sw[6 + v] * L[a] * L[b];
  }
}

__device__ __forceinline__ void rot(const double* R, const double* x, double* y) {
  for (int i = 0; i < 3; ++i) y[i] = R[3 * i] * x[0] + R[3 * i + 1] * x[1] + R[3 * i + 2] * x[2];
}
__device__ __forceinline__ void rot_t(const double* R, const double* x, double* y) {
  for (int i = 0; i < 3; ++i) y[i] = R[i] * x[0] + R[3 + i] * x[1] + R[6 + i] * x[2];
}

__device__ __forceinline__ int64_t orig_of(const int32_t* __restrict__ od, int64_t s) {
  return od ? (int64_t)od[s] : s;
}

// DD of the currently active DoF at their original numbers; slots fixed in
// the original handle are -1 in the new handle; the pressure at the
// injection node is kept for the update.
__global__ void k_cf_ddv(int64_t n_el, const int32_t* __restrict__ od, const int32_t* __restrict__ cur,
                         const double* __restrict__ dd, int64_t ldd, const double* __restrict__ pp, int64_t inj,
                         double* __restrict__ ddv, int32_t* __restrict__ nxt, double* __restrict__ scal) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s == 0) scal[0] = pp[inj];
  if (s >= 18 * n_el) return;
  const int64_t o = orig_of(od, s);
  if (o < 0) {
    nxt[s] = -1;
    return;
  }
  const int64_t e = s / 18, l = s % 18;
  ddv[o] = cur[s] >= 0 ? dd[6 * e + l / 3 + ldd * (l % 3)] : 0.0;
}

// Admissibility of the traction at every collocation node: one thread per
// node g = 6 e + q.  eneg = -A d (row o the traction component of original
// DoF o).  Writes dtr (the right-hand side, -R^T dt) and act (1 / 0) for
// the node's original DoF.
__global__ __launch_bounds__(256) void k_cf_nodes(
    const SrcElem<double>* __restrict__ elem, const double* __restrict__ cp, int64_t n_el,
    const int32_t* __restrict__ od, const int32_t* __restrict__ cur, const double* __restrict__ eneg,
    const double* __restrict__ pp, const double* __restrict__ fc, const double* __restrict__ sc,
    const double* __restrict__ oc, const double* __restrict__ mr_open, const double* __restrict__ mr_slip, CfLoad F,
    double* __restrict__ dtr, int32_t* __restrict__ act) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= 6 * n_el) return;
  const int64_t e = g / 6;
  const int q = (int)(g % 6);
  const SrcElem<double>& E = elem[e];
  double R[9];
  for (int k = 0; k < 9; ++k) R[k] = E.R[k];
  // far-field traction on the element (normal -e3), element frame
  const double nv[3] = {-R[6], -R[7], -R[8]};
  const double* s = F.s;
  const double tg[3] = {s[0] * nv[0] + s[3] * nv[1] + s[4] * nv[2], s[3] * nv[0] + s[1] * nv[1] + s[5] * nv[2],
                        s[4] * nv[0] + s[5] * nv[1] + s[2] * nv[2]};
  double ti[3], tr[3], tt[3];
  rot(R, tg, ti);
  int64_t o[3];
  for (int i = 0; i < 3; ++i) {
    o[i] = orig_of(od, 18 * e + 3 * q + i);
    tt[i] = o[i] >= 0 ? -eneg[o[i]] : 0.0;
  }
  rot(R, tt, tr);
  double N[6];
  sf_at_cp(E, cp, 6 * n_el, e, q, N);
  double pr = 0.0;
  for (int j = 0; j < 6; ++j) pr += pp[6 * e + j] * N[j];
  const double nt = pr - ti[2] - tr[2];
  const double st0 = -ti[0] - tr[0], st1 = -ti[1] - tr[1];
  const double sh = sqrt(st0 * st0 + st1 * st1);
  const double adm_nt = oc[g];
  const double adm_st = sc[g] - fc[g] * nt;
  double d0 = 0.0, d1 = 0.0;
  if (sh > 0) {
    d0 = st0 / sh;
    d1 = st1 / sh;
  }
  double dt[3] = {0.0, 0.0, 0.0};
  int a[3] = {2, 2, 2};  // 1 active, 0 released, 2 as in the current handle
  if (nt > adm_nt) {
    // full separation: all traction released
    dt[0] = -st0;
    dt[1] = -st1;
    dt[2] = -nt;
    a[0] = a[1] = a[2] = 1;
  } else {
    if (mr_open[g] < 1.0 && nt > 0.0) {
      // cohesive zone: normal traction to the cohesion
      dt[2] = adm_nt - nt;
      a[2] = 1;
    }
    if (mr_slip[g] > 0 || sh > adm_st) {
      // slip: shear traction to the friction bound
      dt[0] = adm_st * d0 - st0;
      dt[1] = adm_st * d1 - st1;
      a[0] = a[1] = 1;
    } else if (nt <= 0.0) {
      // intact: the node leaves the system
      a[0] = a[1] = a[2] = 0;
    }
  }
  double gdt[3];
  rot_t(R, dt, gdt);
  for (int i = 0; i < 3; ++i) {
    if (o[i] < 0) continue;
    dtr[o[i]] = -gdt[i];
    act[o[i]] = a[i] == 2 ? (cur[18 * e + 3 * q + i] >= 0 ? 1 : 0) : a[i];
  }
}

// active original DoF per block of CF_BLK
__global__ __launch_bounds__(CF_BLK) void k_cf_count(const int32_t* __restrict__ act, int64_t n0,
                                                     int32_t* __restrict__ bsum) {
  const int64_t o = (int64_t)blockIdx.x * CF_BLK + threadIdx.x;
  const int c = __syncthreads_count(o < n0 && act[o] != 0);
  if (threadIdx.x == 0) bsum[blockIdx.x] = c;
}

// exclusive prefix of the block counts (one thread, nblk entries), total in
// iscal[0]; clears the changed flag iscal[1]
__global__ void k_cf_scan(int32_t* __restrict__ bsum, int64_t nblk, int32_t* __restrict__ iscal) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int32_t run = 0;
  for (int64_t b = 0; b < nblk; ++b) {
    const int32_t c = bsum[b];
    bsum[b] = run;
    run += c;
  }
  iscal[0] = run;
  iscal[1] = 0;
}

// New number of every active original DoF (block offset + rank in the
// block, ranks from wave ballots), the gather list and the new handle;
// iscal[1] |= 1 where the activity differs from act_f (compare != 0).
__global__ __launch_bounds__(CF_BLK) void k_cf_number(const int32_t* __restrict__ act,
                                                      const int32_t* __restrict__ act_f, int compare, int64_t n0,
                                                      const int32_t* __restrict__ bsum,
                                                      const int32_t* __restrict__ slot_of,
                                                      int32_t* __restrict__ idx, int32_t* __restrict__ nxt,
                                                      int32_t* __restrict__ iscal) {
  __shared__ int32_t wsum[CF_BLK / 64];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int64_t o = (int64_t)blockIdx.x * CF_BLK + t;
  const bool in = o < n0;
  const bool a = in && act[o] != 0;
  const uint64_t m = __ballot(a);
  const int rank = __popcll(m & ((1ull << lane) - 1ull));
  if (lane == 0) wsum[w] = __popcll(m);
  const bool diff = compare && in && ((act_f[o] != 0) != a);
  const uint64_t dm = __ballot(diff);
  if (lane == 0 && dm) atomicOr(&iscal[1], 1);
  __syncthreads();
  int off = bsum[blockIdx.x];
  for (int k = 0; k < w; ++k) off += wsum[k];
  if (!in) return;
  const int32_t num = a ? off + rank : -1;
  if (a) idx[num] = (int32_t)o;
  nxt[slot_of[o]] = num;
}

// right-hand side of the truncated system; closes the gather list with the
// border index n0
__global__ void k_cf_rhs(int32_t* __restrict__ idx, int64_t nu, const double* __restrict__ dtr,
                         const double* __restrict__ eneg, int64_t n0, double new_vol, double* __restrict__ rhs) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t > nu) return;
  if (t < nu) {
    rhs[t] = dtr[idx[t]];
  } else {
    idx[nu] = (int32_t)n0;
    rhs[nu] = new_vol + eneg[n0];  // new volume - current volume
  }
}

// DD and pressure update, one thread per element.  x: increments by new
// DoF, x[nu] the pressure increment.
__global__ void k_cf_update(const SrcElem<double>* __restrict__ elem, const double* __restrict__ cp, int64_t n_el,
                            bool dd_local, const int32_t* __restrict__ od, const int32_t* __restrict__ nxt,
                            const double* __restrict__ x, int64_t nu, const int32_t* __restrict__ mat_id,
                            const double* __restrict__ cr_open, const double* __restrict__ mr_open,
                            double* __restrict__ dd, int64_t ldd, double* __restrict__ pp,
                            double* __restrict__ scal) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_el) return;
  const double pressure = scal[0] + x[nu];
  if (e == 0) scal[1] = pressure;
  const SrcElem<double>& E = elem[e];
  double R[9];
  for (int k = 0; k < 9; ++k) R[k] = E.R[k];
  double op[6];  // local opening of the six nodes
  for (int q = 0; q < 6; ++q) {
    const int64_t g = 6 * e + q;
    double dn[3], dl[3];
    int64_t o[3];
    for (int i = 0; i < 3; ++i) {
      const int64_t s = 18 * e + 3 * q + i;
      o[i] = orig_of(od, s);
      const int32_t d = nxt[s];
      dn[i] = dd[g + ldd * i] + ((o[i] >= 0 && d >= 0) ? x[d] : 0.0);
    }
    if (dd_local) {
      for (int i = 0; i < 3; ++i) dl[i] = dn[i];
    } else {
      rot(R, dn, dl);
    }
    if (dl[2] < 0.0) dl[2] = 0.0;  // no overlap
    op[q] = dl[2];
    if (!dd_local) rot_t(R, dl, dn);
    else
      for (int i = 0; i < 3; ++i) dn[i] = dl[i];
    for (int i = 0; i < 3; ++i)
      if (o[i] >= 0) dd[g + ldd * i] = dn[i];
  }
  for (int q = 0; q < 6; ++q) {
    const int64_t g = 6 * e + q;
    double N[6];
    sf_at_cp(E, cp, 6 * n_el, e, q, N);
    double w = 0.0;
    for (int j = 0; j < 6; ++j) w += op[j] * N[j];
    const double fr = mr_open[g];
    double st = w / cr_open[mat_id[g]];
    if (st >= 1.0 || fr >= 1.0) st = 1.0;
    if (st >= 1.0 || (st > 0.0 && fr >= 1.0)) pp[g] = pressure;
  }
}

// partial sums of |x[0..n1)| followed by |y[0..n2)|, CF_RED entries per
// block, in a fixed tree order
__global__ __launch_bounds__(256) void k_cf_l1(const double* __restrict__ x, int64_t n1,
                                               const double* __restrict__ y, int64_t n2,
                                               double* __restrict__ part) {
  __shared__ double s[256];
  const int t = threadIdx.x;
  double v = 0.0;
  for (int k = 0; k < CF_RED / 256; ++k) {
    const int64_t i = (int64_t)blockIdx.x * CF_RED + k * 256 + t;
    if (i < n1) v += fabs(x[i]);
    else if (i < n1 + n2) v += fabs(y[i - n1]);
  }
  s[t] = v;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (t < h) s[t] += s[t + h];
    __syncthreads();
  }
  if (t == 0) part[blockIdx.x] = s[0];
}

__global__ void k_cf_l1_sum(const double* __restrict__ part, int64_t nb, double* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double r = 0.0;
  for (int64_t b = 0; b < nb; ++b) r += part[b];
  *out = r;
}

__global__ void k_cf_fill(double* __restrict__ a, int64_t n, double v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

__global__ void k_cf_copy_i32(const int32_t* __restrict__ a, int32_t* __restrict__ b, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

inline unsigned blocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

template <typename T>
int cf_alloc(hfpx_ctx* ctx, T** p, size_t count) {
  HFPX_TRY_HIP(ctx, hipMalloc((void**)p, sizeof(T) * std::max<size_t>(count, 1)));
  return HFPX_OK;
}

void cf_free(hfpx_cf_sys* s) {
  void* ptrs[] = {s->slot_of, s->cur, s->nxt, s->ddv, s->eneg, s->dtr, s->act, s->act_f, s->idx, s->bsum,
                  s->iscal, s->scal, s->rhs, s->x, s->part, s->cr_open, s->M64, s->M32, s->ipiv};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
}

}  // namespace

// gather of the truncated volume-control system from a device list
// (hfpx_vc.hip)
int launch_vc_gather_idx(hfpx_ctx* ctx, const double* A, int64_t lda, const int32_t* idx, int64_t nu, double* M64,
                         int64_t ldm64, float* M32, int64_t ldm32);

}  // namespace hfpx

using namespace hfpx;

extern "C" {

int hfpx_cf_create(hfpx_ctx* ctx, hfpx_mesh* mesh, const double* A_vc, int64_t lda, hfpx_cf_sys** out) {
  if (!ctx) return HFPX_ERR_INVALID;
  if (!mesh || !A_vc || !out) return set_error(ctx, HFPX_ERR_INVALID, "cf_create: null argument");
  *out = nullptr;
  const int64_t n0 = mesh->n_dof, n_el = mesh->n_el;
  if (n0 < 1 || lda < n0 + 1) return set_error(ctx, HFPX_ERR_INVALID, "cf_create: lda below n_dof + 1");
  if (n0 + 1 >= INT32_MAX || 18 * n_el >= INT32_MAX)
    return set_error(ctx, HFPX_ERR_INVALID, "cf_create: more than 2^31 DoF");
  HFPX_TRY_HIP(ctx, hipSetDevice(ctx->device));
  hfpx_cf_sys* s = new hfpx_cf_sys();
  s->mesh = mesh;
  s->A = A_vc;
  s->lda = lda;
  s->n0 = n0;
  s->n_el = n_el;
  s->nblk = (n0 + CF_BLK - 1) / CF_BLK;
  // the original handle on the host: one slot per DoF (every handle
  // make_dof_h_crack builds), so each original DoF has one slot
  s->orig.resize((size_t)(18 * n_el));
  if (mesh->identity_dof) {
    for (int64_t k = 0; k < 18 * n_el; ++k) s->orig[(size_t)k] = (int32_t)k;
  } else {
    const hipError_t e = hipMemcpy(s->orig.data(), mesh->dof, sizeof(int32_t) * s->orig.size(),
                                   hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
      delete s;
      return check_hip(ctx, e, "cf_create: DoF map");
    }
  }
  std::vector<int32_t> slot_of((size_t)n0, -1);
  for (int64_t k = 0; k < 18 * n_el; ++k) {
    const int32_t o = s->orig[(size_t)k];
    if (o < 0) continue;
    if (o >= n0 || slot_of[(size_t)o] >= 0) {
      delete s;
      return set_error(ctx, HFPX_ERR_INVALID, "cf_create: the mesh's DoF map has DoF %d on more than one slot "
                       "or out of range", o);
    }
    slot_of[(size_t)o] = (int32_t)k;
  }
  for (int64_t o = 0; o < n0; ++o)
    if (slot_of[(size_t)o] < 0) {
      delete s;
      return set_error(ctx, HFPX_ERR_INVALID, "cf_create: DoF %lld has no slot", (long long)o);
    }
  const size_t nu = (size_t)n0 + 1;
  int rc = HFPX_OK;
  if (!rc) rc = cf_alloc(ctx, &s->slot_of, (size_t)n0);
  if (!rc) rc = cf_alloc(ctx, &s->cur, (size_t)(18 * n_el));
  if (!rc) rc = cf_alloc(ctx, &s->nxt, (size_t)(18 * n_el));
  if (!rc) rc = cf_alloc(ctx, &s->ddv, (size_t)n0);
  if (!rc) rc = cf_alloc(ctx, &s->eneg, nu);
  if (!rc) rc = cf_alloc(ctx, &s->dtr, (size_t)n0);
  if (!rc) rc = cf_alloc(ctx, &s->act, (size_t)n0);
  if (!rc) rc = cf_alloc(ctx, &s->act_f, (size_t)n0);
  if (!rc) rc = cf_alloc(ctx, &s->idx, nu);
  if (!rc) rc = cf_alloc(ctx, &s->bsum, (size_t)s->nblk);
  if (!rc) rc = cf_alloc(ctx, &s->iscal, 2);
  if (!rc) rc = cf_alloc(ctx, &s->scal, 3);
  if (!rc) rc = cf_alloc(ctx, &s->rhs, nu);
  if (!rc) rc = cf_alloc(ctx, &s->x, nu);
  if (!rc) rc = cf_alloc(ctx, &s->part, (nu + (size_t)n0 + CF_RED - 1) / CF_RED + 1);
  if (!rc) rc = cf_alloc(ctx, &s->M64, nu * nu);
  if (!rc) rc = cf_alloc(ctx, &s->M32, nu * nu);
  if (!rc) rc = cf_alloc(ctx, &s->ipiv, nu);
  if (!rc) {
    const hipError_t e = hipMemcpy(s->slot_of, slot_of.data(), sizeof(int32_t) * (size_t)n0, hipMemcpyHostToDevice);
    if (e != hipSuccess) rc = check_hip(ctx, e, "cf_create: upload");
  }
  if (rc) {
    cf_free(s);
    delete s;
    return rc;
  }
  *out = s;
  return HFPX_OK;
}

int hfpx_cf_destroy(hfpx_cf_sys* s) {
  if (!s) return HFPX_OK;
  cf_free(s);
  delete s;
  return HFPX_OK;
}

int hfpx_vc_cf_iteration(hfpx_ctx* ctx, hfpx_cf_sys* S, const int32_t* mat_id, const int32_t* c_model_id_host,
                         const hfpx_fc_param* f_c_param_host, int64_t n_mat, const double* sinf_host,
                         int64_t inj_el, int64_t inj_node, double new_vol, double* dd, int64_t ldd, double* pp,
                         const double* mr_open, const double* mr_slip, double* friction_coef,
                         double* slip_cohesion, double* open_cohesion, int64_t* dof_h_host, int64_t* n_dof,
                         double tol, int max_outer, hfpx_cf_report* rep) {
  if (!ctx) return HFPX_ERR_INVALID;
  if (!S || !mat_id || !f_c_param_host || !sinf_host || !dd || !pp || !mr_open || !mr_slip || !friction_coef ||
      !slip_cohesion || !open_cohesion || !dof_h_host || !n_dof || !rep)
    return set_error(ctx, HFPX_ERR_INVALID, "vc_cf_iteration: null argument");
  const int64_t n0 = S->n0, n_el = S->n_el, n_nod = 6 * n_el, full = 18 * n_el;
  if (ldd < n_nod) return set_error(ctx, HFPX_ERR_INVALID, "vc_cf_iteration: ldd below 6 n_el");
  if (inj_el < 0 || inj_el >= n_el || inj_node < 0 || inj_node >= 6)
    return set_error(ctx, HFPX_ERR_INVALID, "vc_cf_iteration: injection point outside the mesh");
  for (int k = 0; k < 6; ++k)
    if (!std::isfinite(sinf_host[k])) return set_error(ctx, HFPX_ERR_INVALID, "vc_cf_iteration: far field not finite");
  if (!std::isfinite(new_vol)) return set_error(ctx, HFPX_ERR_INVALID, "vc_cf_iteration: volume not finite");
  *rep = hfpx_cf_report{};
  // the current handle: every slot it uses must be free in the original one
  std::vector<int32_t> cur((size_t)full);
  for (int64_t e = 0; e < n_el; ++e)
    for (int64_t l = 0; l < 18; ++l) {
      const int64_t d = dof_h_host[e + n_el * l];
      if (d < -1 || d >= n0 || (d >= 0 && S->orig[(size_t)(18 * e + l)] < 0))
        return set_error(ctx, HFPX_ERR_INVALID, "vc_cf_iteration: current handle slot (%lld, %lld) is %lld, not a "
                         "DoF of the original handle", (long long)e, (long long)l, (long long)d);
      cur[(size_t)(18 * e + l)] = (int32_t)d;
    }
  HFPX_TRY_HIP(ctx, hipSetDevice(ctx->device));
  hfpx_mesh* m = S->mesh;
  const int32_t* od = m->identity_dof ? nullptr : m->dof;
  HFPX_TRY_HIP(ctx, hipMemcpyAsync(S->cur, cur.data(), sizeof(int32_t) * cur.size(), hipMemcpyHostToDevice,
                                   ctx->stream));
  // friction and cohesion of the current DD (synchronous; validates the
  // material ids and parameters)
  int rc = hfpx_match_f_c(ctx, n_nod, dd, ldd, mat_id, c_model_id_host, f_c_param_host, n_mat, mr_open, mr_slip,
                          friction_coef, slip_cohesion, open_cohesion);
  if (rc) return rc;
  if (S->cr_cap < n_mat) {
    if (S->cr_open) HFPX_TRY_HIP(ctx, hipFree(S->cr_open));
    S->cr_open = nullptr;
    S->cr_cap = 0;
    if ((rc = cf_alloc(ctx, &S->cr_open, (size_t)n_mat))) return rc;
    S->cr_cap = n_mat;
  }
  std::vector<double> cro((size_t)n_mat);
  for (int64_t k = 0; k < n_mat; ++k) cro[(size_t)k] = f_c_param_host[k].cr_open;
  HFPX_TRY_HIP(ctx, hipMemcpyAsync(S->cr_open, cro.data(), sizeof(double) * cro.size(), hipMemcpyHostToDevice,
                                   ctx->stream));
  const int tok = prof_begin(ctx, HFPX_KC_VC);
  // 1. DD of the active DoF and their tractions / the current volume
  hipLaunchKernelGGL(k_cf_ddv, dim3(blocks(full, 256)), dim3(256), 0, ctx->stream, n_el, od, S->cur, dd, ldd, pp,
                     6 * inj_el + inj_node, S->ddv, S->nxt, S->scal);
  hipLaunchKernelGGL(k_cf_fill, dim3(blocks(n0 + 1, 256)), dim3(256), 0, ctx->stream, S->eneg, n0 + 1, 0.0);
  HFPX_TRY_HIP(ctx, hipGetLastError());
  if ((rc = gemm_sub_any(ctx, HFPX_F64, S->eneg, n0 + 1, S->A, S->lda, S->ddv, n0, n0 + 1, 1, n0))) return rc;
  // algorithmic HBM bytes of the traction product: the n0 + 1 rows of the
  // first n0 columns of the matrix, read once
  const double nn = (double)n0 * (double)(n0 + 1);
  prof_end(ctx, tok, HFPX_KC_VC, 2.0 * nn, 8.0 * nn);
  // 2. admissibility node by node
  CfLoad F;
  for (int k = 0; k < 6; ++k) F.s[k] = sinf_host[k];
  hipLaunchKernelGGL(k_cf_nodes, dim3(blocks(n_nod, 256)), dim3(256), 0, ctx->stream, m->elem, m->cp, n_el, od,
                     S->cur, S->eneg, pp, friction_coef, slip_cohesion, open_cohesion, mr_open, mr_slip, F, S->dtr,
                     S->act);
  // 3. renumbering
  hipLaunchKernelGGL(k_cf_count, dim3((unsigned)S->nblk), dim3(CF_BLK), 0, ctx->stream, S->act, n0, S->bsum);
  hipLaunchKernelGGL(k_cf_scan, dim3(1), dim3(64), 0, ctx->stream, S->bsum, S->nblk, S->iscal);
  hipLaunchKernelGGL(k_cf_number, dim3((unsigned)S->nblk), dim3(CF_BLK), 0, ctx->stream, S->act, S->act_f,
                     S->cached ? 1 : 0, n0, S->bsum, S->slot_of, S->idx, S->nxt, S->iscal);
  HFPX_TRY_HIP(ctx, hipGetLastError());
  int32_t hs[2] = {0, 0};
  std::vector<int32_t> nxt((size_t)full);
  HFPX_TRY_HIP(ctx, hipMemcpyAsync(hs, S->iscal, sizeof(hs), hipMemcpyDeviceToHost, ctx->stream));
  HFPX_TRY_HIP(ctx, hipMemcpyAsync(nxt.data(), S->nxt, sizeof(int32_t) * nxt.size(), hipMemcpyDeviceToHost,
                                   ctx->stream));
  HFPX_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const int64_t nu = hs[0];
  const bool changed = hs[1] != 0 || !S->cached || nu != S->n_cached;
  for (int64_t e = 0; e < n_el; ++e)
    for (int64_t l = 0; l < 18; ++l) dof_h_host[e + n_el * l] = nxt[(size_t)(18 * e + l)];
  *n_dof = nu;
  rep->n_dof = nu;
  int solve_rc = HFPX_OK;
  if (nu == 0) {
    // no DoF stays active: no system, zero increments
    hipLaunchKernelGGL(k_cf_fill, dim3(1), dim3(64), 0, ctx->stream, S->x, (int64_t)1, 0.0);
    HFPX_TRY_HIP(ctx, hipGetLastError());
  } else {
    const int64_t n1 = nu + 1;
    // the list is read for rows < nu and closed with the border at nu
    hipLaunchKernelGGL(k_cf_rhs, dim3(blocks(n1, 256)), dim3(256), 0, ctx->stream, S->idx, nu, S->dtr, S->eneg, n0,
                       new_vol, S->rhs);
    HFPX_TRY_HIP(ctx, hipGetLastError());
    if (changed) {
      // 4. truncated system and its FP32 factors
      S->cached = false;
      if ((rc = launch_vc_gather_idx(ctx, S->A, S->lda, S->idx, n1, S->M64, n1, S->M32, n1))) return rc;
      int64_t info = 0;
      rc = lu_getrf(ctx, HFPX_F32, S->M32, n1, n1, S->ipiv, &info);
      if (rc == HFPX_ERR_SINGULAR)
        return set_error(ctx, rc, "vc_cf_iteration: truncated system singular in FP32 (U(%lld,%lld) = 0)",
                         (long long)info, (long long)info);
      if (rc) return rc;
      hipLaunchKernelGGL(k_cf_copy_i32, dim3(blocks(n0, 256)), dim3(256), 0, ctx->stream, S->act, S->act_f, n0);
      HFPX_TRY_HIP(ctx, hipGetLastError());
      S->cached = true;
      S->n_cached = nu;
      rep->refactored = 1;
    }
    // 5. FP64 refinement on the truncated system
    solve_rc = solve_ir(ctx, S->M64, n1, S->M32, n1, S->ipiv, n1, S->rhs, n1, S->x, n1, 1, tol, max_outer, 20,
                        &rep->ir);
    if (solve_rc != HFPX_OK && solve_rc != HFPX_ERR_NOT_CONVERGED) return solve_rc;
  }
  // 6. DD and pressure
  hipLaunchKernelGGL(k_cf_update, dim3(blocks(n_el, 64)), dim3(64), 0, ctx->stream, m->elem, m->cp, n_el, m->dd_local,
                     od, S->nxt, S->x, nu, mat_id, S->cr_open, mr_open, dd, ldd, pp, S->scal);
  // 7. residual
  const int64_t nr = nu + 1 + n0, nb = (nr + CF_RED - 1) / CF_RED;
  hipLaunchKernelGGL(k_cf_l1, dim3((unsigned)nb), dim3(256), 0, ctx->stream, S->x, nu + 1, S->dtr, n0, S->part);
  hipLaunchKernelGGL(k_cf_l1_sum, dim3(1), dim3(64), 0, ctx->stream, S->part, nb, S->scal + 2);
  HFPX_TRY_HIP(ctx, hipGetLastError());
  double hsc[3];
  HFPX_TRY_HIP(ctx, hipMemcpyAsync(hsc, S->scal, sizeof(hsc), hipMemcpyDeviceToHost, ctx->stream));
  HFPX_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  rep->pressure = hsc[1];
  rep->residual = hsc[2];
  if (solve_rc == HFPX_ERR_NOT_CONVERGED)
    return set_error(ctx, solve_rc, "vc_cf_iteration: refinement of the truncated system did not converge "
                     "(backward error %.3e)", rep->ir.backward_error);
  return HFPX_OK;
}

}  // extern "C"
