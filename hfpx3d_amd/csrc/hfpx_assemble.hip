// hfpx_assemble.hip -- element geometry, dense influence-matrix assembly and
// right-hand sides on gfx950.
//
// Replaces, on the device:
//   hfp3d::make_el_r_tensor / make_el_tau_crd / make_el_sfm_uniform /
//   el_cp_uniform          src/Core/element_utilities.cpp:62-238, 364-395
//   hfp3d::make_el_sfm_nonuniform / el_cp_nonuniform (per-node vertex
//   weights)               src/Core/element_utilities.cpp:240-322, 397-431
//   hfp3d::make_3dbem_matrix_s   src/Solvers/system_assembly.cpp:30-219
//   hfp3d::add_s_inf_to_3dbem_rhs src/Solvers/system_assembly.cpp:222-274
//
// Layout in HBM: A is column-major (il::Array2D), row = target DoF
// (collocation point x traction component), column = source DoF (node x DD
// component), both global-frame, 18x18 blocks per element pair.
//
// Work decomposition: one lane = one (collocation point, source element)
// pair, i.e. a 3 x 18 slab of A.  A wavefront holds 64 consecutive
// collocation points against ONE source element, so the source geometry is
// wave-uniform (read once through the scalar path) while the target data is
// per lane; the 64 lanes together write 192 contiguous rows of each of the
// source's 18 columns.
#include "hfpx_internal.h"

namespace hfpx {

// ---------------------------------------------------------------------
// element geometry: one thread per element
// ---------------------------------------------------------------------
// wts: per-node vertex weights of a non-uniform edge partition, or null for
// the uniform element (edge nodes at mid-edge).
__global__ void k_geometry(const double* __restrict__ nods, const int32_t* __restrict__ conn,
                           const double* __restrict__ wts, int64_t n_el, double beta,
                           SrcElem<double>* __restrict__ elem, double* __restrict__ cp) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_el) return;
  double v[3][3], w[3] = {1.0, 1.0, 1.0};
  for (int j = 0; j < 3; ++j) {
    const int64_t n = conn[3 * e + j];
    for (int k = 0; k < 3; ++k) v[j][k] = nods[3 * n + k];
    if (wts) w[j] = wts[n];
  }
  // frame: e1 = (v1-v0)/|.|, e3 = normalize(e1 x (v2-v0)), e2 = e3 x e1
  double a1[3], a2[3], e1[3], e2[3], e3[3];
  for (int k = 0; k < 3; ++k) {
    a1[k] = v[1][k] - v[0][k];
    a2[k] = v[2][k] - v[0][k];
  }
  double n1 = sqrt(a1[0] * a1[0] + a1[1] * a1[1] + a1[2] * a1[2]);
  for (int k = 0; k < 3; ++k) e1[k] = a1[k] / n1;
  e3[0] = e1[1] * a2[2] - e1[2] * a2[1];
  e3[1] = e1[2] * a2[0] - e1[0] * a2[2];
  e3[2] = e1[0] * a2[1] - e1[1] * a2[0];
  double n3 = sqrt(e3[0] * e3[0] + e3[1] * e3[1] + e3[2] * e3[2]);
  for (int k = 0; k < 3; ++k) e3[k] /= n3;
  e2[0] = e3[1] * e1[2] - e3[2] * e1[1];
  e2[1] = e3[2] * e1[0] - e3[0] * e1[2];
  e2[2] = e3[0] * e1[1] - e3[1] * e1[0];
  double n2 = sqrt(e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2]);
  for (int k = 0; k < 3; ++k) e2[k] /= n2;

  SrcElem<double> E;
  for (int k = 0; k < 3; ++k) {
    E.R[k] = e1[k];
    E.R[3 + k] = e2[k];
    E.R[6 + k] = e3[k];
    E.v0[k] = v[0][k];
  }
  double tr[3], ti[3];
  for (int j = 0; j < 3; ++j) {
    double d[3] = {v[j][0] - v[0][0], v[j][1] - v[0][1], v[j][2] - v[0][2]};
    tr[j] = e1[0] * d[0] + e1[1] * d[1] + e1[2] * d[2];
    ti[j] = e2[0] * d[0] + e2[1] * d[1] + e2[2] * d[2];
    E.t[2 * j] = tr[j];
    E.t[2 * j + 1] = ti[j];
  }
  for (int m = 0; m < 3; ++m) {
    const int q = (m + 1) % 3;
    const double dx = tr[q] - tr[m], dy = ti[q] - ti[m];
    const double L = sqrt(dx * dx + dy * dy);
    E.len[m] = L;
    E.th[2 * m] = dx / L;
    E.th[2 * m + 1] = dy / L;
  }
  // barycentric L_k(t) = Im(conj(t_a - t)(t_b - t)) / area2, (a, b) = (k+1, k+2)
  const double area2 = (tr[1] - tr[0]) * (ti[2] - ti[0]) - (ti[1] - ti[0]) * (tr[2] - tr[0]);
  for (int k = 0; k < 3; ++k) {
    const int a = (k + 1) % 3, b = (k + 2) % 3;
    E.l0[k] = (tr[a] * ti[b] - ti[a] * tr[b]) / area2;
    // l1 = (conj(t_b) - conj(t_a)) / (2 i area2)
    const double xr = tr[b] - tr[a], xi = -(ti[b] - ti[a]);
    E.l1[2 * k] = xi / (2.0 * area2);
    E.l1[2 * k + 1] = -xr / (2.0 * area2);
  }
  // Lagrange basis of the six nodes: the edge node opposite v sits at
  // L_a = w_a / (w_a + w_b) on the edge (a, b) = (v+1, v+2)
  for (int q = 0; q < 3; ++q) {
    const int a = (q + 1) % 3, b = (q + 2) % 3;
    E.sw[2 * q] = w[q] / w[a];
    E.sw[2 * q + 1] = w[q] / w[b];
    E.sw[6 + q] = (w[a] + w[b]) * (w[a] + w[b]) / (w[a] * w[b]);
  }
  elem[e] = E;

  // collocation points (vertex-ward and mid-edge, pulled towards the centroid)
  const int64_t ncp = 6 * n_el;
  double c[3];
  for (int k = 0; k < 3; ++k) c[k] = (v[0][k] + v[1][k] + v[2][k]) / 3.0;
  for (int q = 0; q < 3; ++q) {
    const int m = (q + 1) % 3, l = (m + 1) % 3;
    for (int k = 0; k < 3; ++k) {
      cp[k * ncp + 6 * e + q] = (1.0 - beta) * v[q][k] + beta * c[k];
      cp[k * ncp + 6 * e + 3 + q] =
          wts ? (1.0 - beta) * (w[m] * v[m][k] + w[l] * v[l][k]) / (w[m] + w[l]) + beta * c[k]
              : 0.5 * (1.0 - beta) * (v[m][k] + v[l][k]) + beta * c[k];
    }
  }
  for (int q = 0; q < 6; ++q)
    for (int k = 0; k < 3; ++k) cp[(3 + k) * ncp + 6 * e + q] = -e3[k];
}

// ---------------------------------------------------------------------
// assembly
// ---------------------------------------------------------------------
struct StoreBoth {
  double* a64;
  int64_t lda64;
  float* a32;
  int64_t lda32;
  __device__ __forceinline__ void operator()(int64_t r, int64_t c, double v) const {
    if (a64) a64[r + c * lda64] = v;
    if (a32) a32[r + c * lda32] = (float)v;
  }
};

// One wavefront's share of the matrix: collocation points 64 bx .. 64 bx + 63
// of a mesh of n_el elements against source element 4 by + wave.  cp holds
// the six coordinate planes of the points with plane stride cps (>= 6 n_el;
// larger when the mesh is one crack of a batch).
template <bool NU, typename Store>
__device__ __forceinline__ void assemble_wave(const SrcElem<double>* __restrict__ elem,
                                              const double* __restrict__ cp, int64_t cps, int64_t n_el,
                                              const int32_t* __restrict__ dof, double shear_m, double nu,
                                              bool dd_local, const Store& store, int64_t bx, int64_t by) {
  const int lane = threadIdx.x & 63;
  const int64_t es = __builtin_amdgcn_readfirstlane((int)(by * 4 + (threadIdx.x >> 6)));
  if (es >= n_el) return;
  const int64_t ncp = 6 * n_el;
  const int64_t icp = bx * 64 + lane;
  if (icp >= ncp) return;
  const SrcElem<double>& E = elem[es];

  const double xg[3] = {cp[icp], cp[cps + icp], cp[2 * cps + icp]};
  const double ng[3] = {cp[3 * cps + icp], cp[4 * cps + icp], cp[5 * cps + icp]};
  const double d[3] = {xg[0] - E.v0[0], xg[1] - E.v0[1], xg[2] - E.v0[2]};
  const double y0 = E.R[0] * d[0] + E.R[1] * d[1] + E.R[2] * d[2];
  const double y1 = E.R[3] * d[0] + E.R[4] * d[1] + E.R[5] * d[2];
  const double y2 = E.R[6] * d[0] + E.R[7] * d[1] + E.R[8] * d[2];
  const double nl[3] = {E.R[0] * ng[0] + E.R[1] * ng[1] + E.R[2] * ng[2],
                        E.R[3] * ng[0] + E.R[4] * ng[1] + E.R[5] * ng[2],
                        E.R[6] * ng[0] + E.R[7] * ng[1] + E.R[8] * ng[2]};
  int64_t row[3];
  if (dof) {
    const int64_t et = icp / 6, nt = icp % 6;
    for (int i = 0; i < 3; ++i) row[i] = dof[18 * et + 3 * nt + i];
  } else {
    for (int i = 0; i < 3; ++i) row[i] = 3 * icp + i;
  }
  pair_influence<double, NU>(E, y0, y1, y2, shear_m, nu, [&](int j, const double (&S)[6][3]) {
    // traction in the source frame for each local DD component k
    double tl[3][3];
    for (int k = 0; k < 3; ++k) {
      tl[0][k] = nl[0] * S[0][k] + nl[1] * S[3][k] + nl[2] * S[4][k];
      tl[1][k] = nl[0] * S[3][k] + nl[1] * S[1][k] + nl[2] * S[5][k];
      tl[2][k] = nl[0] * S[4][k] + nl[1] * S[5][k] + nl[2] * S[2][k];
    }
    // T = R^T tl R  (traction to global, DD component to global)
    double tg[3][3];
    for (int i = 0; i < 3; ++i)
      for (int k = 0; k < 3; ++k)
        tg[i][k] = E.R[i] * tl[0][k] + E.R[3 + i] * tl[1][k] + E.R[6 + i] * tl[2][k];
    for (int jg = 0; jg < 3; ++jg) {
      int64_t col;
      if (dof) col = dof[18 * es + 3 * j + jg];
      else col = 18 * es + 3 * j + jg;
      if (col < 0) continue;
      for (int i = 0; i < 3; ++i) {
        if (row[i] < 0) continue;
        // global DD: T R per source node; local DD keep T (wave-uniform branch)
        const double v =
            dd_local ? tg[i][jg] : tg[i][0] * E.R[jg] + tg[i][1] * E.R[3 + jg] + tg[i][2] * E.R[6 + jg];
        store(row[i], col, v);
      }
    }
  });
}

template <bool NU, typename Store>
__global__ __launch_bounds__(256) void k_assemble(const SrcElem<double>* __restrict__ elem,
                                                  const double* __restrict__ cp, int64_t n_el,
                                                  const int32_t* __restrict__ dof, double shear_m,
                                                  double nu, bool dd_local, Store store) {
  assemble_wave<NU>(elem, cp, 6 * n_el, n_el, dof, shear_m, nu, dd_local, store, blockIdx.x, blockIdx.y);
}

// ---------------------------------------------------------------------
// batched assembly: workgroup w belongs to the crack c with
// cr[c].wg_off <= w < cr[c + 1].wg_off (binary search; the table is in
// workgroup order) and does that crack's tile (local % gx, local / gx) of
// the same decomposition as k_assemble.  FP64 matrix always, FP32 copy for
// the cracks that have one.
// ---------------------------------------------------------------------
__device__ __forceinline__ int find_crack(const BCrack* __restrict__ cr, int ncr, int64_t w, bool by_el) {
  int lo = 0, hi = ncr - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    const int64_t start = by_el ? cr[mid].el_off : cr[mid].wg_off;
    if (start <= w)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

__global__ __launch_bounds__(256) void k_bassemble(const SrcElem<double>* __restrict__ elem,
                                                   const double* __restrict__ cp, int64_t cps,
                                                   const BCrack* __restrict__ cr, int ncr, double* __restrict__ a64,
                                                   float* __restrict__ a32) {
  const int64_t w = blockIdx.x;
  const BCrack c = cr[find_crack(cr, ncr, w, false)];
  const int64_t gx = (6 * c.n_el + 63) / 64;
  const int64_t local = w - c.wg_off;
  StoreBoth st{a64 + c.a_off, c.n, c.a32_off >= 0 ? a32 + c.a32_off : nullptr, c.n};
  assemble_wave<false>(elem + c.el_off, cp + 6 * c.el_off, cps, c.n_el, nullptr, c.G, c.nu, false, st, local % gx,
                local / gx);
}

// batched right-hand sides: one thread per element of the batch,
// b[x_off + DoF] = -(traction of the crack's far-field stress); every DoF is
// written exactly once (all DoF free)
__global__ void k_brhs(const SrcElem<double>* __restrict__ elem, int64_t n_el_total, const BCrack* __restrict__ cr,
                       int ncr, double* __restrict__ b) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_el_total) return;
  const BCrack& c = cr[find_crack(cr, ncr, e, true)];
  const double* s = c.sinf;
  const double n[3] = {-elem[e].R[6], -elem[e].R[7], -elem[e].R[8]};
  const double tr[3] = {n[0] * s[0] + n[1] * s[3] + n[2] * s[4], n[0] * s[3] + n[1] * s[1] + n[2] * s[5],
                        n[0] * s[4] + n[1] * s[5] + n[2] * s[2]};
  double* bc = b + c.x_off + 18 * (e - c.el_off);
  for (int ln = 0; ln < 6; ++ln)
    for (int l = 0; l < 3; ++l) bc[3 * ln + l] = -tr[l];
}

// ---------------------------------------------------------------------
// right-hand side: one thread per (element, rhs)
// ---------------------------------------------------------------------
__global__ void k_rhs(const SrcElem<double>* __restrict__ elem, int64_t n_el,
                      const int32_t* __restrict__ dof, const double* __restrict__ sinf, int nrhs,
                      double* __restrict__ b, int64_t ldb) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_el * nrhs) return;
  const int64_t e = t % n_el, r = t / n_el;
  const double* s = sinf + 6 * r;
  const double n[3] = {-elem[e].R[6], -elem[e].R[7], -elem[e].R[8]};
  const double tr[3] = {n[0] * s[0] + n[1] * s[3] + n[2] * s[4],
                        n[0] * s[3] + n[1] * s[1] + n[2] * s[5],
                        n[0] * s[4] + n[1] * s[5] + n[2] * s[2]};
  for (int ln = 0; ln < 6; ++ln)
    for (int l = 0; l < 3; ++l) {
      const int64_t d = dof ? (int64_t)dof[18 * e + 3 * ln + l] : 18 * e + 3 * ln + l;
      if (d >= 0) b[d + r * ldb] -= tr[l];
    }
}

// ---------------------------------------------------------------------
// the element integral alone, on the reference's own arguments: one thread
// per case (h, z, tau, sfm) in the source element's frame -> the 6 x 18
// stress block (rows S11 S22 S33 S12 S13 S23, column 3 node + DD component).
// tau: vertex coordinates (re, im) x 3; sfm: 6 x 6 complex column-major
// (re, im interleaved), row = shape function, column = monomial 1, tau,
// conj(tau), tau^2, conj(tau)^2, tau conj(tau), re-centred at z by the
// binomial shift of shift_el_sfm (element_utilities.cpp:339-360).
// ---------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_local_submatrix(double shear_m, double nu, int64_t n_cases,
                                                        const double* __restrict__ h, const double* __restrict__ z,
                                                        const double* __restrict__ tau,
                                                        const double* __restrict__ sfm, double* __restrict__ out) {
  const int64_t c = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (c >= n_cases) return;
  SrcElem<double> E;  // the integral reads the edge data only
  for (int k = 0; k < 6; ++k) E.t[k] = tau[6 * c + k];
  for (int m = 0; m < 3; ++m) {
    const int q = (m + 1) % 3;
    const double dx = E.t[2 * q] - E.t[2 * m], dy = E.t[2 * q + 1] - E.t[2 * m + 1];
    const double L = sqrt(dx * dx + dy * dy);
    E.len[m] = L;
    E.th[2 * m] = dx / L;
    E.th[2 * m + 1] = dy / L;
  }
  const cx<double> zz = mk(z[2 * c], z[2 * c + 1]), zb = conj(zz);
  const double* q = sfm + 72 * c;
  auto dens = [&](int j, double& p00, cx<double>& p10, cx<double>& p20, double& p11) {
    cx<double> a[6];
    for (int col = 0; col < 6; ++col) a[col] = mk(q[2 * (j + 6 * col)], q[2 * (j + 6 * col) + 1]);
    p00 = (a[0] + a[1] * zz + a[2] * zb + a[3] * (zz * zz) + a[4] * (zb * zb) + a[5] * (zz * zb)).re;
    p10 = a[1] + 2.0 * (a[3] * zz) + a[5] * zb;
    p20 = a[3];
    p11 = a[5].re;
  };
  double* o = out + 108 * c;
  pair_influence_dens<double>(E, zz.re, zz.im, -h[c], shear_m, nu, dens, [&](int j, const double (&S)[6][3]) {
    for (int k = 0; k < 3; ++k)
      for (int r = 0; r < 6; ++r) o[r + 6 * (3 * j + k)] = S[r][k];
  });
}

// ---------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------
int launch_geometry(hfpx_ctx* ctx, hfpx_mesh* m) {
  const int bs = 128;
  const unsigned gs = (unsigned)((m->n_el + bs - 1) / bs);
  hipLaunchKernelGGL(k_geometry, dim3(gs), dim3(bs), 0, ctx->stream, m->nods, m->conn, m->wts, m->n_el,
                     m->beta, m->elem, m->cp);
  HFPX_TRY_HIP(ctx, hipGetLastError());
  return HFPX_OK;
}

int launch_assemble(hfpx_ctx* ctx, const hfpx_mesh* m, double shear_m, double poiss_r,
                    double* a64, int64_t lda64, float* a32, int64_t lda32) {
  const int64_t ncp = 6 * m->n_el;
  const unsigned gx = (unsigned)((ncp + 63) / 64);
  const unsigned gy = (unsigned)((m->n_el + 3) / 4);
  if (gy > 65535u) return set_error(ctx, HFPX_ERR_INVALID, "too many elements for one launch");
  StoreBoth st{a64, lda64, a32, lda32};
  const int tok = prof_begin(ctx, HFPX_KC_ASSEMBLE);
  if (m->wts)
    hipLaunchKernelGGL((k_assemble<true, StoreBoth>), dim3(gx, gy), dim3(256), 0, ctx->stream, m->elem,
                       m->cp, m->n_el, m->identity_dof ? nullptr : m->dof, shear_m, poiss_r, m->dd_local, st);
  else
    hipLaunchKernelGGL((k_assemble<false, StoreBoth>), dim3(gx, gy), dim3(256), 0, ctx->stream, m->elem,
                       m->cp, m->n_el, m->identity_dof ? nullptr : m->dof, shear_m, poiss_r, m->dd_local, st);
  // algorithmic HBM bytes: every matrix entry written once (FP64 and/or FP32)
  const double entries = (double)m->n_dof * (double)m->n_dof;
  prof_end(ctx, tok, HFPX_KC_ASSEMBLE, 0.0, entries * ((a64 ? 8.0 : 0.0) + (a32 ? 4.0 : 0.0)));
  HFPX_TRY_HIP(ctx, hipGetLastError());
  return HFPX_OK;
}

int launch_local_submatrix(hfpx_ctx* ctx, double shear_m, double poiss_r, int64_t n_cases, const double* h,
                           const double* z, const double* tau, const double* sfm, double* out) {
  const int64_t nb = (n_cases + 63) / 64;
  if (nb > 0x7fffffff) return set_error(ctx, HFPX_ERR_INVALID, "local submatrix: too many cases");
  hipLaunchKernelGGL(k_local_submatrix, dim3((unsigned)nb), dim3(64), 0, ctx->stream, shear_m, poiss_r, n_cases, h,
                     z, tau, sfm, out);
  HFPX_TRY_HIP(ctx, hipGetLastError());
  return HFPX_OK;
}

int launch_rhs(hfpx_ctx* ctx, const hfpx_mesh* m, const double* sinf_dev, int nrhs, double* b,
               int64_t ldb) {
  const int bs = 128;
  const int64_t nt = m->n_el * nrhs;
  hipLaunchKernelGGL(k_rhs, dim3((unsigned)((nt + bs - 1) / bs)), dim3(bs), 0, ctx->stream,
                     m->elem, m->n_el, m->identity_dof ? nullptr : m->dof, sinf_dev, nrhs, b, ldb);
  HFPX_TRY_HIP(ctx, hipGetLastError());
  return HFPX_OK;
}

}  // namespace hfpx

namespace hfpx {

int launch_bassemble(hfpx_ctx* ctx, const SrcElem<double>* elem, const double* cp, int64_t cps, const BCrack* cr,
                     int ncr, int64_t total_wg, double* a64, float* a32) {
  if (total_wg <= 0) return HFPX_OK;
  hipLaunchKernelGGL(k_bassemble, dim3((unsigned)total_wg), dim3(256), 0, ctx->stream, elem, cp, cps, cr, ncr, a64,
                     a32);
  HFPX_TRY_HIP(ctx, hipGetLastError());
  return HFPX_OK;
}

int launch_brhs(hfpx_ctx* ctx, const SrcElem<double>* elem, int64_t n_el_total, const BCrack* cr, int ncr,
                double* b) {
  hipLaunchKernelGGL(k_brhs, dim3((unsigned)((n_el_total + 127) / 128)), dim3(128), 0, ctx->stream, elem,
                     n_el_total, cr, ncr, b);
  HFPX_TRY_HIP(ctx, hipGetLastError());
  return HFPX_OK;
}

}  // namespace hfpx
