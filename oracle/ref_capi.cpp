// Test-infrastructure harness: a thin extern "C" layer over the REFERENCE
// HFPx3D sources, compiled in place from the reference tree (see Makefile in
// this directory).  It is the strongest pin for the oracle and for the HIP
// path: every value it returns is computed by the reference's own code.
//
// Only tests and the CPU-baseline leg of bench.py load the library built from
// this file; the product never links it.
//
// Functions wrapped (paths relative to the reference root):
//   hfp3d::make_local_3dbem_submatrix  src/Elasticity/elasticity_kernel_integration.cpp:34-205
//   hfp3d::make_el_sfm_uniform          src/Core/element_utilities.cpp:170-238
//   hfp3d::make_el_tau_crd              src/Core/element_utilities.cpp:89-118
//   hfp3d::el_cp_uniform                src/Core/element_utilities.cpp:364-395
//   hfp3d::make_el_sfm_nonuniform       src/Core/element_utilities.cpp:240-322
//   hfp3d::el_cp_nonuniform             src/Core/element_utilities.cpp:397-431
//   hfp3d::make_dof_h_crack             src/Core/surface_mesh_utilities.cpp:28-112
//   hfp3d::make_3dbem_matrix_s          src/Solvers/system_assembly.cpp:30-219
//   hfp3d::add_s_inf_to_3dbem_rhs       src/Solvers/system_assembly.cpp:222-274
//   hfp3d::make_3dbem_stress_f_s        src/Solvers/system_assembly.cpp:279-455
//   hfp3d::el_p2_sf_integral            src/Core/element_utilities.cpp:573-595
//   hfp3d::make_3dbem_matrix_vc         src/Solvers/system_assembly.cpp:458-557
//   hfp3d::mod_3dbem_system_vc          src/Solvers/system_assembly.cpp:560-629
//   hfp3d::match_f_c, hfp3d::update_mrd src/Development/cohesion_friction.cpp:22-189
//                                       (run on one thread: the reference shares
//                                       its nodal_dd scratch between threads)
//
// All arrays are column-major, like il::Array2D.
#include <chrono>
#include <complex>
#include <cstdint>
#include <cstring>
#include <il/Array.h>
#include <il/Array2D.h>
#include <il/StaticArray.h>
#include <il/StaticArray2D.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "src/Core/element_utilities.h"
#include "src/Core/model_parameters.h"
#include "src/Core/surface_mesh_utilities.h"
#include "src/Elasticity/elasticity_kernel_integration.h"
#include "src/Solvers/system_assembly.h"
#include "src/Development/cohesion_friction.h"

namespace {

void set_threads(int nthreads) {
#ifdef _OPENMP
  omp_set_num_threads(nthreads > 0 ? nthreads : 1);
#else
  (void)nthreads;
#endif
}

hfp3d::Mesh_Geom_T make_mesh(const double* nods, int64_t nods_rows,
                             int64_t n_nodes, const int64_t* conn,
                             int64_t conn_rows, int64_t n_el) {
  hfp3d::Mesh_Geom_T mesh;
  mesh.nods = il::Array2D<double>{nods_rows, n_nodes, 0.0};
  for (int64_t j = 0; j < n_nodes; ++j)
    for (int64_t i = 0; i < nods_rows; ++i)
      mesh.nods(i, j) = nods[i + nods_rows * j];
  mesh.conn = il::Array2D<il::int_t>{conn_rows, n_el, 0};
  for (int64_t j = 0; j < n_el; ++j)
    for (int64_t i = 0; i < conn_rows; ++i)
      mesh.conn(i, j) = conn[i + conn_rows * j];
  return mesh;
}

}  // namespace

extern "C" {

int ref_threads_available(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

// 6x18 local influence of one element on one point.
// tau: 3 complex (re, im interleaved); sfm: 6x6 complex col-major
// (entry (r, c) at 2*(r + 6*c)); out: 6x18 col-major.
int ref_local_submatrix(double shear_m, double poiss_r, double h, double z_re,
                        double z_im, const double* tau, const double* sfm,
                        double* out) {
  il::StaticArray<std::complex<double>, 3> t;
  for (int k = 0; k < 3; ++k) t[k] = std::complex<double>(tau[2 * k], tau[2 * k + 1]);
  il::StaticArray2D<std::complex<double>, 6, 6> s;
  for (int c = 0; c < 6; ++c)
    for (int r = 0; r < 6; ++r)
      s(r, c) = std::complex<double>(sfm[2 * (r + 6 * c)], sfm[2 * (r + 6 * c) + 1]);
  il::StaticArray2D<double, 6, 18> m = hfp3d::make_local_3dbem_submatrix(
      1, shear_m, poiss_r, h, std::complex<double>(z_re, z_im), t, s);
  for (int c = 0; c < 18; ++c)
    for (int r = 0; r < 6; ++r) out[r + 6 * c] = m(r, c);
  return 0;
}

// Element geometry: vert is 3x3 col-major (coordinate k of vertex j at
// k + 3*j).  Outputs: sfm (6x6 complex col-major), rotation tensor (3x3
// col-major), tau (3 complex), collocation points for beta (3x6 col-major).
int ref_el_geom(const double* vert, double beta, double* sfm, double* rt,
                double* tau, double* cp) {
  il::StaticArray2D<double, 3, 3> v;
  for (int j = 0; j < 3; ++j)
    for (int k = 0; k < 3; ++k) v(k, j) = vert[k + 3 * j];
  il::StaticArray2D<double, 3, 3> r;
  il::StaticArray2D<std::complex<double>, 6, 6> s =
      hfp3d::make_el_sfm_uniform(v, il::io, r);
  il::StaticArray<std::complex<double>, 3> t = hfp3d::make_el_tau_crd(v, r);
  il::StaticArray<il::StaticArray<double, 3>, 6> c = hfp3d::el_cp_uniform(v, beta);
  for (int cc = 0; cc < 6; ++cc)
    for (int rr = 0; rr < 6; ++rr) {
      sfm[2 * (rr + 6 * cc)] = s(rr, cc).real();
      sfm[2 * (rr + 6 * cc) + 1] = s(rr, cc).imag();
    }
  for (int j = 0; j < 3; ++j)
    for (int k = 0; k < 3; ++k) rt[k + 3 * j] = r(k, j);
  for (int k = 0; k < 3; ++k) {
    tau[2 * k] = t[k].real();
    tau[2 * k + 1] = t[k].imag();
  }
  for (int n = 0; n < 6; ++n)
    for (int k = 0; k < 3; ++k) cp[k + 3 * n] = c[n][k];
  return 0;
}

// Same for a non-uniform edge partition with vertex weights wts (3).
int ref_el_geom_nonuniform(const double* vert, const double* wts, double beta, double* sfm,
                           double* rt, double* cp) {
  il::StaticArray2D<double, 3, 3> v;
  for (int j = 0; j < 3; ++j)
    for (int k = 0; k < 3; ++k) v(k, j) = vert[k + 3 * j];
  il::StaticArray<double, 3> w;
  for (int j = 0; j < 3; ++j) w[j] = wts[j];
  il::StaticArray2D<double, 3, 3> r;
  il::StaticArray2D<std::complex<double>, 6, 6> s = hfp3d::make_el_sfm_nonuniform(v, w, il::io, r);
  il::StaticArray<il::StaticArray<double, 3>, 6> c = hfp3d::el_cp_nonuniform(v, w, beta);
  for (int cc = 0; cc < 6; ++cc)
    for (int rr = 0; rr < 6; ++rr) {
      sfm[2 * (rr + 6 * cc)] = s(rr, cc).real();
      sfm[2 * (rr + 6 * cc) + 1] = s(rr, cc).imag();
    }
  for (int j = 0; j < 3; ++j)
    for (int k = 0; k < 3; ++k) rt[k + 3 * j] = r(k, j);
  for (int n = 0; n < 6; ++n)
    for (int k = 0; k < 3; ++k) cp[k + 3 * n] = c[n][k];
  return 0;
}

// DoF handle: dof_h is n_el x 18 col-major; returns n_dof.
int64_t ref_dof_h(const double* nods, int64_t nods_rows, int64_t n_nodes,
                  const int64_t* conn, int64_t conn_rows, int64_t n_el,
                  int tip_type, int64_t* dof_h) {
  hfp3d::Mesh_Geom_T mesh = make_mesh(nods, nods_rows, n_nodes, conn, conn_rows, n_el);
  hfp3d::DoF_Handle_T d = hfp3d::make_dof_h_crack(mesh, 2, tip_type);
  for (int64_t c = 0; c < 18; ++c)
    for (int64_t e = 0; e < n_el; ++e) dof_h[e + n_el * c] = d.dof_h(e, c);
  return d.n_dof;
}

// Dense influence matrix.  A must hold n_dof*n_dof doubles (n_dof as returned
// by ref_dof_h for the same tip_type); col-major, ld = n_dof.
int64_t ref_assemble(const double* nods, int64_t nods_rows, int64_t n_nodes,
                     const int64_t* conn, int64_t conn_rows, int64_t n_el,
                     double shear_m, double poiss_r, double beta, int tip_type,
                     int is_dd_local, int nthreads, double* A) {
  set_threads(nthreads);
  hfp3d::Mesh_Geom_T mesh = make_mesh(nods, nods_rows, n_nodes, conn, conn_rows, n_el);
  hfp3d::Num_Param_T np;
  np.beta = beta;
  np.tip_type = tip_type;
  np.is_dd_local = is_dd_local != 0;
  hfp3d::DoF_Handle_T d;
  il::Array2D<double> m = hfp3d::make_3dbem_matrix_s(shear_m, poiss_r, mesh, np, il::io, d);
  std::memcpy(A, m.data(), sizeof(double) * m.size(0) * m.size(1));
  return d.n_dof;
}

// Right-hand side: rhs (n_dof) is accumulated into (caller zero-fills).
int64_t ref_rhs(const double* nods, int64_t nods_rows, int64_t n_nodes,
                const int64_t* conn, int64_t conn_rows, int64_t n_el,
                int tip_type, const double* s_inf, double* rhs) {
  hfp3d::Mesh_Data_T md;
  md.mesh = make_mesh(nods, nods_rows, n_nodes, conn, conn_rows, n_el);
  md.dof_h_dd = hfp3d::make_dof_h_crack(md.mesh, 2, tip_type);
  hfp3d::Load_T load;
  for (int k = 0; k < 6; ++k) load.s_inf[k] = s_inf[k];
  hfp3d::SAE_T sae;
  sae.rhs_v = il::Array<double>{md.dof_h_dd.n_dof, 0.0};
  hfp3d::add_s_inf_to_3dbem_rhs(md, load, il::io, sae);
  for (il::int_t i = 0; i < md.dof_h_dd.n_dof; ++i) rhs[i] = sae.rhs_v[i];
  return md.dof_h_dd.n_dof;
}

// Stresses at n_pts points.  dd: (6*n_el) x 3 col-major (global comps);
// pts: 3 x n_pts col-major; out: n_pts x 6 col-major.  Call with nthreads=1
// for deterministic values (the reference accumulation is unsynchronised).
int ref_stress(const double* nods, int64_t nods_rows, int64_t n_nodes,
               const int64_t* conn, int64_t conn_rows, int64_t n_el,
               double shear_m, double poiss_r, const double* dd,
               const double* pts, int64_t n_pts, const double* s_inf,
               int is_dd_local, int nthreads, double* out) {
  set_threads(nthreads);
  hfp3d::Mesh_Data_T md;
  md.mesh = make_mesh(nods, nods_rows, n_nodes, conn, conn_rows, n_el);
  md.dd = il::Array2D<double>{6 * n_el, 3, 0.0};
  for (int64_t j = 0; j < 3; ++j)
    for (int64_t i = 0; i < 6 * n_el; ++i) md.dd(i, j) = dd[i + 6 * n_el * j];
  hfp3d::Load_T load;
  for (int k = 0; k < 6; ++k) load.s_inf[k] = s_inf[k];
  hfp3d::Num_Param_T np;
  np.is_dd_local = is_dd_local != 0;
  il::Array2D<double> p{3, n_pts, 0.0};
  for (int64_t j = 0; j < n_pts; ++j)
    for (int64_t i = 0; i < 3; ++i) p(i, j) = pts[i + 3 * j];
  il::Array2D<double> s = hfp3d::make_3dbem_stress_f_s(shear_m, poiss_r, md, load, np, p);
  std::memcpy(out, s.data(), sizeof(double) * n_pts * 6);
  return 0;
}

// Area integrals of the six shape functions; tau, sfm as in ref_el_geom.
int ref_sf_integral(const double* tau, const double* sfm, double* out) {
  il::StaticArray<std::complex<double>, 3> t;
  for (int k = 0; k < 3; ++k) t[k] = std::complex<double>(tau[2 * k], tau[2 * k + 1]);
  il::StaticArray2D<std::complex<double>, 6, 6> s;
  for (int c = 0; c < 6; ++c)
    for (int r = 0; r < 6; ++r)
      s(r, c) = std::complex<double>(sfm[2 * (r + 6 * c)], sfm[2 * (r + 6 * c) + 1]);
  il::StaticArray<double, 6> v = hfp3d::el_p2_sf_integral(s, t);
  for (int k = 0; k < 6; ++k) out[k] = v[k];
  return 0;
}

// Volume-control matrix, (n_dof+1)^2 col-major into A (n_dof as returned by
// ref_dof_h for the same tip_type).  The reference leaves the corner entry
// (n_dof, n_dof) unset; it is written as 0 here.
int64_t ref_matrix_vc(const double* nods, int64_t nods_rows, int64_t n_nodes,
                      const int64_t* conn, int64_t conn_rows, int64_t n_el,
                      double shear_m, double poiss_r, double beta, int tip_type,
                      int is_dd_local, int nthreads, double* A) {
  set_threads(nthreads);
  hfp3d::Mesh_Geom_T mesh = make_mesh(nods, nods_rows, n_nodes, conn, conn_rows, n_el);
  hfp3d::Num_Param_T np;
  np.beta = beta;
  np.tip_type = tip_type;
  np.is_dd_local = is_dd_local != 0;
  hfp3d::DoF_Handle_T d;
  il::Array2D<double> m = hfp3d::make_3dbem_matrix_vc(shear_m, poiss_r, mesh, np, il::io, d);
  const il::int_t n = m.size(0);
  m(n - 1, n - 1) = 0.0;
  for (il::int_t j = 0; j < n; ++j)
    for (il::int_t i = 0; i < n; ++i) A[i + n * j] = m(i, j);
  return d.n_dof;
}

// Truncated volume-control system: A is (orig_ndof+1)^2 col-major; dof
// handles n_el x 18 col-major; M (used+1)^2 col-major, rhs (used+1).  The
// reference's OpenMP loop over source elements is race-free when no used DoF
// is shared between elements.  *seconds (may be NULL) = duration of the
// reference call alone, without the copies in and out.
int ref_mod_system_vc(const double* A, int64_t orig_ndof, const int64_t* orig_dof_h,
                      const int64_t* dof_h, int64_t n_el, int64_t used_ndof,
                      const double* delta_t, int64_t tsize, double delta_v, int nthreads,
                      double* M, double* rhs, double* seconds) {
  set_threads(nthreads);
  const il::int_t no = orig_ndof + 1;
  il::Array2D<double> a{no, no, 0.0};
  for (il::int_t j = 0; j < no; ++j)
    for (il::int_t i = 0; i < no; ++i) a(i, j) = A[i + no * j];
  hfp3d::DoF_Handle_T oh, uh;
  oh.n_dof = orig_ndof;
  uh.n_dof = used_ndof;
  oh.dof_h = il::Array2D<il::int_t>{n_el, 18, 0};
  uh.dof_h = il::Array2D<il::int_t>{n_el, 18, 0};
  for (int64_t c = 0; c < 18; ++c)
    for (int64_t e = 0; e < n_el; ++e) {
      oh.dof_h(e, c) = orig_dof_h[e + n_el * c];
      uh.dof_h(e, c) = dof_h[e + n_el * c];
    }
  il::Array<double> dt{tsize, 0.0};
  for (il::int_t i = 0; i < tsize; ++i) dt[i] = delta_t[i];
  const auto t0 = std::chrono::steady_clock::now();
  hfp3d::SAE_T sae = hfp3d::mod_3dbem_system_vc(a, oh, uh, dt, delta_v);
  if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  const il::int_t nu = used_ndof + 1;
  for (il::int_t j = 0; j < nu; ++j)
    for (il::int_t i = 0; i < nu; ++i) M[i + nu * j] = sae.matrix(i, j);
  for (il::int_t i = 0; i < nu; ++i) rhs[i] = sae.rhs_v[i];
  return 0;
}

// Cohesion / friction state (layouts as hfpo_match_f_c / hfpo_update_mrd).
static void fc_inputs(int64_t n_nod, const double* dd, int64_t ldd, const int32_t* mat_id,
                      const int32_t* cmid, int64_t n_mat, const double* par, hfp3d::Properties_T& prop,
                      il::Array<int>& mid, il::Array2D<double>& d) {
  prop.n_surface = n_mat;
  prop.c_model_id = il::Array2D<int>{n_mat, 2, 0};
  prop.f_c_param = il::Array<hfp3d::F_C_Param_T>{n_mat};
  for (int64_t m = 0; m < n_mat; ++m) {
    prop.c_model_id(m, 0) = cmid ? cmid[m] : 0;
    prop.c_model_id(m, 1) = cmid ? cmid[m + n_mat] : 0;
    hfp3d::F_C_Param_T f;
    f.cr_open = par[6 * m];
    f.peak_ts = par[6 * m + 1];
    f.cr_slip = par[6 * m + 2];
    f.peak_sc = par[6 * m + 3];
    f.peak_sf = par[6 * m + 4];
    f.res_sf = par[6 * m + 5];
    prop.f_c_param[m] = f;
  }
  mid = il::Array<int>{n_nod, 0};
  d = il::Array2D<double>{n_nod, 3, 0.0};
  for (int64_t n = 0; n < n_nod; ++n) {
    mid[n] = mat_id[n];
    for (int i = 0; i < 3; ++i) d(n, i) = dd[n + ldd * i];
  }
}

static hfp3d::Frac_State_T fc_state(int64_t n_nod, const double* mr_open, const double* mr_slip) {
  hfp3d::Frac_State_T fs;
  fs.mr_open = il::Array<double>{n_nod, 0.0};
  fs.mr_slip = il::Array<double>{n_nod, 0.0};
  fs.friction_coef = il::Array<double>{n_nod, 0.0};
  fs.slip_cohesion = il::Array<double>{n_nod, 0.0};
  fs.open_cohesion = il::Array<double>{n_nod, 0.0};
  for (int64_t n = 0; n < n_nod; ++n) {
    fs.mr_open[n] = mr_open[n];
    fs.mr_slip[n] = mr_slip[n];
  }
  return fs;
}

int ref_match_f_c(int64_t n_nod, const double* dd, int64_t ldd, const int32_t* mat_id, const int32_t* cmid,
                  int64_t n_mat, const double* par, const double* mr_open, const double* mr_slip,
                  double* friction_coef, double* slip_cohesion, double* open_cohesion) {
  set_threads(1);
  hfp3d::Properties_T prop;
  il::Array<int> mid;
  il::Array2D<double> d;
  fc_inputs(n_nod, dd, ldd, mat_id, cmid, n_mat, par, prop, mid, d);
  hfp3d::Frac_State_T fs = fc_state(n_nod, mr_open, mr_slip);
  hfp3d::match_f_c(prop, mid, d, false, il::io, fs);
  for (int64_t n = 0; n < n_nod; ++n) {
    friction_coef[n] = fs.friction_coef[n];
    slip_cohesion[n] = fs.slip_cohesion[n];
    open_cohesion[n] = fs.open_cohesion[n];
  }
  return 0;
}

int ref_update_mrd(int64_t n_nod, const double* dd, int64_t ldd, const int32_t* mat_id, int64_t n_mat,
                   const double* par, double* mr_open, double* mr_slip) {
  set_threads(1);
  hfp3d::Properties_T prop;
  il::Array<int> mid;
  il::Array2D<double> d;
  fc_inputs(n_nod, dd, ldd, mat_id, nullptr, n_mat, par, prop, mid, d);
  hfp3d::Frac_State_T fs = fc_state(n_nod, mr_open, mr_slip);
  hfp3d::update_mrd(prop, mid, d, false, il::io, fs);
  for (int64_t n = 0; n < n_nod; ++n) {
    mr_open[n] = fs.mr_open[n];
    mr_slip[n] = fs.mr_slip[n];
  }
  return 0;
}

}  // extern "C"
