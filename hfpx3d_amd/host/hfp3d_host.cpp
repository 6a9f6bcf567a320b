// hfp3d_host.cpp -- implementation of the C++ host mirror (hfp3d.hpp).
#include <hip/hip_runtime_api.h>

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>

#include "hfp3d.hpp"

namespace hfp3d {

// ------------------------------------------------------------------ basics
void Status::abortOnError() const {
  if (!ok()) {
    std::cerr << "hfp3d error: " << error << " (info " << info << ")" << std::endl;
    std::abort();
  }
}

static int g_device = 0;
static hfpx_ctx* g_ctx = nullptr;

void set_device(int device) { g_device = device; }

hfpx_ctx* context() {
  if (!g_ctx) {
    if (hfpx_create(&g_ctx, g_device, nullptr) != HFPX_OK) {
      std::cerr << "hfp3d: no usable GPU (device " << g_device << ")" << std::endl;
      std::abort();
    }
  }
  return g_ctx;
}

static void check(int rc, const char* what) {
  if (rc != HFPX_OK) {
    std::cerr << "hfp3d: " << what << ": " << hfpx_status_string(rc) << ": " << hfpx_last_error(context())
              << std::endl;
    std::abort();
  }
}

template <typename T>
static T* dalloc(size_t n) {
  void* p = nullptr;
  if (hipMalloc(&p, sizeof(T) * (n ? n : 1)) != hipSuccess) {
    std::cerr << "hfp3d: device allocation of " << sizeof(T) * n << " bytes failed" << std::endl;
    std::abort();
  }
  return (T*)p;
}

static void h2d(void* d, const void* h, size_t bytes) {
  if (hipMemcpy(d, h, bytes, hipMemcpyHostToDevice) != hipSuccess) std::abort();
}
static void d2h(void* h, const void* d, size_t bytes) {
  if (hipMemcpy(h, d, bytes, hipMemcpyDeviceToHost) != hipSuccess) std::abort();
}

DeviceMatrix::DeviceMatrix(int64_t n) : d_(dalloc<double>((size_t)n * n)), n_(n) {}
DeviceMatrix::~DeviceMatrix() {
  if (d_) (void)hipFree(d_);
}
DeviceMatrix& DeviceMatrix::operator=(DeviceMatrix&& o) noexcept {
  if (this != &o) {
    if (d_) (void)hipFree(d_);
    d_ = o.d_;
    n_ = o.n_;
    o.d_ = nullptr;
    o.n_ = 0;
  }
  return *this;
}
Array2D<double> DeviceMatrix::to_host() const {
  Array2D<double> a(n_, n_);
  (void)hipDeviceSynchronize();
  d2h(a.data(), d_, sizeof(double) * n_ * n_);
  return a;
}

// RAII device mesh
struct MeshHandle {
  hfpx_mesh* m = nullptr;
  MeshHandle(const Mesh_Geom_T& mesh, const DoF_Handle_T* dof, double beta, bool dd_local = false) {
    const int64_t* dh = (dof && dof->n_dof > 0 && dof->dof_h.size(0) == mesh.conn.size(1)) ? dof->dof_h.data()
                                                                                            : nullptr;
    check(hfpx_mesh_create(context(), mesh.nods.data(), mesh.nods.size(0), mesh.nods.size(1), mesh.conn.data(),
                           mesh.conn.size(0), mesh.conn.size(1), dh, dh ? dof->n_dof : 0, beta, &m),
          "mesh upload");
    check(hfpx_mesh_set_dd_local(m, dd_local ? 1 : 0), "mesh upload");
    if (!mesh.vert_wts.empty())
      check(hfpx_mesh_set_vertex_weights(context(), m, mesh.vert_wts.data(), (int64_t)mesh.vert_wts.size()),
            "vertex weights");
  }
  ~MeshHandle() { hfpx_mesh_destroy(m); }
};

// ------------------------------------------------------------------ DoF
DoF_Handle_T make_dof_h_crack(const Mesh_Geom_T& mesh, int ap_order, int tip_type) {
  DoF_Handle_T d;
  if (ap_order != 2) return d;
  const int64_t n_el = mesh.conn.size(1);
  const bool ext = mesh.nods.size(0) >= 5 && mesh.conn.size(0) >= 5;
  d.n_dof = 18 * n_el;
  d.dof_h = Array2D<int64_t>(n_el, 18, 0);
  int64_t dec = 0;
  for (int64_t e = 0; e < n_el; ++e) {
    bool tip[3] = {false, false, false};
    if (ext && mesh.conn(3, e) == mesh.conn(4, e))
      for (int v = 0; v < 3; ++v) {
        const int64_t n = mesh.conn(v, e);
        tip[v] = mesh.nods(3, n) == 1.0 && mesh.nods(4, n) == 1.0;
      }
    for (int v = 0; v < 6; ++v) {
      const bool fixed = (tip_type >= 1 && v < 3 && tip[v]) ||
                         (tip_type == 2 && v >= 3 && tip[(v - 2) % 3] && tip[(v - 1) % 3]);
      for (int l = 0; l < 3; ++l) {
        const int ld = 3 * v + l;
        if (fixed) {
          d.dof_h(e, ld) = -1;
          --d.n_dof;
          ++dec;
        } else {
          d.dof_h(e, ld) = 18 * e + ld - dec;
        }
      }
    }
  }
  return d;
}

static void expect(bool cond, const char* what) {
  if (!cond) {
    std::cerr << "hfp3d: " << what << std::endl;
    std::abort();
  }
}

void write_dd_vector_to_md(const std::vector<double>& sol_v, const DoF_Handle_T& dof_h_dd, bool include_p,
                           const DoF_Handle_T& dof_h_pp, Mesh_Data_T& m_data) {
  const int64_t n_el = dof_h_dd.dof_h.size(0);
  const int64_t n_dd = dof_h_dd.n_dof, n_pp = include_p ? dof_h_pp.n_dof : 0;
  expect(dof_h_dd.dof_h.size(1) == 18, "write_dd_vector_to_md: dof_h_dd must be n_el x 18");
  if (m_data.dd.size(0) == 0 || m_data.dd.size(1) == 0)
    m_data.dd = Array2D<double>(6 * n_el, 3, 0.0);
  else
    expect(m_data.dd.size(0) == 6 * n_el && m_data.dd.size(1) == 3, "write_dd_vector_to_md: dd must be 6 n_el x 3");
  if (include_p) {
    expect(dof_h_pp.dof_h.size(0) == n_el && dof_h_pp.dof_h.size(1) == 6,
           "write_dd_vector_to_md: dof_h_pp must be n_el x 6");
    if (m_data.pp.empty()) m_data.pp.assign((size_t)(6 * n_el), 0.0);
    expect((int64_t)m_data.pp.size() == 6 * n_el, "write_dd_vector_to_md: pp must hold 6 n_el values");
  }
  expect((int64_t)sol_v.size() == n_dd + n_pp, "write_dd_vector_to_md: sol_v size != n_dof(dd) + n_dof(pp)");
  for (int64_t e = 0; e < n_el; ++e) {
    bool is_a = false, is_f = false;
    for (int en = 0; en < 6; ++en) {
      for (int i = 0; i < 3; ++i) {
        const int64_t j = dof_h_dd.dof_h(e, 3 * en + i);
        if (j >= 0) {
          expect(j < n_dd, "write_dd_vector_to_md: DD DoF out of range");
          m_data.dd(6 * e + en, i) = sol_v[(size_t)j];
          is_a = true;
        }
      }
      if (include_p) {
        const int64_t j = dof_h_pp.dof_h(e, en);
        if (j >= 0) {
          expect(j < n_pp, "write_dd_vector_to_md: pressure DoF out of range");
          m_data.pp[(size_t)(6 * e + en)] = sol_v[(size_t)(n_dd + j)];
          is_f = true;
        }
      }
    }
    if (is_a) m_data.ae_set.push_back(e);
    if (is_f) m_data.fe_set.push_back(e);
  }
}

void write_dd_vector_to_md(const std::vector<double>& sol_v, const DoF_Handle_T& dof_h_dd, Mesh_Data_T& m_data) {
  m_data.dd = Array2D<double>();
  write_dd_vector_to_md(sol_v, dof_h_dd, false, DoF_Handle_T(), m_data);
}

Array2D<double> make_local_3dbem_submatrix(int kernel_id, double shear_m, double poiss_r, double h,
                                           std::complex<double> z, const std::array<std::complex<double>, 3>& tau,
                                           const Array2D<std::complex<double>>& sfm) {
  expect(sfm.size(0) == 6 && sfm.size(1) == 6, "make_local_3dbem_submatrix: sfm must be 6 x 6");
  double hv[1] = {h}, zv[2] = {z.real(), z.imag()}, tv[6], sv[72];
  for (int k = 0; k < 3; ++k) {
    tv[2 * k] = tau[(size_t)k].real();
    tv[2 * k + 1] = tau[(size_t)k].imag();
  }
  for (int c = 0; c < 6; ++c)
    for (int r = 0; r < 6; ++r) {
      sv[2 * (r + 6 * c)] = sfm(r, c).real();
      sv[2 * (r + 6 * c) + 1] = sfm(r, c).imag();
    }
  double* d = dalloc<double>(1 + 2 + 6 + 72 + 108);
  h2d(d, hv, sizeof hv);
  h2d(d + 1, zv, sizeof zv);
  h2d(d + 3, tv, sizeof tv);
  h2d(d + 9, sv, sizeof sv);
  check(hfpx_local_submatrix(context(), kernel_id, shear_m, poiss_r, 1, d, d + 1, d + 3, d + 9, d + 81),
        "make_local_3dbem_submatrix");
  check(hfpx_synchronize(context()), "make_local_3dbem_submatrix");
  Array2D<double> out(6, 18);
  d2h(out.data(), d + 81, sizeof(double) * 108);
  (void)hipFree(d);
  return out;
}

// ------------------------------------------------------------------ GPU stages
DeviceMatrix make_3dbem_matrix_s(double shear_m, double poiss_r, const Mesh_Geom_T& mesh,
                                 const Num_Param_T& n_par, DoF_Handle_T& dof_hndl) {
  if (dof_hndl.n_dof == 0 || dof_hndl.dof_h.size(0) == 0) dof_hndl = make_dof_h_crack(mesh, 2, n_par.tip_type);
  MeshHandle mh(mesh, &dof_hndl, n_par.beta, n_par.is_dd_local);
  DeviceMatrix A(dof_hndl.n_dof);
  check(hfpx_assemble(context(), mh.m, shear_m, poiss_r, A.data(), dof_hndl.n_dof), "assembly");
  check(hfpx_synchronize(context()), "assembly");
  return A;
}

// ---- cohesion / friction ---------------------------------------------------
static std::vector<hfpx_fc_param> fc_params(const Properties_T& prop) {
  std::vector<hfpx_fc_param> v;
  for (const F_C_Param_T& p : prop.f_c_param)
    v.push_back({p.cr_open, p.peak_ts, p.cr_slip, p.peak_sc, p.peak_sf, p.res_sf});
  return v;
}

// Device copies of the node arrays of one call; freed on scope exit.
struct FcArrays {
  int64_t n;
  double* dd;
  int32_t* mid;
  double* st[5];
  FcArrays(const std::vector<int32_t>& mat_id, const Array2D<double>& dd_h) : n(dd_h.size(0)) {
    if (dd_h.size(1) != 3 || (int64_t)mat_id.size() != n) {
      std::cerr << "hfp3d: cohesion/friction: dd must be n_nod x 3 with one material id per row" << std::endl;
      std::abort();
    }
    dd = dalloc<double>(3 * (size_t)n);
    mid = dalloc<int32_t>((size_t)n);
    for (double*& p : st) p = dalloc<double>((size_t)n);
    h2d(dd, dd_h.data(), sizeof(double) * 3 * n);
    h2d(mid, mat_id.data(), sizeof(int32_t) * n);
  }
  ~FcArrays() {
    (void)hipFree(dd);
    (void)hipFree(mid);
    for (double* p : st) (void)hipFree(p);
  }
  void up(int k, const std::vector<double>& v) {
    if ((int64_t)v.size() != n) {
      std::cerr << "hfp3d: cohesion/friction: state arrays need one entry per DD row" << std::endl;
      std::abort();
    }
    h2d(st[k], v.data(), sizeof(double) * n);
  }
  void down(int k, std::vector<double>& v) const {
    v.resize(n);
    d2h(v.data(), st[k], sizeof(double) * n);
  }
};

void match_f_c(const Properties_T& prop, const std::vector<int32_t>& mat_id, const Array2D<double>& dd,
               bool is_dd_local, Frac_State_T& f_state) {
  (void)is_dd_local;  // the reference rotates nothing either way
  const std::vector<hfpx_fc_param> par = fc_params(prop);
  const int64_t n_mat = (int64_t)par.size();
  if (prop.c_model_id.size(0) != n_mat || prop.c_model_id.size(1) != 2) {
    std::cerr << "hfp3d: match_f_c: c_model_id must be n_surface x 2" << std::endl;
    std::abort();
  }
  FcArrays a(mat_id, dd);
  if (a.n == 0) return;
  a.up(0, f_state.mr_open);
  a.up(1, f_state.mr_slip);
  check(hfpx_match_f_c(context(), a.n, a.dd, a.n, a.mid, prop.c_model_id.data(), par.data(), n_mat, a.st[0],
                       a.st[1], a.st[2], a.st[3], a.st[4]),
        "match_f_c");
  a.down(2, f_state.friction_coef);
  a.down(3, f_state.slip_cohesion);
  a.down(4, f_state.open_cohesion);
}

void update_mrd(const Properties_T& prop, const std::vector<int32_t>& mat_id, const Array2D<double>& dd,
                bool is_dd_local, Frac_State_T& f_state) {
  (void)is_dd_local;
  const std::vector<hfpx_fc_param> par = fc_params(prop);
  FcArrays a(mat_id, dd);
  if (a.n == 0) return;
  a.up(0, f_state.mr_open);
  a.up(1, f_state.mr_slip);
  check(hfpx_update_mrd(context(), a.n, a.dd, a.n, a.mid, par.data(), (int64_t)par.size(), a.st[0], a.st[1]),
        "update_mrd");
  a.down(0, f_state.mr_open);
  a.down(1, f_state.mr_slip);
}

// device mesh of the original handle and the library's iteration state
struct CfState {
  MeshHandle mh;
  hfpx_cf_sys* sys = nullptr;
  const double* A = nullptr;
  CfState(const Mesh_Geom_T& mesh, const DoF_Handle_T& orig_dof_h, const Num_Param_T& n_par, const DeviceMatrix& M)
      : mh(mesh, &orig_dof_h, n_par.beta, n_par.is_dd_local), A(M.data()) {
    check(hfpx_cf_create(context(), mh.m, M.data(), M.size(0), &sys), "vc_cf_iteration: iteration state");
  }
  ~CfState() { hfpx_cf_destroy(sys); }
  CfState(const CfState&) = delete;
  CfState& operator=(const CfState&) = delete;
};

double vc_cf_iteration(const Mesh_Data_T& m_data_p, const SAE_T& orig_vc_sys, const DoF_Handle_T& orig_dof_h,
                       const Properties_T& prop, const Num_Param_T& n_par, const Load_T& load, double new_vol,
                       Mesh_Data_T& m_data) {
  (void)m_data_p;
  const int64_t n_el = m_data.mesh.conn.size(1), n_nod = 6 * n_el;
  auto fail = [](const char* what) {
    std::cerr << "hfp3d: vc_cf_iteration: " << what << std::endl;
    std::abort();
  };
  if (orig_vc_sys.matrix.size(0) != orig_dof_h.n_dof + 1) fail("the matrix is not orig n_dof + 1 square");
  if (m_data.dd.size(0) != n_nod || m_data.dd.size(1) != 3) fail("m_data.dd must be (6 n_el) x 3");
  if ((int64_t)m_data.pp.size() != n_nod || (int64_t)m_data.mat_id.size() != n_nod)
    fail("m_data.pp and m_data.mat_id need one entry per node");
  if (m_data.dof_h_dd.dof_h.size(0) != n_el || m_data.dof_h_dd.dof_h.size(1) != 18)
    fail("m_data.dof_h_dd must be n_el x 18");
  if (load.inj_loc.size(0) < 1 || load.inj_loc.size(1) != 2) fail("load.inj_loc needs a row (element, node)");
  Frac_State_T& fs = m_data.frac_state;
  if (fs.mr_open.empty()) fs.mr_open.assign((size_t)n_nod, 0.0);
  if (fs.mr_slip.empty()) fs.mr_slip.assign((size_t)n_nod, 0.0);
  if ((int64_t)fs.mr_open.size() != n_nod || (int64_t)fs.mr_slip.size() != n_nod)
    fail("the fracture state needs one entry per node");
  const std::vector<hfpx_fc_param> par = fc_params(prop);
  const int64_t n_mat = (int64_t)par.size();
  if (prop.c_model_id.size(0) != n_mat || prop.c_model_id.size(1) != 2) fail("c_model_id must be n_surface x 2");
  if (!orig_vc_sys.cf || orig_vc_sys.cf->A != orig_vc_sys.matrix.data())
    orig_vc_sys.cf = std::make_shared<CfState>(m_data.mesh, orig_dof_h, n_par, orig_vc_sys.matrix);
  CfState& cf = *orig_vc_sys.cf;
  check(hfpx_mesh_set_dd_local(cf.mh.m, n_par.is_dd_local ? 1 : 0), "vc_cf_iteration");
  double* dd = dalloc<double>(3 * (size_t)n_nod);
  double* st[6];  // pp, mr_open, mr_slip, friction, slip cohesion, opening cohesion
  for (double*& p : st) p = dalloc<double>((size_t)n_nod);
  int32_t* mid = dalloc<int32_t>((size_t)n_nod);
  h2d(dd, m_data.dd.data(), sizeof(double) * 3 * n_nod);
  h2d(st[0], m_data.pp.data(), sizeof(double) * n_nod);
  h2d(st[1], fs.mr_open.data(), sizeof(double) * n_nod);
  h2d(st[2], fs.mr_slip.data(), sizeof(double) * n_nod);
  h2d(mid, m_data.mat_id.data(), sizeof(int32_t) * n_nod);
  Array2D<int64_t> dh = m_data.dof_h_dd.dof_h;
  int64_t nd = m_data.dof_h_dd.n_dof;
  hfpx_cf_report rep;
  const int rc = hfpx_vc_cf_iteration(context(), cf.sys, mid, prop.c_model_id.data(), par.data(), n_mat, load.s_inf,
                                      load.inj_loc(0, 0), load.inj_loc(0, 1), new_vol, dd, n_nod, st[0], st[1],
                                      st[2], st[3], st[4], st[5], dh.data(), &nd, 1e-12, 10, &rep);
  check(rc, "vc_cf_iteration");
  d2h(m_data.dd.data(), dd, sizeof(double) * 3 * n_nod);
  d2h(m_data.pp.data(), st[0], sizeof(double) * n_nod);
  fs.friction_coef.resize((size_t)n_nod);
  fs.slip_cohesion.resize((size_t)n_nod);
  fs.open_cohesion.resize((size_t)n_nod);
  d2h(fs.friction_coef.data(), st[3], sizeof(double) * n_nod);
  d2h(fs.slip_cohesion.data(), st[4], sizeof(double) * n_nod);
  d2h(fs.open_cohesion.data(), st[5], sizeof(double) * n_nod);
  (void)hipFree(dd);
  (void)hipFree(mid);
  for (double* p : st) (void)hipFree(p);
  m_data.dof_h_dd.dof_h = std::move(dh);
  m_data.dof_h_dd.n_dof = nd;
  return rep.residual;
}

DeviceMatrix make_3dbem_matrix_vc(double shear_m, double poiss_r, const Mesh_Geom_T& mesh,
                                  const Num_Param_T& n_par, DoF_Handle_T& dof_hndl) {
  if (dof_hndl.n_dof == 0 || dof_hndl.dof_h.size(0) == 0) dof_hndl = make_dof_h_crack(mesh, 2, n_par.tip_type);
  MeshHandle mh(mesh, &dof_hndl, n_par.beta, n_par.is_dd_local);
  const int64_t n1 = dof_hndl.n_dof + 1;
  DeviceMatrix A(n1);
  check(hfpx_assemble_vc(context(), mh.m, shear_m, poiss_r, A.data(), n1, nullptr, 0), "volume-control assembly");
  check(hfpx_synchronize(context()), "volume-control assembly");
  return A;
}

SAE_T mod_3dbem_system_vc(const DeviceMatrix& orig_matrix, const DoF_Handle_T& orig_dof_hndl,
                          const DoF_Handle_T& dof_hndl, const std::vector<double>& delta_t, double delta_v) {
  const int64_t n_el = dof_hndl.dof_h.size(0), ndpe = dof_hndl.dof_h.size(1);
  if (orig_dof_hndl.dof_h.size(0) != n_el || orig_dof_hndl.dof_h.size(1) != ndpe)
    check(HFPX_ERR_INVALID, "mod_3dbem_system_vc: DoF handles of different shapes");
  if (orig_matrix.size(0) != orig_dof_hndl.n_dof + 1)
    check(HFPX_ERR_INVALID, "mod_3dbem_system_vc: matrix is not orig n_dof + 1 square");
  const int64_t nu = dof_hndl.n_dof + 1;
  SAE_T sae;
  sae.n_dof = dof_hndl.n_dof;
  sae.matrix = DeviceMatrix(nu);
  sae.rhs_v.assign((size_t)nu, 0.0);
  double* rhs = dalloc<double>((size_t)nu);
  check(hfpx_vc_truncate(context(), orig_matrix.data(), orig_matrix.size(0), orig_dof_hndl.n_dof,
                         orig_dof_hndl.dof_h.data(), dof_hndl.dof_h.data(), n_el, ndpe, dof_hndl.n_dof,
                         delta_t.data(), (int64_t)delta_t.size(), delta_v, sae.matrix.data(), nu, nullptr, 0, rhs),
        "volume-control truncation");
  d2h(sae.rhs_v.data(), rhs, sizeof(double) * nu);
  (void)hipFree(rhs);
  return sae;
}

void add_s_inf_to_3dbem_rhs(const Mesh_Data_T& mesh_data, const Load_T& load, SAE_T& sae) {
  const int64_t n = mesh_data.dof_h_dd.n_dof;
  if ((int64_t)sae.rhs_v.size() != n) sae.rhs_v.assign((size_t)n, 0.0);
  MeshHandle mh(mesh_data.mesh, &mesh_data.dof_h_dd, 0.125);
  double* b = dalloc<double>((size_t)n);
  h2d(b, sae.rhs_v.data(), sizeof(double) * n);
  check(hfpx_rhs_sinf(context(), mh.m, load.s_inf, 1, b, n), "rhs");
  d2h(sae.rhs_v.data(), b, sizeof(double) * n);
  (void)hipFree(b);
  sae.n_dof = n;
}

Array2D<double> make_3dbem_stress_f_s(double shear_m, double poiss_r, const Mesh_Data_T& m_data,
                                      const Load_T& load, const Num_Param_T& n_par,
                                      const Array2D<double>& m_pts_crd) {
  const int64_t n_el = m_data.mesh.conn.size(1), np = m_pts_crd.size(1);
  Array2D<double> out(np, 6, 0.0);
  if (np == 0) return out;
  MeshHandle mh(m_data.mesh, nullptr, 0.0, n_par.is_dd_local);
  double* dd = dalloc<double>(18 * (size_t)n_el);
  double* pts = dalloc<double>(3 * (size_t)np);
  double* s = dalloc<double>(6 * (size_t)np);
  h2d(dd, m_data.dd.data(), sizeof(double) * 18 * n_el);
  h2d(pts, m_pts_crd.data(), sizeof(double) * 3 * np);
  check(hfpx_stress(context(), mh.m, shear_m, poiss_r, dd, 6 * n_el, pts, np, load.s_inf, s, np), "stress");
  d2h(out.data(), s, sizeof(double) * 6 * np);
  (void)hipFree(dd);
  (void)hipFree(pts);
  (void)hipFree(s);
  return out;
}

LU::LU(DeviceMatrix A, Status& status) : A_(std::move(A)) {
  const int64_t n = A_.size(0);
  ipiv_ = dalloc<int32_t>((size_t)n);
  int64_t info = 0;
  const int rc = hfpx_getrf(context(), HFPX_F64, A_.data(), n, n, ipiv_, &info);
  if (rc == HFPX_ERR_SINGULAR) {
    status.error = "MatrixSingular";
    status.info = info - 1;
  } else {
    check(rc, "LU factorisation");
  }
}
LU::~LU() {
  if (ipiv_) (void)hipFree(ipiv_);
}
Array2D<double> LU::solve(const Array2D<double>& y) const {
  const int64_t n = A_.size(0), nrhs = y.size(1);
  expect(y.size(0) == n, "LU::solve: right-hand sides must have n rows");
  Array2D<double> x = y;
  if (nrhs == 0) return x;
  double* b = dalloc<double>((size_t)n * nrhs);
  h2d(b, x.data(), sizeof(double) * n * nrhs);
  check(hfpx_getrs(context(), HFPX_F64, A_.data(), n, ipiv_, n, b, n, nrhs), "LU solve");
  check(hfpx_synchronize(context()), "LU solve");
  d2h(x.data(), b, sizeof(double) * n * nrhs);
  (void)hipFree(b);
  return x;
}
std::vector<double> LU::solve(const std::vector<double>& y) const {
  const int64_t n = A_.size(0);
  std::vector<double> x(y);
  double* b = dalloc<double>((size_t)n);
  h2d(b, x.data(), sizeof(double) * n);
  check(hfpx_getrs(context(), HFPX_F64, A_.data(), n, ipiv_, n, b, n, 1), "LU solve");
  check(hfpx_synchronize(context()), "LU solve");
  d2h(x.data(), b, sizeof(double) * n);
  (void)hipFree(b);
  return x;
}

// ------------------------------------------------------------------ npy
// Reads a 2-D .npy array (little-endian <i4, <i8 or <f8, either order) into a
// column-major array of T.  Read-only: unlike the reference's il::load, the
// file is opened "rb" (il opens "r+b", which fails on read-only files).
template <typename T>
static bool load_npy(const std::string& path, Array2D<T>& out, std::string& err) {
  std::ifstream f(path, std::ios::binary);
  if (!f) {
    err = "cannot open " + path;
    return false;
  }
  char magic[6];
  f.read(magic, 6);
  if (!f || std::memcmp(magic, "\x93NUMPY", 6) != 0) {
    err = path + " is not an npy file";
    return false;
  }
  unsigned char ver[2];
  f.read((char*)ver, 2);
  uint32_t hlen = 0;
  if (ver[0] == 1) {
    unsigned char b[2];
    f.read((char*)b, 2);
    hlen = b[0] | (b[1] << 8);
  } else {
    unsigned char b[4];
    f.read((char*)b, 4);
    hlen = b[0] | (b[1] << 8) | (b[2] << 16) | ((uint32_t)b[3] << 24);
  }
  if (!f || hlen > (1u << 20)) {
    err = path + ": bad npy header";
    return false;
  }
  std::string hdr(hlen, '\0');
  f.read(&hdr[0], hlen);
  if (!f) {
    err = path + ": truncated npy header";
    return false;
  }
  const std::streamoff data_start = f.tellg();
  f.seekg(0, std::ios::end);
  const std::streamoff data_bytes = f.tellg() - data_start;
  f.seekg(data_start);
  // value of a header key; npos-safe on malformed headers (empty string)
  auto field = [&](const std::string& key) {
    size_t p = hdr.find("'" + key + "'");
    if (p == std::string::npos) return std::string();
    p = hdr.find(':', p);
    if (p == std::string::npos) return std::string();
    size_t q = p + 1;
    while (q < hdr.size() && std::isspace((unsigned char)hdr[q])) ++q;
    if (q >= hdr.size()) return std::string();
    if (hdr[q] == '\'' || hdr[q] == '(') {
      const size_t e = hdr.find(hdr[q] == '(' ? ')' : '\'', q + 1);
      if (e == std::string::npos) return std::string();
      return hdr.substr(q + 1, e - q - 1);
    }
    const size_t e = hdr.find_first_of(",}", q);
    return e == std::string::npos ? std::string() : hdr.substr(q, e - q);
  };
  const std::string descr = field("descr"), fo = field("fortran_order"), shape = field("shape");
  const bool fortran = fo.find("True") != std::string::npos;
  std::vector<int64_t> dims;
  {
    std::stringstream ss(shape);
    std::string tok;
    while (std::getline(ss, tok, ',')) {
      const size_t a = tok.find_first_not_of(" \t");
      if (a == std::string::npos) continue;
      const size_t b = tok.find_last_not_of(" \t");
      const std::string t = tok.substr(a, b - a + 1);
      if (t.empty() || t.size() > 12 || t.find_first_not_of("0123456789") != std::string::npos) {
        err = path + ": bad shape (" + shape + ")";
        return false;
      }
      dims.push_back(std::strtoll(t.c_str(), nullptr, 10));
    }
  }
  if (dims.size() == 1) dims.push_back(1);
  if (dims.size() != 2) {
    err = path + ": only 1-D and 2-D arrays are supported";
    return false;
  }
  const int64_t n0 = dims[0], n1 = dims[1];
  const int64_t esz = descr == "<i4" ? 4 : 8;
  if (n1 > 0 && n0 > data_bytes / esz / n1) {
    err = path + ": shape (" + shape + ") larger than the file's data";
    return false;
  }
  const int64_t cnt = n0 * n1;
  out.resize(n0, n1);
  auto fill = [&](auto dummy) {
    using S = decltype(dummy);
    std::vector<S> buf((size_t)cnt);
    f.read((char*)buf.data(), sizeof(S) * cnt);
    if (!f) return false;
    for (int64_t i = 0; i < n0; ++i)
      for (int64_t j = 0; j < n1; ++j) out(i, j) = (T)buf[(size_t)(fortran ? i + n0 * j : i * n1 + j)];
    return true;
  };
  bool ok;
  if (descr == "<f8") ok = fill(double());
  else if (descr == "<i4") ok = fill(int32_t());
  else if (descr == "<i8") ok = fill(int64_t());
  else {
    err = path + ": unsupported dtype " + descr;
    return false;
  }
  if (!ok) err = path + ": truncated data";
  return ok;
}

static std::string join(const std::string& dir, const std::string& f) {
  if (dir.empty()) return f;
  return dir.back() == '/' ? dir + f : dir + "/" + f;
}

static void shift_origin(Mesh_Geom_T& mesh, int64_t origin) {
  if (origin == 0) return;
  const int64_t nr = mesh.conn.size(0) >= 3 ? 3 : mesh.conn.size(0);
  for (int64_t k = 0; k < mesh.conn.size(1); ++k)
    for (int64_t j = 0; j < nr; ++j) mesh.conn(j, k) -= origin;
}

static void load_npy_mesh(const std::string& src_dir, const std::string& cf, const std::string& nf, int64_t origin,
                          Mesh_Geom_T& mesh, Status& status) {
  std::string err;
  if (!load_npy(join(src_dir, cf), mesh.conn, err) || !load_npy(join(src_dir, nf), mesh.nods, err)) {
    status.error = err;
    return;
  }
  shift_origin(mesh, origin);
}

}  // namespace hfp3d

namespace hfpx {
int set_error(hfpx_ctx* ctx, int code, const char* fmt, ...);
}

// npy files straight to a device mesh (C ABI): the reference's npy mesh
// loader, index-origin shift of the vertex rows, DoF handle and upload.
extern "C" int hfpx_mesh_create_npy(hfpx_ctx* ctx, const char* nodes_npy, const char* elems_npy,
                                    int64_t index_origin, int tip_type, double beta, hfpx_mesh** mesh) {
  if (!ctx || !nodes_npy || !elems_npy || !mesh) return HFPX_ERR_INVALID;
  *mesh = nullptr;
  if (tip_type < 0 || tip_type > 2) return hfpx::set_error(ctx, HFPX_ERR_INVALID, "tip_type %d not in 0..2", tip_type);
  try {
    hfp3d::Mesh_Geom_T m;
    hfp3d::Status st;
    hfp3d::load_npy_mesh_files(elems_npy, nodes_npy, index_origin, m, st);
    if (!st.ok()) return hfpx::set_error(ctx, HFPX_ERR_INVALID, "%s", st.error.c_str());
    if (m.nods.size(0) < 3 || m.conn.size(0) < 3 || m.conn.size(1) < 1 || m.nods.size(1) < 3)
      return hfpx::set_error(ctx, HFPX_ERR_INVALID,
                             "mesh needs >= 3 coordinate rows, >= 3 nodes, >= 3 connectivity rows, >= 1 element");
    // vertex ids in range before the DoF handle reads the nodes they name
    const int64_t n_nodes = m.nods.size(1);
    for (int64_t e = 0; e < m.conn.size(1); ++e)
      for (int k = 0; k < 3; ++k)
        if (m.conn(k, e) < 0 || m.conn(k, e) >= n_nodes)
          return hfpx::set_error(ctx, HFPX_ERR_INVALID, "element %lld references node %lld out of range",
                                 (long long)e, (long long)m.conn(k, e));
    const hfp3d::DoF_Handle_T d = hfp3d::make_dof_h_crack(m, 2, tip_type);
    return hfpx_mesh_create(ctx, m.nods.data(), m.nods.size(0), m.nods.size(1), m.conn.data(), m.conn.size(0),
                            m.conn.size(1), d.dof_h.data(), d.n_dof, beta, mesh);
  } catch (const std::exception& ex) {
    return hfpx::set_error(ctx, HFPX_ERR_INVALID, "npy mesh: %s", ex.what());
  } catch (...) {
    return hfpx::set_error(ctx, HFPX_ERR_INVALID, "npy mesh: unreadable input");
  }
}

namespace hfp3d {

void load_npy_mesh_files(const std::string& conn_path, const std::string& nods_path, int64_t origin,
                         Mesh_Geom_T& mesh, Status& status) {
  load_npy_mesh("", conn_path, nods_path, origin, mesh, status);
}

void load_mesh_from_numpy_32(const std::string& d, const std::string& cf, const std::string& nf, int64_t origin,
                             Mesh_Geom_T& mesh, Status& status) {
  load_npy_mesh(d, cf, nf, origin, mesh, status);
}
void load_mesh_from_numpy_64(const std::string& d, const std::string& cf, const std::string& nf, int64_t origin,
                             Mesh_Geom_T& mesh, Status& status) {
  load_npy_mesh(d, cf, nf, origin, mesh, status);
}

// CSV: one row per node / element / point; the array is stored transposed
// (rows of the file become columns), as in mesh_file_io.cpp:56-116.
template <typename T>
static bool load_csv(const std::string& path, Array2D<T>& out, std::string& err) {
  std::ifstream f(path);
  if (!f) {
    err = "cannot open " + path;
    return false;
  }
  std::vector<std::vector<T>> rows;
  std::string line;
  size_t ncol = 0;
  while (std::getline(f, line)) {
    if (line.find_first_not_of(" \t\r") == std::string::npos) continue;
    std::vector<T> r;
    std::stringstream ls(line);
    std::string cell;
    while (std::getline(ls, cell, ',')) {
      if (cell.find_first_not_of(" \t\r") == std::string::npos) continue;
      std::stringstream cs(cell);
      double v = 0;
      cs >> v;
      r.push_back((T)v);
    }
    if (ncol && r.size() != ncol) {
      err = path + ": row size is not constant";
      return false;
    }
    ncol = r.size();
    rows.push_back(r);
  }
  out.resize((int64_t)ncol, (int64_t)rows.size());
  for (size_t k = 0; k < rows.size(); ++k)
    for (size_t j = 0; j < ncol; ++j) out((int64_t)j, (int64_t)k) = rows[k][j];
  return true;
}

void load_mesh_from_csv(const std::string& d, const std::string& cf, const std::string& nf, int64_t origin,
                        Mesh_Geom_T& mesh, Status& status) {
  std::string err;
  if (!load_csv(join(d, cf), mesh.conn, err) || !load_csv(join(d, nf), mesh.nods, err)) {
    status.error = err;
    return;
  }
  shift_origin(mesh, origin);
}

Array2D<double> load_crd_from_csv(const std::string& d, const std::string& f, Status& status) {
  Array2D<double> a;
  std::string err;
  if (!load_csv(join(d, f), a, err)) status.error = err;
  return a;
}

Array2D<double> load_crd_from_numpy(const std::string& d, const std::string& f, Status& status) {
  Array2D<double> a;
  std::string err;
  if (!load_npy(join(d, f), a, err)) status.error = err;
  return a;
}

bool save_data_to_csv(const Array2D<double>& a, const std::string& trg_dir, const std::string& of_name) {
  std::ofstream o(join(trg_dir, of_name));
  if (!o) return false;
  for (int64_t j = 0; j < a.size(0); ++j) {
    for (int64_t k = 0; k < a.size(1); ++k) {
      o << a(j, k);  // default stream precision (6 significant digits), like the reference
      if (k < a.size(1) - 1) o << ",";
    }
    o << "\n";
  }
  return (bool)o;
}

bool save_data_to_npy(const Array2D<double>& a, const std::string& trg_dir, const std::string& of_name, bool as_f32) {
  std::ofstream o(join(trg_dir, of_name), std::ios::binary);
  if (!o) return false;
  std::string dict = std::string("{'descr': '") + (as_f32 ? "<f4" : "<f8") + "', 'fortran_order': True, 'shape': (" +
                     std::to_string(a.size(0)) + ", " + std::to_string(a.size(1)) + "), }";
  // magic (6) + version (2) + header length (2) + dict + padding + '\n' is a multiple of 64
  const size_t base = 10 + dict.size() + 1;
  dict.append((64 - base % 64) % 64, ' ');
  dict.push_back('\n');
  const uint16_t hl = (uint16_t)dict.size();
  o.write("\x93NUMPY\x01\x00", 8);
  const char hb[2] = {(char)(hl & 0xff), (char)(hl >> 8)};
  o.write(hb, 2);
  o.write(dict.data(), (std::streamsize)dict.size());
  const size_t n = (size_t)(a.size(0) * a.size(1));
  if (as_f32) {
    std::vector<float> f(n);
    for (size_t i = 0; i < n; ++i) f[i] = (float)a.data()[i];
    o.write((const char*)f.data(), (std::streamsize)(n * sizeof(float)));
  } else {
    o.write((const char*)a.data(), (std::streamsize)(n * sizeof(double)));
  }
  return (bool)o;
}

// ------------------------------------------------------------------ config
// A TOML subset: key = value pairs (strings, integers, floats, booleans),
// '#' comments and [table] headers.  Values keep their type: like the
// reference's reader, a key read as a float is ignored when written as an
// integer literal and vice versa (config_file_io.cpp:147-275).
namespace {
struct TVal {
  enum Kind { Str, Int, Dbl, Bool } kind = Str;
  std::string s;
  int64_t i = 0;
  double d = 0;
  bool b = false;
};
using Table = std::map<std::string, TVal>;

std::string trim(const std::string& x) {
  size_t a = x.find_first_not_of(" \t\r"), b = x.find_last_not_of(" \t\r");
  return a == std::string::npos ? std::string() : x.substr(a, b - a + 1);
}

bool parse_toml(const std::string& path, std::map<std::string, Table>& out, std::string& err) {
  std::ifstream f(path);
  if (!f) {
    err = "cannot open config file " + path;
    return false;
  }
  std::string line, table;
  int ln = 0;
  while (std::getline(f, line)) {
    ++ln;
    // strip comments outside strings
    bool in_s = false;
    for (size_t k = 0; k < line.size(); ++k) {
      if (line[k] == '"') in_s = !in_s;
      if (line[k] == '#' && !in_s) {
        line = line.substr(0, k);
        break;
      }
    }
    line = trim(line);
    if (line.empty()) continue;
    if (line.front() == '[') {
      table = trim(line.substr(1, line.find(']') - 1));
      continue;
    }
    size_t eq = line.find('=');
    if (eq == std::string::npos) {
      err = path + ":" + std::to_string(ln) + ": expected key = value";
      return false;
    }
    std::string key = trim(line.substr(0, eq)), v = trim(line.substr(eq + 1));
    TVal t;
    if (!v.empty() && v.front() == '"') {
      t.kind = TVal::Str;
      t.s = v.substr(1, v.rfind('"') - 1);
    } else if (v == "true" || v == "false") {
      t.kind = TVal::Bool;
      t.b = v == "true";
    } else {
      bool is_int = !v.empty();
      for (size_t k = 0; k < v.size(); ++k)
        if (!(std::isdigit((unsigned char)v[k]) || ((v[k] == '-' || v[k] == '+') && k == 0) || v[k] == '_'))
          is_int = false;
      char* end = nullptr;
      if (is_int) {
        std::string w;
        for (char c : v)
          if (c != '_') w += c;
        t.kind = TVal::Int;
        t.i = std::strtoll(w.c_str(), &end, 10);
      } else {
        t.kind = TVal::Dbl;
        t.d = std::strtod(v.c_str(), &end);
        if (!end || *end != '\0') {
          err = path + ":" + std::to_string(ln) + ": cannot parse value of " + key;
          return false;
        }
      }
    }
    out[table][key] = t;
  }
  return true;
}

const TVal* get(const Table& t, const std::string& k, TVal::Kind kind) {
  auto it = t.find(k);
  if (it == t.end() || it->second.kind != kind) return nullptr;
  return &it->second;
}

std::string dir_of(const std::string& base, const std::string& d) {
  std::string r = (!d.empty() && d.front() == '/') ? d : join(base, d);
  if (r.empty() || r.back() != '/') r += '/';
  return r;
}
}  // namespace

IO_param_T set_default_paths(const std::string& src_path) {
  IO_param_T p;
  p.src_dir = src_path.empty() ? std::string(".") : src_path;
  p.cf_name = p.src_dir + "/config.toml";
  p.input_dir = p.src_dir + "/Mesh_Files/";
  p.output_dir = p.src_dir + "/Test_Output/";
  p.matr_f_name = "test_assembly";
  p.out_f_name = "test_solution";
  p.strs_f_name = "test_stresses";
  return p;
}

void read_config(const std::string& f_path, Mesh_Data_T& mesh_data, Properties_T& props, Load_T& load,
                 Num_Param_T& num_param, Sim_Param_T& sim_param, IO_param_T& io_param, Status& status) {
  std::string p = f_path;
  for (auto& c : p)
    if (c == '\\') c = '/';
  const size_t slash = p.rfind('/');
  const std::string src_dir = slash == std::string::npos ? std::string(".") : p.substr(0, slash);
  io_param = set_default_paths(src_dir);
  std::map<std::string, Table> doc;
  std::string err;
  if (!parse_toml(p, doc, err)) {
    status.error = err;
    return;
  }
  const Table& c = doc[""];
  if (auto v = get(c, "mesh_input_directory", TVal::Str)) io_param.input_dir = dir_of(src_dir, v->s);
  if (auto v = get(c, "mesh_conn_fname", TVal::Str)) io_param.conn_f_name = v->s;
  else {
    status.error = "config: mesh_conn_fname missing";
    return;
  }
  if (auto v = get(c, "nodes_crd_fname", TVal::Str)) io_param.node_f_name = v->s;
  else {
    status.error = "config: nodes_crd_fname missing";
    return;
  }
  io_param.in_f_format = "npy32";
  if (auto v = get(c, "mesh_input_format", TVal::Str)) io_param.in_f_format = v->s;
  int64_t origin = 0;
  if (auto v = get(c, "array_origin", TVal::Int)) origin = v->i;
  sim_param = Sim_Param_T();
  if (auto v = get(c, "do_save_matrix", TVal::Bool)) sim_param.do_save_matrix = v->b;
  if (auto v = get(c, "do_save_solution", TVal::Bool)) sim_param.do_save_solution = v->b;
  if (auto v = get(c, "do_postprocess", TVal::Bool)) sim_param.do_postprocess = v->b;
  // material (config_file_io.cpp:171-219): optional [solid] table first
  props = Properties_T();
  if (auto v = get(c, "number_solids", TVal::Int)) props.n_solid = v->i ? v->i : 1;
  props.shear_m.assign((size_t)props.n_solid, 1.0);
  props.poiss_r.assign((size_t)props.n_solid, 0.0);
  const bool solid_table = doc.count("solid") > 0;
  const Table& mat = solid_table ? doc["solid"] : c;
  if (auto v = get(mat, "Poisson_ratio", TVal::Dbl)) props.poiss_r[0] = v->d;
  // in [solid] a non-float Shear_modulus falls back to Young_modulus; at the
  // top level any Shear_modulus entry suppresses Young_modulus
  const TVal* g = get(mat, "Shear_modulus", TVal::Dbl);
  if (g) {
    props.shear_m[0] = g->d;
  } else if (solid_table || !mat.count("Shear_modulus")) {
    if (auto v = get(mat, "Young_modulus", TVal::Dbl)) props.shear_m[0] = v->d / 2.0 / (1.0 + props.poiss_r[0]);
  }
  num_param = Num_Param_T();
  if (auto v = get(c, "cp_offset", TVal::Dbl)) num_param.beta = v->d;
  if (auto v = get(c, "fix_tip", TVal::Int)) num_param.tip_type = (int)v->i;
  if (auto v = get(c, "is_dd_local", TVal::Bool)) num_param.is_dd_local = v->b;
  // far-field stress: compression-positive input, normal components flipped
  const char* keys[6] = {"S_xx", "S_yy", "S_zz", "S_xy", "S_xz", "S_yz"};
  for (int k = 0; k < 6; ++k) {
    load.s_inf[k] = 0.0;
    if (auto v = get(c, keys[k], TVal::Dbl)) load.s_inf[k] = k < 3 ? -v->d : v->d;
  }
  // mesh
  if (io_param.in_f_format == "csv")
    load_mesh_from_csv(io_param.input_dir, io_param.conn_f_name, io_param.node_f_name, origin, mesh_data.mesh, status);
  else if (io_param.in_f_format == "npy64")
    load_mesh_from_numpy_64(io_param.input_dir, io_param.conn_f_name, io_param.node_f_name, origin, mesh_data.mesh,
                            status);
  else
    load_mesh_from_numpy_32(io_param.input_dir, io_param.conn_f_name, io_param.node_f_name, origin, mesh_data.mesh,
                            status);
  if (!status.ok()) return;
  // observation points and outputs
  io_param.obs_p_dir = io_param.input_dir;
  if (auto v = get(c, "observ_crd_directory", TVal::Str)) io_param.obs_p_dir = dir_of(src_dir, v->s);
  if (auto v = get(c, "observ_crd_fname", TVal::Str)) io_param.obs_p_f_name = v->s;
  else if (sim_param.do_postprocess) {
    status.error = "config: observ_crd_fname missing";
    return;
  }
  io_param.obs_p_f_format = "npy32";
  if (auto v = get(c, "observ_input_format", TVal::Str)) io_param.obs_p_f_format = v->s;
  if (auto v = get(c, "output_directory", TVal::Str)) io_param.output_dir = dir_of(src_dir, v->s);
  if (auto v = get(c, "output_signature", TVal::Str)) {
    io_param.matr_f_name += v->s;
    io_param.out_f_name += v->s;
    io_param.strs_f_name += v->s;
  }
  io_param.out_f_format = "csv";
  if (auto v = get(c, "output_format", TVal::Str)) io_param.out_f_format = v->s;
  io_param.matr_f_name += ".csv";
  io_param.out_f_name += ".csv";
  io_param.strs_f_name += ".csv";
}

}  // namespace hfp3d
