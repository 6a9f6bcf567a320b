// hfp3d.hpp -- C++ host mirror of the HFPx3D hot-path call surface, backed by
// the HIP library through its C ABI (include/hfpx3d.h).
//
// Same names, argument meaning and stage semantics as the reference's
// namespace hfp3d:
//   Mesh_Geom_T, DoF_Handle_T, Mesh_Data_T      src/Core/surface_mesh_utilities.h:25-104
//   Properties_T, Load_T, Num_Param_T, Sim_Param_T  src/Core/model_parameters.h:34-104
//   SAE_T                                       src/Solvers/system_assembly.h:26-30
//   IO_param_T, read_config                     src/IO/config_file_io.h:22-53
//   make_dof_h_crack, write_dd_vector_to_md     src/Core/surface_mesh_utilities.cpp:28-112, 256-328
//   make_3dbem_matrix_s, add_s_inf_to_3dbem_rhs,
//   make_3dbem_stress_f_s                       src/Solvers/system_assembly.cpp:30-455
//   make_3dbem_matrix_vc, mod_3dbem_system_vc   src/Solvers/system_assembly.cpp:458-629
//   make_local_3dbem_submatrix                  src/Elasticity/elasticity_kernel_integration.h:25-29
//   load_mesh_from_numpy_32/_64, load_mesh_from_csv,
//   load_crd_from_csv/_numpy                    src/IO/mesh_file_io.cpp:28-471
//   save_data_to_csv                            src/IO/data_file_io.h:171-204
//   LU (factor + solve)                         il/linear_algebra/dense/factorization/LU.h:46-138
//
// Differences, by design: the dense matrix lives in device memory
// (DeviceMatrix, moved into LU like il::LU takes its matrix by value);
// failures that the reference handles with abort() are reported through
// Status instead.
#pragma once
#include <array>
#include <complex>
#include <cstdint>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "../../include/hfpx3d.h"

namespace hfp3d {

// Column-major 2D array (ld = size(0)), the layout of il::Array2D.
template <typename T>
class Array2D {
 public:
  Array2D() = default;
  Array2D(int64_t n0, int64_t n1, T v = T()) : n0_(n0), n1_(n1), d_((size_t)(n0 * n1), v) {}
  int64_t size(int k) const { return k == 0 ? n0_ : n1_; }
  T& operator()(int64_t i, int64_t j) { return d_[(size_t)(i + n0_ * j)]; }
  const T& operator()(int64_t i, int64_t j) const { return d_[(size_t)(i + n0_ * j)]; }
  T* data() { return d_.data(); }
  const T* data() const { return d_.data(); }
  void resize(int64_t n0, int64_t n1) {
    n0_ = n0;
    n1_ = n1;
    d_.assign((size_t)(n0 * n1), T());
  }

 private:
  int64_t n0_ = 0, n1_ = 0;
  std::vector<T> d_;
};

struct Status {
  std::string error;  // empty: ok
  int64_t info = 0;   // e.g. rank for MatrixSingular (LU.h:94-98)
  bool ok() const { return error.empty(); }
  void abortOnError() const;
};

struct Mesh_Geom_T {
  Array2D<double> nods;   // (3 or 6) x n_nodes
  Array2D<int64_t> conn;  // (3 or 5) x n_el, 0-based vertex ids in rows 0..2
  // Optional per-node vertex weights (n_nodes, > 0) of a non-uniform edge
  // partition (make_el_sfm_nonuniform / el_cp_nonuniform,
  // element_utilities.cpp:240-322, 397-431); empty = uniform element.
  std::vector<double> vert_wts;
};

struct DoF_Handle_T {
  int64_t n_dof = 0;
  Array2D<int64_t> dof_h;  // n_el x 18, -1 = fixed
};

// Node-wise crack-surface state (src/Core/surface_mesh_utilities.h:51-62).
struct Frac_State_T {
  std::vector<double> mr_open;        // largest relative opening reached
  std::vector<double> mr_slip;        // largest relative slip reached
  std::vector<double> friction_coef;
  std::vector<double> slip_cohesion;
  std::vector<double> open_cohesion;
};

struct Mesh_Data_T {
  Mesh_Geom_T mesh;
  DoF_Handle_T dof_h_dd;
  DoF_Handle_T dof_h_pp;   // n_el x 6 pressure DoF (one per node), -1 = none
  Array2D<double> dd;      // (6 n_el) x 3, global components
  std::vector<double> pp;  // fluid pressure per node (6 n_el)
  std::vector<int64_t> ae_set;  // elements with a DD DoF ("active")
  std::vector<int64_t> fe_set;  // elements with a pressure DoF ("filled")
  std::vector<int32_t> mat_id;  // surface material id per node (6 n_el)
  Frac_State_T frac_state;      // node-wise cohesion / friction state
};

// Friction / cohesion parameters of one surface material
// (src/Core/model_parameters.h:22-30).
struct F_C_Param_T {
  double cr_open = 1.0;  // critical opening
  double peak_ts = 0.0;  // peak tensile stress
  double cr_slip = 1.0;  // critical slip
  double peak_sc = 0.0;  // peak shear cohesion
  double peak_sf = 0.0;  // peak friction coefficient
  double res_sf = 0.0;   // residual friction coefficient
};

struct Properties_T {
  int64_t n_solid = 1;
  std::vector<double> shear_m{1.0};
  std::vector<double> poiss_r{0.0};
  int64_t n_surface = 1;
  Array2D<int32_t> c_model_id;          // n_surface x 2: cohesion model, friction model
  std::vector<F_C_Param_T> f_c_param;   // per surface material
};

struct Load_T {
  double s_inf[6] = {0, 0, 0, 0, 0, 0};  // xx yy zz xy xz yz, tension positive
  // injection locations, n x 2 (element, local node); row 0 is the injection
  // point of the volume-control iteration (model_parameters.h:64)
  Array2D<int64_t> inj_loc;
};

struct Num_Param_T {
  double beta = 0.125;
  int tip_type = 1;
  bool is_dd_local = false;
};

struct Sim_Param_T {
  bool do_save_matrix = false;
  bool do_save_solution = false;
  bool do_postprocess = false;
};

struct IO_param_T {
  std::string src_dir, cf_name, input_dir, conn_f_name, node_f_name, in_f_format;
  std::string obs_p_dir, obs_p_f_name, obs_p_f_format;
  std::string output_dir, out_f_name, out_f_format, matr_f_name, strs_f_name;
};

// One GPU context per process (device 0 unless set_device is called first).
hfpx_ctx* context();
void set_device(int device);

// Dense n x n FP64 matrix in device memory (move-only).
class DeviceMatrix {
 public:
  DeviceMatrix() = default;
  explicit DeviceMatrix(int64_t n);
  ~DeviceMatrix();
  DeviceMatrix(DeviceMatrix&& o) noexcept { *this = std::move(o); }
  DeviceMatrix& operator=(DeviceMatrix&& o) noexcept;
  DeviceMatrix(const DeviceMatrix&) = delete;
  DeviceMatrix& operator=(const DeviceMatrix&) = delete;
  int64_t size(int) const { return n_; }
  double* data() { return d_; }
  const double* data() const { return d_; }
  Array2D<double> to_host() const;

 private:
  double* d_ = nullptr;
  int64_t n_ = 0;
};

// Device state of the cohesion / friction iteration on a resident
// volume-control matrix (the device mesh of the original handle and the
// cached FP32 factors of the truncated system).
struct CfState;

struct SAE_T {
  int64_t n_dof = 0;
  DeviceMatrix matrix;
  std::vector<double> rhs_v;
  // set by vc_cf_iteration on the volume-control system it iterates on
  mutable std::shared_ptr<CfState> cf;
};

// ---- mesh / DoF -----------------------------------------------------------
DoF_Handle_T make_dof_h_crack(const Mesh_Geom_T& mesh, int ap_order, int tip_type);
// Solution vector -> nodal DD (m_data.dd, (6 n_el) x 3; fixed DoF keep their
// value, 0 when dd is created here) and, with include_p, nodal pressures
// m_data.pp[6 el + node] = sol_v[dof_h_dd.n_dof + dof_h_pp(el, node)];
// m_data.ae_set / fe_set get the elements with a DD / pressure DoF.  As
// write_dd_vector_to_md (src/Core/surface_mesh_utilities.cpp:256-328), except
// that pp holds one value per node (the reference sizes it per element and
// then indexes it per node).  Size mismatches abort, like IL_EXPECT_FAST.
void write_dd_vector_to_md(const std::vector<double>& sol_v, const DoF_Handle_T& dof_h_dd, bool include_p,
                           const DoF_Handle_T& dof_h_pp, Mesh_Data_T& m_data);
// The DD-only form (include_p = false).
void write_dd_vector_to_md(const std::vector<double>& sol_v, const DoF_Handle_T& dof_h_dd, Mesh_Data_T& m_data);

// ---- the element integral ---------------------------------------------------
// 6 x 18 stress influence (rows S11 S22 S33 S12 S13 S23 in the element
// frame, column 3 node + DD component) at the point (h, z) of the element
// tau / sfm (6 x 6 complex: row = node, column = monomial 1, tau, conj(tau),
// tau^2, conj(tau)^2, tau conj(tau)), on the device.  kernel_id 1 (the
// hypersingular kernel) only.  As hfp3d::make_local_3dbem_submatrix
// (src/Elasticity/elasticity_kernel_integration.cpp:34-205).
Array2D<double> make_local_3dbem_submatrix(int kernel_id, double shear_m, double poiss_r, double h,
                                           std::complex<double> z, const std::array<std::complex<double>, 3>& tau,
                                           const Array2D<std::complex<double>>& sfm);

// ---- the GPU stages --------------------------------------------------------
DeviceMatrix make_3dbem_matrix_s(double shear_m, double poiss_r, const Mesh_Geom_T& mesh,
                                 const Num_Param_T& n_par, DoF_Handle_T& dof_hndl);
// Volume control: (n_dof+1)^2 bordered matrix (corner 0) and the truncated
// system gathered from it on the device (rhs_v copied to the host).
DeviceMatrix make_3dbem_matrix_vc(double shear_m, double poiss_r, const Mesh_Geom_T& mesh,
                                  const Num_Param_T& n_par, DoF_Handle_T& dof_hndl);
SAE_T mod_3dbem_system_vc(const DeviceMatrix& orig_matrix, const DoF_Handle_T& orig_dof_hndl,
                          const DoF_Handle_T& dof_hndl, const std::vector<double>& delta_t, double delta_v);
void add_s_inf_to_3dbem_rhs(const Mesh_Data_T& mesh_data, const Load_T& load, SAE_T& sae);
Array2D<double> make_3dbem_stress_f_s(double shear_m, double poiss_r, const Mesh_Data_T& m_data,
                                      const Load_T& load, const Num_Param_T& n_par,
                                      const Array2D<double>& m_pts_crd);

// Cohesion / friction laws, node by node, on the device
// (src/Development/cohesion_friction.cpp:22-189).  dd: n_nod x 3 (shear 0, 1,
// opening 2, used as given, as in the reference); mat_id: n_nod.
void match_f_c(const Properties_T& prop, const std::vector<int32_t>& mat_id, const Array2D<double>& dd,
               bool is_dd_local, Frac_State_T& f_state);
void update_mrd(const Properties_T& prop, const std::vector<int32_t>& mat_id, const Array2D<double>& dd,
                bool is_dd_local, Frac_State_T& f_state);

// One step of the volume-control scheme with cohesion and friction on a
// pre-existing mesh (src/Development/c_f_iteration.cpp:27-436): returns the
// residual |DD and pressure increments|_1 + |traction adjustments|_1 and
// updates m_data.dd, m_data.pp (collocation points that opened take the new
// pressure), m_data.dof_h_dd (the new active handle) and m_data.frac_state
// (friction / cohesion of the entry DD).  orig_vc_sys.matrix is the
// volume-control matrix of orig_dof_h (make_3dbem_matrix_vc); the device
// state is kept on orig_vc_sys between steps, and the FP32 factors of the
// truncated system are reused while the active set stays the same.
// m_data_p is unused, as in the reference.  The traction adjustment enters
// the system as -R^T dt in all three components (the reference adds +R^T dt
// in two), and the opening of a collocation point is taken in the element
// frame (the reference rotates it twice).
double vc_cf_iteration(const Mesh_Data_T& m_data_p, const SAE_T& orig_vc_sys, const DoF_Handle_T& orig_dof_h,
                       const Properties_T& prop, const Num_Param_T& n_par, const Load_T& load, double new_vol,
                       Mesh_Data_T& m_data);

class LU {
 public:
  LU(DeviceMatrix A, Status& status);
  ~LU();
  LU(const LU&) = delete;
  LU& operator=(const LU&) = delete;
  std::vector<double> solve(const std::vector<double>& y) const;
  // Many right-hand sides at once (n x nrhs, column-major), as
  // il::LU::solve(il::Array2D<double>) (LU.h:122-138, dgetrs with nrhs).
  Array2D<double> solve(const Array2D<double>& y) const;

 private:
  DeviceMatrix A_;
  int32_t* ipiv_ = nullptr;
};

// ---- IO ---------------------------------------------------------------------
void load_mesh_from_numpy_32(const std::string& src_dir, const std::string& conn_f_name,
                             const std::string& node_f_name, int64_t origin, Mesh_Geom_T& mesh,
                             Status& status);
void load_mesh_from_numpy_64(const std::string& src_dir, const std::string& conn_f_name,
                             const std::string& node_f_name, int64_t origin, Mesh_Geom_T& mesh,
                             Status& status);
// Same from two file paths (no directory join).
void load_npy_mesh_files(const std::string& conn_path, const std::string& nods_path, int64_t origin,
                         Mesh_Geom_T& mesh, Status& status);
void load_mesh_from_csv(const std::string& src_dir, const std::string& conn_f_name,
                        const std::string& node_f_name, int64_t origin, Mesh_Geom_T& mesh, Status& status);
Array2D<double> load_crd_from_csv(const std::string& src_dir, const std::string& f_name, Status& status);
Array2D<double> load_crd_from_numpy(const std::string& src_dir, const std::string& f_name, Status& status);
bool save_data_to_csv(const Array2D<double>& a, const std::string& trg_dir, const std::string& of_name);
// Binary output for output_format "npy64" / "npy32" (the reference's branches
// for these formats are empty, main.cpp:132-152, 215-230, 270-285): NumPy
// .npy v1.0, '<f8' or '<f4', fortran_order True (the array's own layout).
bool save_data_to_npy(const Array2D<double>& a, const std::string& trg_dir, const std::string& of_name, bool as_f32);

IO_param_T set_default_paths(const std::string& src_path);
void read_config(const std::string& f_path, Mesh_Data_T& mesh_data, Properties_T& props, Load_T& load,
                 Num_Param_T& num_param, Sim_Param_T& sim_param, IO_param_T& io_param, Status& status);

}  // namespace hfp3d
