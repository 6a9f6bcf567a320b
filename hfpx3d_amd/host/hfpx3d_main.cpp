// hfpx3d_main.cpp -- command-line driver equivalent to the reference's main()
// (main.cpp:44-294): read config.toml, then the timed stages
//   Assembly       = DoF handle + influence matrix + right-hand side
//   Solution       = LU factorisation + solve
//   Postprocessing = stresses at the observation points
// and the CSV (or, for output_format npy64 / npy32, .npy) outputs the
// configuration asks for.  The configuration path is
// the first argument (default ./config.toml) instead of being derived from
// the source file's location.
#include <chrono>
#include <iostream>

#include "hfp3d.hpp"

using clk = std::chrono::steady_clock;
static double secs(clk::time_point a) { return std::chrono::duration<double>(clk::now() - a).count(); }

// output_format "npy64" / "npy32" writes NumPy files (name with ".npy" in
// place of the configured ".csv"); anything else writes CSV, like the
// reference's default branch.
static bool save_out(const hfp3d::Array2D<double>& a, const hfp3d::IO_param_T& io, std::string& name) {
  if (io.out_f_format == "npy64" || io.out_f_format == "npy32") {
    if (name.size() >= 4 && name.compare(name.size() - 4, 4, ".csv") == 0) name.resize(name.size() - 4);
    name += ".npy";
    return hfp3d::save_data_to_npy(a, io.output_dir, name, io.out_f_format == "npy32");
  }
  return hfp3d::save_data_to_csv(a, io.output_dir, name);
}

int main(int argc, char** argv) {
  const std::string cfg = argc > 1 ? argv[1] : "config.toml";
  hfp3d::Mesh_Data_T md;
  hfp3d::Properties_T props;
  hfp3d::Load_T load;
  hfp3d::Num_Param_T np;
  hfp3d::Sim_Param_T sp;
  hfp3d::IO_param_T io;
  hfp3d::Status status;
  hfp3d::read_config(cfg, md, props, load, np, sp, io, status);
  status.abortOnError();
  const int64_t n_el = md.mesh.conn.size(1);

  std::cout << "Assembly";
  auto t = clk::now();
  md.dof_h_dd = hfp3d::make_dof_h_crack(md.mesh, 2, np.tip_type);
  hfp3d::SAE_T sae;
  sae.matrix = hfp3d::make_3dbem_matrix_s(props.shear_m[0], props.poiss_r[0], md.mesh, np, md.dof_h_dd);
  sae.n_dof = md.dof_h_dd.n_dof;
  hfp3d::add_s_inf_to_3dbem_rhs(md, load, sae);
  std::cout << ": " << secs(t) << "s" << std::endl;
  std::cout << 18 * n_el << " DoF Total" << std::endl;
  std::cout << 18 * n_el - sae.n_dof << " Fixed DoF" << std::endl;

  if (sp.do_save_matrix) {
    std::string name = io.matr_f_name;
    const bool ok = save_out(sae.matrix.to_host(), io, name);
    std::cout << (ok ? "Matrix saved to " + io.output_dir + name : std::string("Cannot save the matrix"))
              << std::endl;
  }

  std::cout << "Solution";
  t = clk::now();
  hfp3d::LU lu(std::move(sae.matrix), status);
  status.abortOnError();
  const std::vector<double> dd_v = lu.solve(sae.rhs_v);
  std::cout << ": " << secs(t) << "s" << std::endl;
  hfp3d::write_dd_vector_to_md(dd_v, md.dof_h_dd, md);

  if (sp.do_save_solution) {
    // 6 nodes per element (beta = 0) and their DD; column 3 is unused
    hfp3d::Array2D<double> out(6 * n_el, 7, 0.0);
    for (int64_t e = 0; e < n_el; ++e) {
      double v[3][3];
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) v[k][l] = md.mesh.nods(l, md.mesh.conn(k, e));
      for (int q = 0; q < 6; ++q) {
        const int a = q < 3 ? q : (q - 3 + 1) % 3, b = q < 3 ? q : (q - 3 + 2) % 3;
        for (int l = 0; l < 3; ++l) {
          out(6 * e + q, l) = 0.5 * (v[a][l] + v[b][l]);
          out(6 * e + q, l + 4) = md.dd(6 * e + q, l);
        }
      }
    }
    std::string name = io.out_f_name;
    const bool ok = save_out(out, io, name);
    std::cout << (ok ? "Solution (nodes + DD) saved to " + io.output_dir + name
                     : std::string("Cannot save the solution"))
              << std::endl;
  }

  if (sp.do_postprocess) {
    // the reference selects the point loader by the OUTPUT format (main.cpp:246)
    hfp3d::Array2D<double> pts = io.out_f_format == "csv"
                                     ? hfp3d::load_crd_from_csv(io.obs_p_dir, io.obs_p_f_name, status)
                                     : hfp3d::load_crd_from_numpy(io.obs_p_dir, io.obs_p_f_name, status);
    status.abortOnError();
    std::cout << "Postprocessing";
    t = clk::now();
    hfp3d::Array2D<double> s =
        hfp3d::make_3dbem_stress_f_s(props.shear_m[0], props.poiss_r[0], md, load, np, pts);
    std::cout << ": " << secs(t) << "s" << std::endl;
    std::string name = io.strs_f_name;
    const bool ok = save_out(s, io, name);
    std::cout << (ok ? "Stresses saved to " + io.output_dir + name : std::string("Cannot save stresses"))
              << std::endl;
  }
  return 0;
}
