// hfpx3d_host_check.cpp -- prints what the host layer parses, for the CPU
// tests (no device call): read_config results, mesh sizes and checksums,
// and the DoF handle.  Usage: hfpx3d_host_check <config.toml>; with
// "--npy <dir>" it writes a small array through save_data_to_npy instead;
// with "--npy-mesh <nodes.npy> <elems.npy>" it runs the npy mesh loader.
// A test helper, not part of the library.
#include <cstdio>
#include <string>

#include "../../hfpx3d_amd/host/hfp3d.hpp"

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  if (argc == 4 && std::string(argv[1]) == "--npy-mesh") {
    // loads nodes (argv[2]) and elements (argv[3]) with index origin 1 and
    // prints the shapes, or the loader's error
    hfp3d::Mesh_Geom_T m;
    hfp3d::Status st;
    hfp3d::load_npy_mesh_files(argv[3], argv[2], 1, m, st);
    if (!st.ok()) {
      std::printf("{\"error\": \"%s\"}\n", st.error.c_str());
      return 1;
    }
    std::printf("{\"nods\": [%lld, %lld], \"conn\": [%lld, %lld]}\n", (long long)m.nods.size(0),
                (long long)m.nods.size(1), (long long)m.conn.size(0), (long long)m.conn.size(1));
    return 0;
  }
  if (argc == 3 && std::string(argv[1]) == "--npy") {
    // writes a 3 x 5 array (entry (i, j) = i + 10 j + 0.1) as check64.npy and check32.npy
    hfp3d::Array2D<double> a(3, 5);
    for (int64_t j = 0; j < 5; ++j)
      for (int64_t i = 0; i < 3; ++i) a(i, j) = (double)i + 10.0 * (double)j + 0.1;
    const bool ok = hfp3d::save_data_to_npy(a, argv[2], "check64.npy", false) &&
                    hfp3d::save_data_to_npy(a, argv[2], "check32.npy", true);
    return ok ? 0 : 1;
  }
  hfp3d::Mesh_Data_T md;
  hfp3d::Properties_T props;
  hfp3d::Load_T load;
  hfp3d::Num_Param_T np;
  hfp3d::Sim_Param_T sp;
  hfp3d::IO_param_T io;
  hfp3d::Status st;
  hfp3d::read_config(argv[1], md, props, load, np, sp, io, st);
  if (!st.ok()) {
    std::printf("{\"error\": \"%s\"}\n", st.error.c_str());
    return 1;
  }
  double ns = 0, cs = 0;
  for (int64_t j = 0; j < md.mesh.nods.size(1); ++j)
    for (int64_t i = 0; i < md.mesh.nods.size(0); ++i) ns += md.mesh.nods(i, j) * (1 + i + 3 * j);
  for (int64_t j = 0; j < md.mesh.conn.size(1); ++j)
    for (int64_t i = 0; i < md.mesh.conn.size(0); ++i) cs += (double)md.mesh.conn(i, j) * (1 + i + 5 * j);
  hfp3d::DoF_Handle_T d = hfp3d::make_dof_h_crack(md.mesh, 2, np.tip_type);
  double ds = 0;
  for (int64_t e = 0; e < d.dof_h.size(0); ++e)
    for (int c = 0; c < 18; ++c) ds += (double)d.dof_h(e, c) * (1 + c + 18 * e);
  std::printf(
      "{\"G\": %.17g, \"nu\": %.17g, \"beta\": %.17g, \"tip_type\": %d, \"is_dd_local\": %d, "
      "\"s_inf\": [%.17g, %.17g, %.17g, %.17g, %.17g, %.17g], \"save_matrix\": %d, \"save_solution\": %d, "
      "\"postprocess\": %d, \"nods\": [%lld, %lld], \"conn\": [%lld, %lld], \"nods_sum\": %.17g, \"conn_sum\": %.17g, "
      "\"n_dof\": %lld, \"dof_sum\": %.17g, \"input_dir\": \"%s\", \"obs_dir\": \"%s\", \"obs_file\": \"%s\", "
      "\"out_dir\": \"%s\", \"out_file\": \"%s\", \"strs_file\": \"%s\", \"matr_file\": \"%s\", \"out_format\": \"%s\"}\n",
      props.shear_m[0], props.poiss_r[0], np.beta, np.tip_type, (int)np.is_dd_local, load.s_inf[0], load.s_inf[1],
      load.s_inf[2], load.s_inf[3], load.s_inf[4], load.s_inf[5], (int)sp.do_save_matrix, (int)sp.do_save_solution,
      (int)sp.do_postprocess, (long long)md.mesh.nods.size(0), (long long)md.mesh.nods.size(1),
      (long long)md.mesh.conn.size(0), (long long)md.mesh.conn.size(1), ns, cs, (long long)d.n_dof, ds,
      io.input_dir.c_str(), io.obs_p_dir.c_str(), io.obs_p_f_name.c_str(), io.output_dir.c_str(),
      io.out_f_name.c_str(), io.strs_f_name.c_str(), io.matr_f_name.c_str(), io.out_f_format.c_str());
  return 0;
}
