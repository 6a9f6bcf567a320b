// hfpx3d_fc_check.cpp -- runs the C++ host mirror of the cohesion/friction
// laws (hfp3d::match_f_c, then hfp3d::update_mrd) on the device over a
// deterministic node set, and prints inputs and outputs as one JSON object
// so a test can recompute them independently.  Usage: hfpx3d_fc_check
// [n_nod] (default 2000).
#include <cstdio>
#include <cstdlib>
#include <string>

#include "../../hfpx3d_amd/host/hfp3d.hpp"

namespace {

// 64-bit LCG (Knuth's MMIX constants), uniform doubles in [0, 1)
struct Lcg {
  uint64_t s;
  double next() {
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    return (double)(s >> 11) * (1.0 / 9007199254740992.0);
  }
};

void print_vec(const char* name, const std::vector<double>& v, bool last = false) {
  std::printf("\"%s\": [", name);
  for (size_t i = 0; i < v.size(); ++i) std::printf(i ? ", %.17g" : "%.17g", v[i]);
  std::printf(last ? "]" : "], ");
}

}  // namespace

int main(int argc, char** argv) {
  const int64_t n = argc > 1 ? std::atoll(argv[1]) : 2000;
  hfp3d::Properties_T prop;
  prop.n_surface = 3;
  prop.c_model_id = hfp3d::Array2D<int32_t>(3, 2);
  const int32_t ids[3][2] = {{0, 0}, {1, 1}, {2, 1}};
  for (int m = 0; m < 3; ++m) {
    prop.c_model_id(m, 0) = ids[m][0];
    prop.c_model_id(m, 1) = ids[m][1];
  }
  prop.f_c_param = {{1.0, 2.0, 0.5, 3.0, 0.6, 0.2}, {0.5, 1.5, 1.0, 0.5, 0.8, 0.1}, {2.0, 1.0, 2.0, 1.0, 0.7, 0.3}};
  Lcg g{12345};
  hfp3d::Array2D<double> dd(n, 3);
  std::vector<int32_t> mat(n);
  hfp3d::Frac_State_T fs;
  fs.mr_open.resize(n);
  fs.mr_slip.resize(n);
  for (int64_t i = 0; i < n; ++i) {
    for (int k = 0; k < 3; ++k) dd(i, k) = 2.4 * g.next() - 1.2;
    mat[i] = (int32_t)(3.0 * g.next());
    fs.mr_open[i] = 1.2 * g.next();
    fs.mr_slip[i] = 1.2 * g.next();
  }
  const std::vector<double> mo0 = fs.mr_open, ms0 = fs.mr_slip;
  hfp3d::match_f_c(prop, mat, dd, false, fs);
  hfp3d::update_mrd(prop, mat, dd, false, fs);
  std::vector<double> ddv(dd.data(), dd.data() + 3 * n), matv(mat.begin(), mat.end());
  std::printf("{");
  print_vec("dd", ddv);
  print_vec("mat", matv);
  print_vec("mr_open0", mo0);
  print_vec("mr_slip0", ms0);
  print_vec("friction_coef", fs.friction_coef);
  print_vec("slip_cohesion", fs.slip_cohesion);
  print_vec("open_cohesion", fs.open_cohesion);
  print_vec("mr_open", fs.mr_open);
  print_vec("mr_slip", fs.mr_slip, true);
  std::printf("}\n");
  return 0;
}
