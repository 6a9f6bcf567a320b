/* oracle_check.c -- drives every oracle routine once on a small mesh (a
 * hexagon of 6 elements, one of them lifted off the plane) so that a build
 * with AddressSanitizer + UndefinedBehaviorSanitizer sees every code path:
 * geometry, the element integral, DoF handles (tip types 0..2), assembly
 * (whole, weighted, rows only), right-hand side, DD write-back, stresses, LU
 * and its solve, volume-control border and truncation, cohesion / friction.
 * Prints a checksum; any sanitizer report makes the run fail.  A test helper
 * (the oracle is test infrastructure). */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "hfpx_oracle.h"

int main(void) {
  enum { NN = 7, NE = 6 };
  /* nodes: centre + 6 on a circle; rows x y z crack-id tip-flag 0 */
  double nods[6 * NN];
  for (int j = 0; j < NN; ++j) {
    const double a = 2.0 * 3.14159265358979 * (j - 1) / 6.0;
    nods[6 * j + 0] = j == 0 ? 0.0 : cos(a);
    nods[6 * j + 1] = j == 0 ? 0.0 : sin(a);
    nods[6 * j + 2] = j == 3 ? 0.2 : 0.0;
    nods[6 * j + 3] = 1.0;
    nods[6 * j + 4] = j == 0 ? 0.0 : 1.0;
    nods[6 * j + 5] = 0.0;
  }
  int64_t conn[5 * NE];
  for (int e = 0; e < NE; ++e) {
    conn[5 * e + 0] = 0;
    conn[5 * e + 1] = 1 + e;
    conn[5 * e + 2] = 1 + (e + 1) % 6;
    conn[5 * e + 3] = 1;
    conn[5 * e + 4] = 1;
  }
  double sum = 0.0;
  int64_t dof[NE * 18];
  for (int tip = 0; tip <= 2; ++tip) {
    const int64_t nd = hfpo_dof_h(nods, 6, NN, conn, 5, NE, tip, dof);
    sum += (double)nd;
  }
  const int64_t n = hfpo_dof_h(nods, 6, NN, conn, 5, NE, 0, dof);
  double* A = calloc((size_t)(n * n), sizeof(double));
  double* A2 = calloc((size_t)(n * n), sizeof(double));
  double* R = calloc((size_t)(36 * n), sizeof(double));
  double wts[NN];
  for (int j = 0; j < NN; ++j) wts[j] = 1.0 + 0.1 * j;
  hfpo_assemble(nods, 6, NN, conn, 5, NE, dof, n, 0.4, 0.3, 0.125, 0, 0, 0, 2, A);
  hfpo_assemble_w(nods, 6, NN, conn, 5, NE, dof, n, 0.4, 0.3, 0.125, wts, 0, 0, 1, 2, A2);
  hfpo_assemble_rows(nods, 6, NN, conn, 5, NE, dof, n, 0.4, 0.3, 0.125, 2, 4, 2, R);
  for (int64_t i = 0; i < n * n; ++i) sum += A[i] + 1e-3 * A2[i];
  for (int64_t i = 0; i < 36 * n; ++i) sum += 1e-6 * R[i];
  double s_inf[6] = {0.1, 0.2, 1.0, 0.05, -0.1, 0.02};
  double* rhs = calloc((size_t)n, sizeof(double));
  hfpo_rhs(nods, 6, NN, conn, 5, NE, dof, s_inf, rhs);
  /* LU + solve of A (copy) */
  int32_t* ipiv = malloc(sizeof(int32_t) * (size_t)n);
  memcpy(A2, A, sizeof(double) * (size_t)(n * n));
  const int64_t info = hfpo_getrf(n, A2, n, ipiv, 2);
  double* x = malloc(sizeof(double) * (size_t)n);
  memcpy(x, rhs, sizeof(double) * (size_t)n);
  hfpo_getrs(n, 1, A2, n, ipiv, x, n);
  for (int64_t i = 0; i < n; ++i) sum += x[i];
  double dd[18 * NE];
  hfpo_write_dd(x, dof, NE, dd);
  double pts[3 * 5] = {0.1, 0.2, 0.3, 2.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.5, 0.5, -0.2, 1.5, -1.0, 0.7};
  double S[5 * 6];
  hfpo_stress(nods, 6, NN, conn, 5, NE, 0.4, 0.3, dd, pts, 5, s_inf, 0, 2, S);
  hfpo_stress_w(nods, 6, NN, conn, 5, NE, 0.4, 0.3, dd, pts, 5, s_inf, wts, 1, 2, S);
  for (int i = 0; i < 30; ++i) sum += S[i];
  /* the element integral and geometry directly */
  double vert[9] = {0, 0.1, 0, 1.8, 0, 0, 1.2, 1.8, 0}, rt[9], tau[6], sfm[72], cp[18], M[108];
  hfpo_el_geom(vert, 0.125, rt, tau, sfm, cp);
  hfpo_local_submatrix(1.0, 0.35, -1.21, 1.0, 1.0, tau, sfm, M);
  for (int i = 0; i < 108; ++i) sum += M[i];
  double sfi[6];
  hfpo_sf_integral(tau, sfm, sfi);
  double sd[3], xyz[3] = {0.3, 0.2, 0.0};
  hfpo_sneddon_dd(0.4, 0.3, 1.0, 1.0, xyz, sd);
  sum += sfi[3] + sd[2];
  /* volume control: border, then truncation to the first half of the DoF */
  const int64_t n1 = n + 1;
  double* V = calloc((size_t)(n1 * n1), sizeof(double));
  for (int64_t c = 0; c < n; ++c) memcpy(V + c * n1, A + c * n, sizeof(double) * (size_t)n);
  double* row = calloc((size_t)n, sizeof(double));
  double* col = calloc((size_t)n, sizeof(double));
  hfpo_vc_border(nods, 6, NN, conn, 5, NE, dof, 0, row, col);
  hfpo_vc_border_w(nods, 6, NN, conn, 5, NE, dof, wts, 0, row, col);
  for (int64_t d = 0; d < n; ++d) {
    V[n + d * n1] = row[d];
    V[d + n * n1] = col[d];
  }
  int64_t used[NE * 18];
  int64_t nu = 0;
  for (int e = 0; e < NE; ++e)
    for (int c = 0; c < 18; ++c) used[e + NE * c] = e < NE / 2 ? nu++ : -1;
  double* Mv = calloc((size_t)((nu + 1) * (nu + 1)), sizeof(double));
  double* rv = calloc((size_t)(nu + 1), sizeof(double));
  double* dt = calloc((size_t)n, sizeof(double));
  for (int64_t d = 0; d < n; ++d) dt[d] = 0.01 * (double)d;
  const int rc = hfpo_mod_system_vc(V, n1, n, dof, used, NE, 18, nu, dt, n, 0.5, Mv, rv);
  for (int64_t i = 0; i < (nu + 1) * (nu + 1); ++i) sum += 1e-3 * Mv[i];
  /* cohesion / friction on the nodal DD */
  const int64_t nn = 6 * NE;
  int32_t mid[6 * NE], cmid[4] = {0, 1, 1, 0};
  double par[12] = {1.0, 0.5, 1.0, 0.3, 0.6, 0.4, 2.0, 0.1, 0.5, 0.2, 0.7, 0.2};
  double mo[6 * NE], ms[6 * NE], fc[6 * NE], sc[6 * NE], oc[6 * NE];
  for (int i = 0; i < nn; ++i) {
    mid[i] = i % 2;
    mo[i] = 0.01 * i;
    ms[i] = 0.02 * i;
  }
  hfpo_match_f_c(nn, dd, nn, mid, cmid, 2, par, mo, ms, fc, sc, oc);
  hfpo_update_mrd(nn, dd, nn, mid, 2, par, mo, ms);
  for (int i = 0; i < nn; ++i) sum += fc[i] + sc[i] + oc[i] + mo[i] + ms[i];
  printf("{\"n_dof\": %lld, \"info\": %lld, \"vc\": %d, \"checksum\": %.17g}\n", (long long)n, (long long)info, rc,
         sum);
  free(A);
  free(A2);
  free(R);
  free(rhs);
  free(ipiv);
  free(x);
  free(V);
  free(row);
  free(col);
  free(Mv);
  free(rv);
  free(dt);
  return isfinite(sum) ? 0 : 1;
}
