/*
 * hfpx_oracle.c -- CPU restatement of the HFPx3D hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker that the HIP path is compared
 * against; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  It is never part of the product path.
 *
 * What it restates (reference paths relative to the reference root):
 *   element geometry        src/Core/element_utilities.cpp:62-431 (uniform and
 *                           non-uniform edge partitions)
 *   per-pair kernel         src/Elasticity/elasticity_kernel_integration.cpp:34-205
 *                           (values specified by src/Elasticity/h_potential.cpp;
 *                            the integrals are re-derived here, see below)
 *   traction / rotation     src/Core/tensor_utilities.cpp:21-91
 *   DoF handle              src/Core/surface_mesh_utilities.cpp:28-112
 *   assembly / rhs / stress src/Solvers/system_assembly.cpp:30-455
 *   DD write-back           src/Core/surface_mesh_utilities.cpp:256-328
 *   LU (dgetrf / dgetrs)    il/linear_algebra/dense/factorization/LU.h:79-138
 *                           (LAPACK right-looking LU with partial pivoting)
 *   Sneddon opening         src/Solvers/test_radial_crack_static.cpp:316-337
 *   volume control          src/Solvers/system_assembly.cpp:458-629,
 *                           src/Core/element_utilities.cpp:484-595
 *   cohesion / friction     src/Development/cohesion_friction.cpp:22-189
 *                           (match_f_c, update_mrd)
 *   cohesion / friction     src/Development/c_f_iteration.cpp:27-436
 *   iteration step          (vc_cf_iteration; parity pinned by the plain
 *                           volume-control solve, the Newton fixed point
 *                           and the friction bound -- the reference file
 *                           does not build here)
 *
 * The per-pair kernel.  The stress at X due to a displacement discontinuity
 * D_i P(xi) on a flat element (plane x3 = 0 of the element frame) follows
 * from Kelvin's solution; each stress component is a combination of second
 * derivatives of the single-layer potential Phi[P](X) = int P/|X - xi| dA and
 * of x3 times its third derivatives (derived symbolically from Kelvin's
 * tensor).  Integrating by parts in the element plane turns every derivative
 * of Phi into line integrals along the three edges of moments
 * int a^j / r^(2l+1) da (a = arc coordinate, r^2 = a^2 + c^2) plus the
 * solid-angle term dPhi[1]/dx3 and Phi[1].  The hypersingular in-plane case
 * (x3 = 0, the point inside the element) is the x3 -> 0 limit of the same
 * formulas with the solid-angle term taken at its principal value (0).
 * Parity status: pinned against the reference build (oracle/_ref), the
 * MATLAB fixtures of the reference tests and brute-force quadrature.
 */
#include "hfpx_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

static const double kPi = 3.14159265358979323846;

/* ---------------------------------------------------------------------- */
/* element geometry (element_utilities.cpp:62-87, 89-118, 170-238, 364-395) */
/* ---------------------------------------------------------------------- */

static void cross3(const double* a, const double* b, double* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
static void normalize3(double* a) {
  double n = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
  a[0] /= n; a[1] /= n; a[2] /= n;
}

/* rows of rt: e1 = (v1-v0)/|.|, e3 = normalize(e1 x (v2-v0)), e2 = e3 x e1
 * (element_utilities.cpp:62-87); rt col-major: rt[row + 3*col]. */
static void el_rt(const double* vert, double* rt) {
  double a1[3], a2[3], e1[3], e2[3], e3[3];
  for (int k = 0; k < 3; ++k) {
    a1[k] = vert[k + 3] - vert[k];
    a2[k] = vert[k + 6] - vert[k];
  }
  memcpy(e1, a1, sizeof e1);
  normalize3(e1);
  cross3(e1, a2, e3);
  normalize3(e3);
  cross3(e3, e1, e2);
  normalize3(e2);
  for (int k = 0; k < 3; ++k) {
    rt[0 + 3 * k] = e1[k];
    rt[1 + 3 * k] = e2[k];
    rt[2 + 3 * k] = e3[k];
  }
}

/* y = R (x - o) */
static void to_local(const double* rt, const double* o, const double* x, double* y) {
  double d[3] = {x[0] - o[0], x[1] - o[1], x[2] - o[2]};
  for (int i = 0; i < 3; ++i)
    y[i] = rt[i] * d[0] + rt[i + 3] * d[1] + rt[i + 6] * d[2];
}

/* Shape functions of the 6-node triangle in complex monomials
 * [1, t, tb, t^2, tb^2, t*tb]: vertices N_v = L_v (2 L_v - 1), edge nodes
 * N_{3+v} = 4 L_{v+1} L_{v+2} (node 3+v sits mid-edge opposite vertex v,
 * element_utilities.cpp:380-393).  The barycentric coordinates are written
 * as L_k = l0 + l1 t + conj(l1) tb. */
static void lin_mul(const cplx* a, const cplx* b, double s, cplx* q) {
  /* q += s * (a0 + a1 t + a2 tb)(b0 + b1 t + b2 tb) */
  q[0] += s * a[0] * b[0];
  q[1] += s * (a[0] * b[1] + a[1] * b[0]);
  q[2] += s * (a[0] * b[2] + a[2] * b[0]);
  q[3] += s * a[1] * b[1];
  q[4] += s * a[2] * b[2];
  q[5] += s * (a[1] * b[2] + a[2] * b[1]);
}

static void el_sfm(const cplx* t, cplx* sfm /* 6x6 col-major */) {
  /* twice the signed area */
  double area2 = cimag(conj(t[1] - t[0]) * (t[2] - t[0]));
  cplx L[3][3];
  for (int k = 0; k < 3; ++k) {
    cplx ta = t[(k + 1) % 3], tb = t[(k + 2) % 3];
    /* L_k(t) = Im(conj(ta - t) (tb - t)) / area2 */
    L[k][0] = cimag(conj(ta) * tb) / area2;
    L[k][1] = (conj(tb) - conj(ta)) / (2.0 * I) / area2;
    L[k][2] = (ta - tb) / (2.0 * I) / area2;
  }
  for (int j = 0; j < 36; ++j) sfm[j] = 0;
  for (int v = 0; v < 3; ++v) {
    cplx q[6] = {0, 0, 0, 0, 0, 0};
    lin_mul(L[v], L[v], 2.0, q);
    q[0] -= L[v][0]; q[1] -= L[v][1]; q[2] -= L[v][2];
    for (int c = 0; c < 6; ++c) sfm[v + 6 * c] = q[c];
    cplx e[6] = {0, 0, 0, 0, 0, 0};
    lin_mul(L[(v + 1) % 3], L[(v + 2) % 3], 4.0, e);
    for (int c = 0; c < 6; ++c) sfm[3 + v + 6 * c] = e[c];
  }
}

/* Shape functions for a non-uniform edge partition
 * (make_el_sfm_nonuniform, element_utilities.cpp:240-322): the edge node
 * opposite vertex v sits at L_a = w_a / (w_a + w_b), L_b = w_b / (w_a + w_b)
 * on the edge (a, b) = (v+1, v+2).  The quadratic Lagrange basis of the six
 * nodes, derived here from N_j(node_k) = delta_jk:
 *   N_{3+v} = (w_a + w_b)^2 / (w_a w_b) L_a L_b,
 *   N_v     = L_v - sum over the edges (v, u) of L_v(edge node) N_edge
 *           = L_v^2 - (w_v / w_a) L_v L_a - (w_v / w_b) L_v L_b
 * (using L_v = L_v (L_0 + L_1 + L_2)).  Equal weights give the uniform basis. */
static void el_sfm_w(const cplx* t, const double* w, cplx* sfm /* 6x6 col-major */) {
  double area2 = cimag(conj(t[1] - t[0]) * (t[2] - t[0]));
  cplx L[3][3];
  for (int k = 0; k < 3; ++k) {
    cplx ta = t[(k + 1) % 3], tb = t[(k + 2) % 3];
    L[k][0] = cimag(conj(ta) * tb) / area2;
    L[k][1] = (conj(tb) - conj(ta)) / (2.0 * I) / area2;
    L[k][2] = (ta - tb) / (2.0 * I) / area2;
  }
  for (int j = 0; j < 36; ++j) sfm[j] = 0;
  for (int v = 0; v < 3; ++v) {
    const int a = (v + 1) % 3, b = (v + 2) % 3;
    cplx q[6] = {0, 0, 0, 0, 0, 0};
    lin_mul(L[v], L[v], 1.0, q);
    lin_mul(L[v], L[a], -w[v] / w[a], q);
    lin_mul(L[v], L[b], -w[v] / w[b], q);
    for (int c = 0; c < 6; ++c) sfm[v + 6 * c] = q[c];
    cplx e[6] = {0, 0, 0, 0, 0, 0};
    lin_mul(L[a], L[b], (w[a] + w[b]) * (w[a] + w[b]) / (w[a] * w[b]), e);
    for (int c = 0; c < 6; ++c) sfm[3 + v + 6 * c] = e[c];
  }
}

void hfpo_el_geom_w(const double* vert, const double* wts, double beta, double* rt,
                    double* tau, double* sfm, double* cp) {
  el_rt(vert, rt);
  cplx t[3];
  for (int j = 0; j < 3; ++j) {
    double y[3];
    to_local(rt, vert, vert + 3 * j, y);
    t[j] = y[0] + I * y[1];
    tau[2 * j] = y[0];
    tau[2 * j + 1] = y[1];
  }
  cplx s[36];
  if (wts) el_sfm_w(t, wts, s);
  else el_sfm(t, s);
  for (int j = 0; j < 36; ++j) {
    sfm[2 * j] = creal(s[j]);
    sfm[2 * j + 1] = cimag(s[j]);
  }
  /* collocation points (el_cp_uniform / el_cp_nonuniform,
   * element_utilities.cpp:364-431): the edge point at the weighted node */
  double c[3];
  for (int k = 0; k < 3; ++k) c[k] = (vert[k] + vert[k + 3] + vert[k + 6]) / 3.0;
  for (int v = 0; v < 3; ++v) {
    int m = (v + 1) % 3, l = (m + 1) % 3;
    for (int k = 0; k < 3; ++k) {
      cp[k + 3 * v] = (1.0 - beta) * vert[k + 3 * v] + beta * c[k];
      if (wts)
        cp[k + 3 * (v + 3)] = (1.0 - beta) * (wts[m] * vert[k + 3 * m] + wts[l] * vert[k + 3 * l]) /
                                  (wts[m] + wts[l]) + beta * c[k];
      else
        cp[k + 3 * (v + 3)] = 0.5 * (1.0 - beta) * (vert[k + 3 * m] + vert[k + 3 * l]) + beta * c[k];
    }
  }
}

void hfpo_el_geom(const double* vert, double beta, double* rt, double* tau,
                  double* sfm, double* cp) {
  hfpo_el_geom_w(vert, NULL, beta, rt, tau, sfm, cp);
}

/* ---------------------------------------------------------------------- */
/* the per-pair kernel                                                      */
/* ---------------------------------------------------------------------- */

/* Edge moments M[j][l] = int_{am}^{an} a^j (a^2 + c2)^(-l - 1/2) da,
 * j <= 4, l <= 2, in forms that keep full relative accuracy when the point
 * is close to the edge line (small c) and the end coordinates share a sign. */
typedef struct { double m[5][3]; } Moments;

static void edge_moments(double am, double an, double c2, Moments* M) {
  double rm = sqrt(am * am + c2), rn = sqrt(an * an + c2);
  /* Lambda = int da / r */
  double lam, f01, f02;
  /* per-end pieces with the 1/c^k constants split off */
  double gm1, gn1, gm2, gn2;
  if (am >= 0.0) {
    lam = log((an + rn) / (am + rm));
    gm1 = -1.0 / (rm * (rm + am));
    gn1 = -1.0 / (rn * (rn + an));
    gm2 = -(2.0 * rm + am) / (3.0 * rm * rm * rm * (rm + am) * (rm + am));
    gn2 = -(2.0 * rn + an) / (3.0 * rn * rn * rn * (rn + an) * (rn + an));
    f01 = gn1 - gm1;
    f02 = gn2 - gm2;
  } else if (an <= 0.0) {
    lam = log((rm - am) / (rn - an));
    gm1 = 1.0 / (rm * (rm - am));
    gn1 = 1.0 / (rn * (rn - an));
    gm2 = (2.0 * rm - am) / (3.0 * rm * rm * rm * (rm - am) * (rm - am));
    gn2 = (2.0 * rn - an) / (3.0 * rn * rn * rn * (rn - an) * (rn - an));
    f01 = gn1 - gm1;
    f02 = gn2 - gm2;
  } else { /* am < 0 < an: the foot of the point lies on the edge span */
    lam = log((an + rn) * (rm - am) / c2);
    gm1 = 1.0 / (rm * (rm - am));
    gn1 = -1.0 / (rn * (rn + an));
    gm2 = (2.0 * rm - am) / (3.0 * rm * rm * rm * (rm - am) * (rm - am));
    gn2 = -(2.0 * rn + an) / (3.0 * rn * rn * rn * (rn + an) * (rn + an));
    f01 = 2.0 / c2 + gn1 - gm1;
    f02 = 4.0 / (3.0 * c2 * c2) + gn2 - gm2;
  }
  /* differences of r, 1/r, 1/r^3 without cancellation */
  double dr = (an - am) * (an + am) / (rn + rm);        /* rn - rm */
  double dinv = -dr / (rn * rm);                        /* 1/rn - 1/rm */
  double dinv3 = -dr * (rn * rn + rn * rm + rm * rm) / (rm * rm * rm * rn * rn * rn);
  M->m[0][0] = lam;
  M->m[1][0] = dr;
  M->m[0][1] = f01;
  M->m[1][1] = -dinv;
  M->m[2][1] = lam - c2 * f01;
  M->m[3][1] = dr + c2 * dinv;
  M->m[0][2] = f02;
  M->m[1][2] = -dinv3 / 3.0;
  M->m[2][2] = f01 - c2 * f02;
  M->m[3][2] = -dinv - c2 * (-dinv3 / 3.0);
  M->m[4][2] = lam - 2.0 * c2 * f01 + c2 * c2 * f02;
  M->m[2][0] = M->m[3][0] = M->m[4][0] = 0.0; /* unused */
  M->m[4][1] = 0.0;
}

/* per-end solid-angle function (sign handled by the caller) */
static double solid_end(double a, double s, double ax3) {
  if (s == 0.0) return 0.0;
  double r = sqrt(a * a + s * s + ax3 * ax3);
  return atan(a * s * (a * a + s * s) / ((r + ax3) * (s * s * r + ax3 * a * a)));
}

static cplx integ(const cplx* p, int np, int l, const Moments* M) {
  cplx acc = 0;
  for (int j = 0; j < np; ++j) acc += p[j] * M->m[j][l];
  return acc;
}

void hfpo_local_submatrix(double shear_m, double poiss_r, double h, double z_re,
                          double z_im, const double* tau, const double* sfm,
                          double* out) {
  const double nu = poiss_r;
  const double x3 = -h; /* h is minus the local x3 coordinate (element_utilities.cpp:163) */
  const cplx z = z_re + I * z_im, zc = conj(z);
  cplx t[3];
  for (int k = 0; k < 3; ++k) t[k] = tau[2 * k] + I * tau[2 * k + 1];

  /* shape functions re-centred at z: coefficients of [1, w, wb, w^2, wb^2, w wb],
   * w = t - z (the binomial shift of element_utilities.cpp:339-360) */
  cplx p[6][6];
  for (int j = 0; j < 6; ++j) {
    cplx q[6];
    for (int c = 0; c < 6; ++c) q[c] = sfm[2 * (j + 6 * c)] + I * sfm[2 * (j + 6 * c) + 1];
    p[j][0] = q[0] + q[1] * z + q[2] * zc + q[3] * z * z + q[4] * zc * zc + q[5] * z * zc;
    p[j][1] = q[1] + 2.0 * q[3] * z + q[5] * zc;
    p[j][2] = q[2] + 2.0 * q[4] * zc + q[5] * z;
    p[j][3] = q[3];
    p[j][4] = q[4];
    p[j][5] = q[5];
  }

  /* edge geometry */
  double es[3];
  cplx enu[3];
  Moments em[3];
  double w1 = 0.0, phi1 = 0.0;
  cplx enu_r = 0, enub_r = 0;
  const double ax3 = fabs(x3);
  const double sgn3 = (x3 > 0) ? 1.0 : ((x3 < 0) ? -1.0 : 0.0);
  for (int m = 0; m < 3; ++m) {
    cplx tm = t[m], tn = t[(m + 1) % 3];
    double L = cabs(tn - tm);
    cplx th = (tn - tm) / L;
    cplx nv = -I * th; /* outward normal of the counter-clockwise boundary */
    double s = creal((tm - z) * conj(nv));
    double am = creal((tm - z) * conj(th));
    double an = am + L;
    double c2 = s * s + x3 * x3;
    edge_moments(am, an, c2, &em[m]);
    es[m] = s;
    enu[m] = nv;
    w1 += -sgn3 * (solid_end(an, s, ax3) - solid_end(am, s, ax3));
    phi1 += s * em[m].m[0][0];
    enu_r += nv * em[m].m[0][0];
    enub_r += conj(nv) * em[m].m[0][0];
  }
  phi1 += x3 * w1;

  const double K = shear_m / (16.0 * kPi * (1.0 - nu));
  for (int j = 0; j < 6; ++j) {
    const cplx p00 = p[j][0], p10 = p[j][1], p01 = p[j][2], p20 = p[j][3],
               p02 = p[j][4], p11 = p[j][5];
    cplx A = 0, B = 0, C = 0, D = 0, F = 0, Gv = 0, H = 0;
    for (int m = 0; m < 3; ++m) {
      const double s = es[m];
      const cplx nv = enu[m], nb = conj(nv), n2 = nv * nv;
      const Moments* M = &em[m];
      /* density along the edge: w = (s + i a) nu */
      cplx P[3], Q[2];
      P[0] = p00 + p10 * s * nv + p01 * s * nb + p20 * s * s * n2 + p02 * s * s * nb * nb + p11 * s * s;
      P[1] = p10 * I * nv - p01 * I * nb + p20 * 2.0 * I * s * n2 - p02 * 2.0 * I * s * nb * nb;
      P[2] = -p20 * n2 - p02 * nb * nb + p11;
      /* d/d(wb) of the density: p01 + 2 p02 wb + p11 w */
      Q[0] = p01 + 2.0 * p02 * s * nb + p11 * s * nv;
      Q[1] = -2.0 * p02 * I * nb + p11 * I * nv;
      cplx smP[4] = {s * P[0], s * P[1] - I * P[0], s * P[2] - I * P[1], -I * P[2]};
      cplx spP[4] = {s * P[0], s * P[1] + I * P[0], s * P[2] + I * P[1], I * P[2]};
      cplx r2P[5] = {s * s * P[0], s * s * P[1], s * s * P[2] + P[0], P[1], P[2]};
      cplx sp2P[5] = {s * s * P[0], s * s * P[1] + 2.0 * I * s * P[0],
                      s * s * P[2] + 2.0 * I * s * P[1] - P[0], 2.0 * I * s * P[2] - P[1], -P[2]};
      cplx smQ[3] = {s * Q[0], s * Q[1] - I * Q[0], -I * Q[1]};
      cplx spQ[3] = {s * Q[0], s * Q[1] + I * Q[0], I * Q[1]};
      const double m00 = M->m[0][0];
      A += -integ(smP, 4, 1, M) - 2.0 * nb * integ(Q, 2, 0, M);
      B += -n2 * integ(spP, 4, 1, M) - 2.0 * nv * integ(Q, 2, 0, M);
      C += x3 * nv * integ(P, 3, 1, M);
      Gv += 3.0 * x3 * integ(smP, 4, 2, M) + 2.0 * x3 * nb * integ(Q, 2, 1, M);
      H += 3.0 * x3 * n2 * integ(spP, 4, 2, M) + 2.0 * x3 * nv * integ(Q, 2, 1, M);
      D += 2.0 * nv * integ(P, 3, 1, M) - 3.0 * nv * integ(r2P, 5, 2, M) -
           2.0 * integ(smQ, 3, 1, M) - 8.0 * p02 * nb * m00;
      F += -3.0 * n2 * nv * integ(sp2P, 5, 2, M) - 2.0 * n2 * integ(spQ, 3, 1, M) -
           8.0 * p02 * nv * m00;
    }
    A += 4.0 * p11 * phi1;
    B += 8.0 * p02 * phi1;
    C += 2.0 * (p01 * w1 + x3 * (p11 * enu_r + 2.0 * p02 * enub_r));
    Gv += 4.0 * p11 * w1;
    H += 8.0 * p02 * w1;
    const double a = creal(A), g = creal(Gv);
    /* stress combos S0 = S11+S22, S2 = S11-S22+2iS12, S33, S1 = S13+iS23 */
    double S0[3], S33[3];
    cplx S2[3], S1[3];
    S0[0] = K * (-8.0 * (1.0 + nu) * creal(C) + 4.0 * x3 * creal(D));
    S2[0] = K * (-8.0 * (1.0 - nu) * C + 2.0 * x3 * (D + F));
    S33[0] = -4.0 * K * x3 * creal(D);
    S1[0] = K * (2.0 * (2.0 - nu) * a + 2.0 * nu * B + 2.0 * x3 * (g + H));
    S0[1] = K * (-8.0 * (1.0 + nu) * cimag(C) + 4.0 * x3 * cimag(D));
    S2[1] = K * (-8.0 * I * (1.0 - nu) * C + 2.0 * I * x3 * (D - F));
    S33[1] = -4.0 * K * x3 * cimag(D);
    S1[1] = I * K * (2.0 * (2.0 - nu) * a - 2.0 * nu * B + 2.0 * x3 * (g - H));
    S0[2] = 4.0 * K * ((1.0 + 2.0 * nu) * a + x3 * g);
    S2[2] = 4.0 * K * ((1.0 - 2.0 * nu) * B + x3 * H);
    S33[2] = 4.0 * K * (a - x3 * g);
    S1[2] = -4.0 * K * x3 * D;
    for (int k = 0; k < 3; ++k) {
      double* col = out + 6 * (3 * j + k);
      col[0] = 0.5 * (S0[k] + creal(S2[k]));
      col[1] = 0.5 * (S0[k] - creal(S2[k]));
      col[2] = S33[k];
      col[3] = 0.5 * cimag(S2[k]);
      col[4] = creal(S1[k]);
      col[5] = cimag(S1[k]);
    }
  }
}

/* ---------------------------------------------------------------------- */
/* DoF handle (surface_mesh_utilities.cpp:28-112)                           */
/* ---------------------------------------------------------------------- */

int64_t hfpo_dof_h(const double* nods, int64_t nods_rows, int64_t n_nodes,
                   const int64_t* conn, int64_t conn_rows, int64_t n_el,
                   int tip_type, int64_t* dof_h) {
  (void)n_nodes;
  const int ext = (nods_rows >= 5 && conn_rows >= 5);
  int64_t n_dof = 18 * n_el, dec = 0;
  for (int64_t e = 0; e < n_el; ++e) {
    int tip[3] = {0, 0, 0};
    if (ext && conn[3 + conn_rows * e] == conn[4 + conn_rows * e]) {
      for (int v = 0; v < 3; ++v) {
        int64_t n = conn[v + conn_rows * e];
        tip[v] = nods[3 + nods_rows * n] == 1.0 && nods[4 + nods_rows * n] == 1.0;
      }
    }
    for (int v = 0; v < 6; ++v) {
      int fixed = 0;
      if (tip_type >= 1 && v < 3 && tip[v]) fixed = 1;
      if (tip_type == 2 && v >= 3 && tip[(v - 3 + 1) % 3] && tip[(v - 3 + 2) % 3]) fixed = 1;
      for (int l = 0; l < 3; ++l) {
        int64_t ld = 3 * v + l;
        if (fixed) {
          dof_h[e + n_el * ld] = -1;
          --n_dof;
          ++dec;
        } else {
          dof_h[e + n_el * ld] = 18 * e + ld - dec;
        }
      }
    }
  }
  return n_dof;
}

/* ---------------------------------------------------------------------- */
/* assembly (system_assembly.cpp:30-219)                                    */
/* ---------------------------------------------------------------------- */

static void el_vert(const double* nods, int64_t nods_rows, const int64_t* conn,
                    int64_t conn_rows, int64_t e, double* v) {
  for (int j = 0; j < 3; ++j) {
    int64_t n = conn[j + conn_rows * e];
    for (int k = 0; k < 3; ++k) v[k + 3 * j] = nods[k + nods_rows * n];
  }
}

/* vertex weights of element e from per-node weights (NULL: uniform) */
static const double* el_wts(const double* node_wts, const int64_t* conn, int64_t conn_rows,
                            int64_t e, double* w) {
  if (!node_wts) return NULL;
  for (int j = 0; j < 3; ++j) w[j] = node_wts[conn[j + conn_rows * e]];
  return w;
}

/* 18x18 traction influence of source element es on the 6 collocation
 * points of target element et, global frame (system_assembly.cpp:88-199). */
static void el2el_block(const double* vs, const double* ws, const double* vt, const double* wt,
                        double shear_m, double poiss_r, double beta, int dd_local,
                        double* blk /* 18x18 col-major */) {
  double rs[9], taus[6], sfms[72], cps[18];
  hfpo_el_geom_w(vs, ws, beta, rs, taus, sfms, cps);
  double rt[9], taut[6], sfmt[72], cpt[18];
  hfpo_el_geom_w(vt, wt, beta, rt, taut, sfmt, cpt);
  double nglob[3] = {-rt[2], -rt[5], -rt[8]};
  double nloc[3];
  for (int i = 0; i < 3; ++i) nloc[i] = rs[i] * nglob[0] + rs[i + 3] * nglob[1] + rs[i + 6] * nglob[2];
  for (int nt = 0; nt < 6; ++nt) {
    double y[3];
    to_local(rs, vs, cpt + 3 * nt, y);
    double S[108];
    hfpo_local_submatrix(shear_m, poiss_r, -y[2], y[0], y[1], taus, sfms, S);
    for (int c = 0; c < 18; ++c) {
      const double* s = S + 6 * c;
      /* traction = N(n) . S  (tensor_utilities.cpp:21-37) */
      double tl[3] = {nloc[0] * s[0] + nloc[1] * s[3] + nloc[2] * s[4],
                      nloc[0] * s[3] + nloc[1] * s[1] + nloc[2] * s[5],
                      nloc[0] * s[4] + nloc[1] * s[5] + nloc[2] * s[2]};
      double tg[3];
      for (int i = 0; i < 3; ++i) tg[i] = rs[0 + 3 * i] * tl[0] + rs[1 + 3 * i] * tl[1] + rs[2 + 3 * i] * tl[2];
      for (int i = 0; i < 3; ++i) blk[3 * nt + i + 18 * c] = tg[i];
    }
    /* DD to global: per source node, T . R_s (system_assembly.cpp:164-185);
     * local DD keep the block as it is (:186-195) */
    for (int ns = 0; ns < 6 && !dd_local; ++ns) {
      double tmp[3][3];
      for (int i = 0; i < 3; ++i)
        for (int jj = 0; jj < 3; ++jj) tmp[i][jj] = blk[3 * nt + i + 18 * (3 * ns + jj)];
      for (int i = 0; i < 3; ++i)
        for (int jj = 0; jj < 3; ++jj)
          blk[3 * nt + i + 18 * (3 * ns + jj)] =
              tmp[i][0] * rs[0 + 3 * jj] + tmp[i][1] * rs[1 + 3 * jj] + tmp[i][2] * rs[2 + 3 * jj];
    }
  }
}

int64_t hfpo_assemble(const double* nods, int64_t nods_rows, int64_t n_nodes,
                      const int64_t* conn, int64_t conn_rows, int64_t n_el,
                      const int64_t* dof_h, int64_t n_dof, double shear_m,
                      double poiss_r, double beta, int64_t t0, int64_t t1,
                      int is_dd_local, int nthreads, double* A) {
  return hfpo_assemble_w(nods, nods_rows, n_nodes, conn, conn_rows, n_el, dof_h, n_dof, shear_m,
                         poiss_r, beta, NULL, t0, t1, is_dd_local, nthreads, A);
}

int64_t hfpo_assemble_w(const double* nods, int64_t nods_rows, int64_t n_nodes,
                        const int64_t* conn, int64_t conn_rows, int64_t n_el,
                        const int64_t* dof_h, int64_t n_dof, double shear_m,
                        double poiss_r, double beta, const double* node_wts, int64_t t0,
                        int64_t t1, int is_dd_local, int nthreads, double* A) {
  (void)n_nodes;
  if (t1 <= t0) { t0 = 0; t1 = n_el; }
  memset(A, 0, sizeof(double) * (size_t)n_dof * (size_t)n_dof);
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic)
#endif
  for (int64_t es = 0; es < n_el; ++es) {
    double vs[9], vt[9], blk[324], wsb[3], wtb[3];
    el_vert(nods, nods_rows, conn, conn_rows, es, vs);
    const double* ws = el_wts(node_wts, conn, conn_rows, es, wsb);
    for (int64_t et = t0; et < t1; ++et) {
      el_vert(nods, nods_rows, conn, conn_rows, et, vt);
      const double* wt = el_wts(node_wts, conn, conn_rows, et, wtb);
      el2el_block(vs, ws, vt, wt, shear_m, poiss_r, beta, is_dd_local, blk);
      for (int c = 0; c < 18; ++c) {
        int64_t gc = dof_h[es + n_el * c];
        if (gc < 0) continue;
        for (int r = 0; r < 18; ++r) {
          int64_t gr = dof_h[et + n_el * r];
          if (gr < 0) continue;
          A[gr + n_dof * gc] += blk[r + 18 * c];
        }
      }
    }
  }
  return n_dof;
}

/* The rows of target elements [t0, t1) only (the loop of
 * make_3dbem_matrix_s, system_assembly.cpp:61-213, restricted to those
 * targets): R is (18 (t1 - t0)) x n_dof column-major, row 18 (et - t0) + r
 * = the element's local traction slot r (the matrix row dof_h(et, r); left
 * 0 when that slot is fixed), column = source DoF.  For meshes whose full
 * matrix does not fit in host memory. */
void hfpo_assemble_rows(const double* nods, int64_t nods_rows, int64_t n_nodes, const int64_t* conn,
                        int64_t conn_rows, int64_t n_el, const int64_t* dof_h, int64_t n_dof, double shear_m,
                        double poiss_r, double beta, int64_t t0, int64_t t1, int nthreads, double* R) {
  (void)n_nodes;
  const int64_t nr = 18 * (t1 - t0);
  memset(R, 0, sizeof(double) * (size_t)nr * (size_t)n_dof);
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic)
#endif
  for (int64_t es = 0; es < n_el; ++es) {
    double vs[9], vt[9], blk[324], wsb[3], wtb[3];
    el_vert(nods, nods_rows, conn, conn_rows, es, vs);
    const double* ws = el_wts(NULL, conn, conn_rows, es, wsb);
    for (int64_t et = t0; et < t1; ++et) {
      el_vert(nods, nods_rows, conn, conn_rows, et, vt);
      const double* wt = el_wts(NULL, conn, conn_rows, et, wtb);
      el2el_block(vs, ws, vt, wt, shear_m, poiss_r, beta, 0, blk);
      for (int c = 0; c < 18; ++c) {
        const int64_t gc = dof_h[es + n_el * c];
        if (gc < 0) continue;
        for (int r = 0; r < 18; ++r)
          if (dof_h[et + n_el * r] >= 0) R[18 * (et - t0) + r + nr * gc] += blk[r + 18 * c];
      }
    }
  }
}

/* ---------------------------------------------------------------------- */
/* right-hand side (system_assembly.cpp:222-274)                            */
/* ---------------------------------------------------------------------- */

void hfpo_rhs(const double* nods, int64_t nods_rows, int64_t n_nodes,
              const int64_t* conn, int64_t conn_rows, int64_t n_el,
              const int64_t* dof_h, const double* s_inf, double* rhs) {
  (void)n_nodes;
  for (int64_t e = 0; e < n_el; ++e) {
    double v[9], rt[9];
    el_vert(nods, nods_rows, conn, conn_rows, e, v);
    el_rt(v, rt);
    double n[3] = {-rt[2], -rt[5], -rt[8]};
    /* s_inf = (xx, yy, zz, xy, xz, yz) */
    double tr[3] = {n[0] * s_inf[0] + n[1] * s_inf[3] + n[2] * s_inf[4],
                    n[0] * s_inf[3] + n[1] * s_inf[1] + n[2] * s_inf[5],
                    n[0] * s_inf[4] + n[1] * s_inf[5] + n[2] * s_inf[2]};
    for (int ln = 0; ln < 6; ++ln)
      for (int l = 0; l < 3; ++l) {
        int64_t d = dof_h[e + n_el * (3 * ln + l)];
        if (d >= 0) rhs[d] -= tr[l];
      }
  }
}

void hfpo_write_dd(const double* x, const int64_t* dof_h, int64_t n_el, double* dd) {
  const int64_t nr = 6 * n_el;
  for (int64_t e = 0; e < n_el; ++e)
    for (int en = 0; en < 6; ++en)
      for (int i = 0; i < 3; ++i) {
        int64_t d = dof_h[e + n_el * (3 * en + i)];
        dd[6 * e + en + nr * i] = d >= 0 ? x[d] : 0.0;
      }
}

/* ---------------------------------------------------------------------- */
/* stress postprocessing (system_assembly.cpp:279-455)                      */
/* ---------------------------------------------------------------------- */

void hfpo_stress(const double* nods, int64_t nods_rows, int64_t n_nodes,
                 const int64_t* conn, int64_t conn_rows, int64_t n_el,
                 double shear_m, double poiss_r, const double* dd,
                 const double* pts, int64_t n_pts, const double* s_inf,
                 int is_dd_local, int nthreads, double* out) {
  hfpo_stress_w(nods, nods_rows, n_nodes, conn, conn_rows, n_el, shear_m, poiss_r, dd, pts, n_pts,
                s_inf, NULL, is_dd_local, nthreads, out);
}

void hfpo_stress_w(const double* nods, int64_t nods_rows, int64_t n_nodes,
                   const int64_t* conn, int64_t conn_rows, int64_t n_el,
                   double shear_m, double poiss_r, const double* dd,
                   const double* pts, int64_t n_pts, const double* s_inf,
                   const double* node_wts, int is_dd_local, int nthreads, double* out) {
  (void)n_nodes;
  const int64_t nr = 6 * n_el;
  double* geo = (double*)malloc(sizeof(double) * 96 * (size_t)n_el);
  for (int64_t e = 0; e < n_el; ++e) {
    double v[9], cp[18], wb[3];
    el_vert(nods, nods_rows, conn, conn_rows, e, v);
    double* g = geo + 96 * e; /* rt 9 | tau 6 | sfm 72 | v0 3 | pad */
    hfpo_el_geom_w(v, el_wts(node_wts, conn, conn_rows, e, wb), 0.0, g, g + 9, g + 15, cp);
    g[87] = v[0]; g[88] = v[1]; g[89] = v[2];
  }
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 4)
#endif
  for (int64_t p = 0; p < n_pts; ++p) {
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int64_t e = 0; e < n_el; ++e) {
      const double* g = geo + 96 * e;
      const double* rs = g;
      double y[3];
      to_local(rs, g + 87, pts + 3 * p, y);
      double S[108];
      hfpo_local_submatrix(shear_m, poiss_r, -y[2], y[0], y[1], g + 9, g + 15, S);
      /* DD of the element in its local frame: d_loc = R d_glob, or the
       * given DD as they are when they are local (system_assembly.cpp:358-387) */
      for (int ns = 0; ns < 6; ++ns) {
        double dg[3] = {dd[6 * e + ns], dd[6 * e + ns + nr], dd[6 * e + ns + 2 * nr]};
        double dl[3];
        for (int i = 0; i < 3; ++i)
          dl[i] = is_dd_local ? dg[i] : rs[i] * dg[0] + rs[i + 3] * dg[1] + rs[i + 6] * dg[2];
        /* local stress, then R^T S R (tensor_utilities.cpp:57-91) */
        double sl[6] = {0, 0, 0, 0, 0, 0};
        for (int k = 0; k < 3; ++k)
          for (int c = 0; c < 6; ++c) sl[c] += S[c + 6 * (3 * ns + k)] * dl[k];
        double m[3][3] = {{sl[0], sl[3], sl[4]}, {sl[3], sl[1], sl[5]}, {sl[4], sl[5], sl[2]}};
        double tg[3][3];
        for (int i = 0; i < 3; ++i)
          for (int jj = 0; jj < 3; ++jj) {
            double v = 0;
            for (int a = 0; a < 3; ++a)
              for (int b = 0; b < 3; ++b) v += rs[a + 3 * i] * m[a][b] * rs[b + 3 * jj];
            tg[i][jj] = v;
          }
        acc[0] += tg[0][0]; acc[1] += tg[1][1]; acc[2] += tg[2][2];
        acc[3] += tg[0][1]; acc[4] += tg[0][2]; acc[5] += tg[1][2];
      }
    }
    for (int c = 0; c < 6; ++c) out[p + n_pts * c] = acc[c] + s_inf[c];
  }
  free(geo);
}

/* ---------------------------------------------------------------------- */
/* LU with partial pivoting (LAPACK dgetrf / dgetrs semantics)              */
/* ---------------------------------------------------------------------- */

int64_t hfpo_getrf(int64_t n, double* A, int64_t lda, int32_t* ipiv, int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  const int64_t nb = 64;
  int64_t info = 0;
  for (int64_t k = 0; k < n; k += nb) {
    int64_t kb = (n - k < nb) ? n - k : nb;
    /* unblocked panel A[k:n, k:k+kb] */
    for (int64_t j = k; j < k + kb; ++j) {
      int64_t p = j;
      double mx = fabs(A[j + lda * j]);
      for (int64_t i = j + 1; i < n; ++i) {
        double v = fabs(A[i + lda * j]);
        if (v > mx) { mx = v; p = i; }
      }
      ipiv[j] = (int32_t)(p + 1);
      if (A[p + lda * j] != 0.0) {
        if (p != j)
          for (int64_t c = 0; c < n; ++c) {
            double t = A[j + lda * c];
            A[j + lda * c] = A[p + lda * c];
            A[p + lda * c] = t;
          }
        double inv = 1.0 / A[j + lda * j];
        for (int64_t i = j + 1; i < n; ++i) A[i + lda * j] *= inv;
      } else if (info == 0) {
        info = j + 1;
      }
      for (int64_t c = j + 1; c < k + kb; ++c) {
        double u = A[j + lda * c];
        if (u != 0.0)
          for (int64_t i = j + 1; i < n; ++i) A[i + lda * c] -= A[i + lda * j] * u;
      }
    }
    if (k + kb >= n) break;
    /* U12 = L11^-1 A12 */
#ifdef _OPENMP
#pragma omp parallel for
#endif
    for (int64_t c = k + kb; c < n; ++c)
      for (int64_t j = k; j < k + kb; ++j) {
        double u = A[j + lda * c];
        for (int64_t i = j + 1; i < k + kb; ++i) A[i + lda * c] -= A[i + lda * j] * u;
      }
    /* A22 -= L21 U12 */
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int64_t c = k + kb; c < n; ++c)
      for (int64_t j = k; j < k + kb; ++j) {
        double u = A[j + lda * c];
        if (u == 0.0) continue;
        const double* l = A + lda * j;
        double* col = A + lda * c;
        for (int64_t i = k + kb; i < n; ++i) col[i] -= l[i] * u;
      }
  }
  return info;
}

void hfpo_getrs(int64_t n, int64_t nrhs, const double* LU, int64_t lda,
                const int32_t* ipiv, double* B, int64_t ldb) {
  for (int64_t r = 0; r < nrhs; ++r) {
    double* b = B + ldb * r;
    for (int64_t i = 0; i < n; ++i) {
      int64_t p = ipiv[i] - 1;
      if (p != i) { double t = b[i]; b[i] = b[p]; b[p] = t; }
    }
    for (int64_t j = 0; j < n; ++j) {
      double v = b[j];
      if (v != 0.0)
        for (int64_t i = j + 1; i < n; ++i) b[i] -= LU[i + lda * j] * v;
    }
    for (int64_t j = n - 1; j >= 0; --j) {
      b[j] /= LU[j + lda * j];
      double v = b[j];
      if (v != 0.0)
        for (int64_t i = 0; i < j; ++i) b[i] -= LU[i + lda * j] * v;
    }
  }
}

/* ---------------------------------------------------------------------- */
/* Sneddon's penny crack opening (test_radial_crack_static.cpp:316-337)     */
/* ---------------------------------------------------------------------- */

void hfpo_sneddon_dd(double shear_m, double poiss_r, double a, double p,
                     const double* xyz, double* dd) {
  double r2 = xyz[0] * xyz[0] + xyz[1] * xyz[1];
  dd[0] = dd[1] = dd[2] = 0.0;
  if (r2 < a * a)
    dd[2] = 4.0 * (1.0 - poiss_r) * a * p / (kPi * shear_m) * sqrt(1.0 - r2 / (a * a));
}

/* ---------------------------------------------------------------------- */
/* volume control (system_assembly.cpp:458-629, element_utilities.cpp:484-595) */
/* ---------------------------------------------------------------------- */

/* Area integrals of the six shape functions over the element: the
 * reference's el_p2_sf_integral (element_utilities.cpp:573-595) sums, over
 * the three edges a -> b, the Cauchy-Pompeiu contour form of the monomial
 * integrals (el_p2_cbp_integral, :484-510) and maps them through sfm.  Here
 * the same contour form is evaluated by Gauss-Legendre quadrature instead of
 * the reference's closed-form edge polynomials: with dG/dtb = t^p tb^q,
 * G = t^p tb^(q+1) / (q+1), the area integral is (1 / 2i) sum over edges of
 * int G dt, a polynomial of degree <= 3 in the edge parameter (3-point rule
 * exact).  Monomial order [1, t, tb, t^2, tb^2, t*tb]. */
void hfpo_sf_integral(const double* tau, const double* sfm, double* out) {
  static const int P[6] = {0, 1, 0, 2, 0, 1}, Q[6] = {0, 0, 1, 0, 2, 1};
  const double gx[3] = {0.5 - 0.5 * sqrt(0.6), 0.5, 0.5 + 0.5 * sqrt(0.6)};
  const double gw[3] = {5.0 / 18.0, 8.0 / 18.0, 5.0 / 18.0};
  cplx t[3], m[6] = {0, 0, 0, 0, 0, 0};
  for (int j = 0; j < 3; ++j) t[j] = tau[2 * j] + I * tau[2 * j + 1];
  for (int e = 0; e < 3; ++e) {
    const cplx a = t[e], b = t[(e + 1) % 3];
    for (int g = 0; g < 3; ++g) {
      const cplx z = a + gx[g] * (b - a), zb = conj(z);
      for (int k = 0; k < 6; ++k)
        m[k] += gw[g] * cpow(z, P[k]) * cpow(zb, Q[k] + 1) / (Q[k] + 1.0) * (b - a) / (2.0 * I);
    }
  }
  for (int r = 0; r < 6; ++r) {
    cplx s = 0;
    for (int c = 0; c < 6; ++c) s += (sfm[2 * (r + 6 * c)] + I * sfm[2 * (r + 6 * c) + 1]) * m[c];
    out[r] = creal(s);
  }
}

/* The border that make_3dbem_matrix_vc (system_assembly.cpp:497-553) adds to
 * the n_dof x n_dof matrix: for each free slot (element e, node n_s,
 * component j) with DoF d,
 *   row[d] = (int N_{n_s} dA) * n_j   (volume vs DD; n = row 2 of the
 *                                      element's rotation tensor)
 *   col[d] = -n_j                     (traction vs pressure),
 * written in element order (the last writer wins for a DoF shared between
 * slots).  is_dd_local: the row holds the opening component only. */
void hfpo_vc_border(const double* nods, int64_t nods_rows, int64_t n_nodes,
                    const int64_t* conn, int64_t conn_rows, int64_t n_el,
                    const int64_t* dof_h, int is_dd_local, double* row, double* col) {
  hfpo_vc_border_w(nods, nods_rows, n_nodes, conn, conn_rows, n_el, dof_h, NULL, is_dd_local, row,
                   col);
}

void hfpo_vc_border_w(const double* nods, int64_t nods_rows, int64_t n_nodes,
                      const int64_t* conn, int64_t conn_rows, int64_t n_el,
                      const int64_t* dof_h, const double* node_wts, int is_dd_local,
                      double* row, double* col) {
  (void)n_nodes;
  for (int64_t e = 0; e < n_el; ++e) {
    double v[9], rt[9], tau[6], sfm[72], cp[18], sf[6], wb[3];
    el_vert(nods, nods_rows, conn, conn_rows, e, v);
    hfpo_el_geom_w(v, el_wts(node_wts, conn, conn_rows, e, wb), 0.0, rt, tau, sfm, cp);
    hfpo_sf_integral(tau, sfm, sf);
    for (int ns = 0; ns < 6; ++ns)
      for (int j = 0; j < 3; ++j) {
        const int64_t d = dof_h[e + n_el * (3 * ns + j)];
        if (d < 0) continue;
        /* local DD: the opening is component 2 (system_assembly.cpp:529-536) */
        row[d] = is_dd_local ? (j == 2 ? sf[ns] : 0.0) : sf[ns] * rt[2 + 3 * j];
        col[d] = -rt[2 + 3 * j];
      }
  }
}

/* Truncated volume-control system, mod_3dbem_system_vc
 * (system_assembly.cpp:560-629).  A: the (orig_ndof + 1)^2 volume-control
 * matrix (lda), its border at row/column orig_ndof (the reference indexes it
 * with n_el * ndpe, the same index whenever the original handle has no fixed
 * DoF, the only case in which the reference stays in bounds).  For used
 * slots (e, j) (dof_h >= 0) in element order:
 *   M(t, s)    += A(o_t, o_s)      over every pair of used slots
 *   M(used, s) += A(orig_ndof, o_s),  M(s, used) += A(o_s, orig_ndof)
 *   rhs[s] = delta_t[f] with f = e * ndpe + j if tsize == n_el * ndpe,
 *            o_s if tsize == orig_ndof, s otherwise;  rhs[used] = delta_v.
 * M ((used+1)^2, ld used+1) and rhs (used+1) are zero-filled here.  Returns
 * 0, or -1 when a used slot is fixed in the original handle or tsize fits
 * none of the three lengths. */
int hfpo_mod_system_vc(const double* A, int64_t lda, int64_t orig_ndof,
                       const int64_t* orig_dof_h, const int64_t* dof_h,
                       int64_t n_el, int64_t ndpe, int64_t used_ndof,
                       const double* delta_t, int64_t tsize, double delta_v,
                       double* M, double* rhs) {
  const int64_t full = n_el * ndpe, ldm = used_ndof + 1;
  if (tsize != full && tsize != orig_ndof && tsize != used_ndof) return -1;
  for (int64_t k = 0; k < n_el * ndpe; ++k)
    if (dof_h[k] >= 0 && orig_dof_h[k] < 0) return -1;
  memset(M, 0, sizeof(double) * (size_t)(ldm * ldm));
  memset(rhs, 0, sizeof(double) * (size_t)ldm);
  rhs[used_ndof] = delta_v;
  for (int64_t se = 0; se < n_el; ++se)
    for (int64_t j = 0; j < ndpe; ++j) {
      const int64_t s = dof_h[se + n_el * j];
      if (s < 0) continue;
      const int64_t os = orig_dof_h[se + n_el * j];
      for (int64_t te = 0; te < n_el; ++te)
        for (int64_t k = 0; k < ndpe; ++k) {
          const int64_t t = dof_h[te + n_el * k];
          if (t < 0) continue;
          M[t + ldm * s] += A[orig_dof_h[te + n_el * k] + lda * os];
        }
      M[used_ndof + ldm * s] += A[orig_ndof + lda * os];
      M[s + ldm * used_ndof] += A[os + lda * orig_ndof];
      if (tsize == full)
        rhs[s] = delta_t[se * ndpe + j];
      else if (tsize == orig_ndof)
        rhs[s] = delta_t[os];
      else
        rhs[s] = delta_t[s];
    }
  return 0;
}

/* ---------------------------------------------------------------------- */
/* cohesion / friction state (src/Development/cohesion_friction.cpp)        */
/* ---------------------------------------------------------------------- */

/* match_f_c (cohesion_friction.cpp:22-127), node by node (the reference
 * shares its nodal_dd scratch between OpenMP threads; this is its
 * one-thread behaviour).  dd: n_nod x 3 col-major (ldd), used as given (the
 * reference's rotation to local coordinates is an empty branch).  par:
 * 6 doubles per material (cr_open, peak_ts, cr_slip, peak_sc, peak_sf,
 * res_sf); cmid: n_mat x 2 col-major model ids (cohesion, friction). */
void hfpo_match_f_c(int64_t n_nod, const double* dd, int64_t ldd, const int32_t* mat_id,
                    const int32_t* cmid, int64_t n_mat, const double* par,
                    const double* mr_open, const double* mr_slip, double* friction_coef,
                    double* slip_cohesion, double* open_cohesion) {
  for (int64_t n = 0; n < n_nod; ++n) {
    const int m = mat_id[n];
    const int c_model = cmid[m], f_model = cmid[m + n_mat];
    const double* p = par + 6 * m;
    const double cr_open = p[0], peak_ts = p[1], cr_slip = p[2], peak_sc = p[3], peak_sf = p[4],
                 res_sf = p[5];
    const double d0 = dd[n], d1 = dd[n + ldd], d2 = dd[n + 2 * ldd];
    const double open = d2 / cr_open;                           /* opening / critical */
    const double slip = sqrt(d0 * d0 + d1 * d1) / cr_slip;      /* slip / critical */
    /* cohesion factor: Dugdale box (model 0) or linear softening (model 1)
     * while partly open; 0 otherwise (model 2 and unknown ids) */
    double f = 0.0;
    if (open > 0.0 && open < 1.0) f = c_model == 0 ? 1.0 : (c_model == 1 ? 1.0 - open : 0.0);
    /* unloading below the largest opening reached: scaled linearly */
    if (open > 0.0 && open < mr_open[n]) f *= open / mr_open[n];
    open_cohesion[n] = peak_ts * f;
    slip_cohesion[n] = peak_sc * f;
    /* friction: none once fully open; residual by default, peak without
     * weakening (0), linear slip weakening from the largest slip (1) */
    double mu = 0.0;
    if (open < 1.0) {
      mu = res_sf;
      if (f_model == 0) {
        mu = peak_sf;
      } else if (f_model == 1 && slip < 1.0) {
        const double s = slip < mr_slip[n] ? mr_slip[n] : slip;
        mu = res_sf + (peak_sf - res_sf) * (1.0 - s);
      }
    }
    friction_coef[n] = mu;
  }
}

/* update_mrd (cohesion_friction.cpp:131-189): largest relative opening and
 * slip reached.  A full opening sets 1 unless the opening itself is larger
 * than the previous maximum, which then wins (the reference's order). */
void hfpo_update_mrd(int64_t n_nod, const double* dd, int64_t ldd, const int32_t* mat_id,
                     int64_t n_mat, const double* par, double* mr_open, double* mr_slip) {
  (void)n_mat;
  for (int64_t n = 0; n < n_nod; ++n) {
    const double* p = par + 6 * mat_id[n];
    const double d0 = dd[n], d1 = dd[n + ldd], d2 = dd[n + 2 * ldd];
    const double open = d2 / p[0];
    const double slip = sqrt(d0 * d0 + d1 * d1) / p[2];
    const double pr_open = mr_open[n], pr_slip = mr_slip[n];
    if (slip > pr_slip) mr_slip[n] = slip;
    if (open >= 1.0) mr_open[n] = 1.0;
    if (open > pr_open) mr_open[n] = open;
  }
}

/* ---------------------------------------------------------------------- */
/* one step of the volume-control cohesion / friction iteration             */
/* (vc_cf_iteration, src/Development/c_f_iteration.cpp:27-436)             */
/* ---------------------------------------------------------------------- */

/* Shape-function values of the uniform 6-node element at its six
 * collocation points (set_ele_struct's sf_cp): sf[q][j] = N_j(cp_q), cp_q at
 * barycentric (1 - beta) L(node q) + beta / 3. */
static void cf_sf_cp(double beta, double sf[6][6]) {
  for (int q = 0; q < 6; ++q) {
    double L[3];
    for (int k = 0; k < 3; ++k) {
      double ln;
      if (q < 3)
        ln = k == q ? 1.0 : 0.0;
      else
        ln = k == q - 3 ? 0.0 : 0.5;
      L[k] = (1.0 - beta) * ln + beta / 3.0;
    }
    for (int v = 0; v < 3; ++v) {
      sf[q][v] = L[v] * (2.0 * L[v] - 1.0);
      sf[q][3 + v] = 4.0 * L[(v + 1) % 3] * L[(v + 2) % 3];
    }
  }
}

/* y = R x (transpose = 0) or R^T x (transpose = 1); rt col-major, rows e1..e3 */
static void cf_rot(const double* rt, int transpose, const double* x, double* y) {
  double t[3];
  for (int i = 0; i < 3; ++i)
    t[i] = transpose ? rt[0 + 3 * i] * x[0] + rt[1 + 3 * i] * x[1] + rt[2 + 3 * i] * x[2]
                     : rt[i] * x[0] + rt[i + 3] * x[1] + rt[i + 6] * x[2];
  y[0] = t[0]; y[1] = t[1]; y[2] = t[2];
}

int hfpo_vc_cf_iteration(const double* nods, int64_t nods_rows, int64_t n_nodes, const int64_t* conn,
                         int64_t conn_rows, int64_t n_el, double beta, int is_dd_local,
                         const int64_t* orig_dof_h, int64_t orig_ndof, const double* A, int64_t lda,
                         const int32_t* mat_id, const int32_t* cmid, int64_t n_mat, const double* par,
                         const double* mr_open, const double* mr_slip, const double* s_inf, int64_t inj_el,
                         int64_t inj_node, double new_vol, double* dd, double* pp, int64_t* dof_h,
                         int64_t* n_dof, double* friction_coef, double* slip_cohesion, double* open_cohesion,
                         double* delta_trac, double* incr, double* pressure, double* residual) {
  (void)n_nodes;
  const int64_t n0 = orig_ndof, n_nod = 6 * n_el, full = 18 * n_el;
  if (n_el < 1 || n0 < 1 || lda < n0 + 1 || inj_el < 0 || inj_el >= n_el || inj_node < 0 || inj_node >= 6)
    return -1;
  for (int64_t k = 0; k < full; ++k)
    if (orig_dof_h[k] < -1 || orig_dof_h[k] >= n0 || (dof_h[k] >= 0 && orig_dof_h[k] < 0)) return -1;
  for (int64_t n = 0; n < n_nod; ++n)
    if (mat_id[n] < 0 || mat_id[n] >= n_mat) return -1;
  /* friction and cohesion of the current DD (c_f_iteration.cpp:88-91) */
  hfpo_match_f_c(n_nod, dd, n_nod, mat_id, cmid, n_mat, par, mr_open, mr_slip, friction_coef, slip_cohesion,
                 open_cohesion);
  /* DD vector of the DoF active in the current handle, at their original
   * numbers (get_dd_vector_from_md, surface_mesh_utilities.cpp:193-240, read
   * into the original numbering the matrix has), its tractions and the
   * current volume (c_f_iteration.cpp:65-77) */
  double* ddv = (double*)calloc((size_t)n0, sizeof(double));
  double* et = (double*)calloc((size_t)n0 + 1, sizeof(double));
  int64_t* act = (int64_t*)malloc(sizeof(int64_t) * (size_t)n0);
  for (int64_t e = 0; e < n_el; ++e)
    for (int l = 0; l < 18; ++l) {
      const int64_t o = orig_dof_h[e + n_el * l];
      if (o >= 0 && dof_h[e + n_el * l] >= 0) ddv[o] = dd[6 * e + l / 3 + n_nod * (l % 3)];
    }
  for (int64_t c = 0; c < n0; ++c) {
    const double x = ddv[c];
    if (x == 0.0) continue;
    for (int64_t r = 0; r <= n0; ++r) et[r] += A[r + lda * c] * x;
  }
  const double curr_vol = et[n0], delta_v = new_vol - curr_vol;
  const double p_old = pp[6 * inj_el + inj_node];
  double sf[6][6];
  cf_sf_cp(beta, sf);
  for (int64_t c = 0; c < n0; ++c) act[c] = 0;
  /* admissibility of the traction at each collocation node
   * (c_f_iteration.cpp:96-276) */
  for (int64_t e = 0; e < n_el; ++e) {
    double vert[9], rt[9];
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k) vert[k + 3 * j] = nods[k + nods_rows * conn[j + conn_rows * e]];
    el_rt(vert, rt);
    /* traction of the far field on the element (normal -e3), local frame */
    const double nv[3] = {-rt[2], -rt[5], -rt[8]};
    double ti[3] = {s_inf[0] * nv[0] + s_inf[3] * nv[1] + s_inf[4] * nv[2],
                    s_inf[3] * nv[0] + s_inf[1] * nv[1] + s_inf[5] * nv[2],
                    s_inf[4] * nv[0] + s_inf[5] * nv[1] + s_inf[2] * nv[2]};
    cf_rot(rt, 0, ti, ti);
    for (int q = 0; q < 6; ++q) {
      const int64_t g = 6 * e + q;
      double tr[3];
      for (int i = 0; i < 3; ++i) {
        const int64_t o = orig_dof_h[e + n_el * (3 * q + i)];
        tr[i] = o >= 0 ? et[o] : 0.0;
      }
      cf_rot(rt, 0, tr, tr);
      double pr = 0.0;
      for (int j = 0; j < 6; ++j) pr += pp[6 * e + j] * sf[q][j];
      const double nt = pr - ti[2] - tr[2];
      const double st[2] = {-ti[0] - tr[0], -ti[1] - tr[1]};
      const double sh = sqrt(st[0] * st[0] + st[1] * st[1]);
      const double adm_nt = open_cohesion[g];
      const double adm_st = slip_cohesion[g] - friction_coef[g] * nt;
      double dir[2] = {0.0, 0.0};
      if (sh > 0) { dir[0] = st[0] / sh; dir[1] = st[1] / sh; }
      double dt[3] = {0.0, 0.0, 0.0};
      int a[3] = {2, 2, 2}; /* 1 active, 0 released, 2 as in the current handle */
      if (nt > adm_nt) {
        /* full separation: all traction released */
        dt[0] = -st[0]; dt[1] = -st[1]; dt[2] = -nt;
        a[0] = a[1] = a[2] = 1;
      } else {
        if (mr_open[g] < 1.0 && nt > 0.0) {
          /* cohesive zone: normal traction to the cohesion */
          dt[2] = adm_nt - nt;
          a[2] = 1;
        }
        if (mr_slip[g] > 0 || sh > adm_st) {
          /* slip: shear traction to the friction bound */
          dt[0] = adm_st * dir[0] - st[0];
          dt[1] = adm_st * dir[1] - st[1];
          a[0] = a[1] = 1;
        } else if (nt <= 0.0) {
          /* intact (closed, no slip): node leaves the system */
          a[0] = a[1] = a[2] = 0;
        }
      }
      /* the adjustment as a right-hand side of the system, global frame:
       * -R^T dt, all three components (the reference adds +R^T dt for
       * components 0 and 1 only, c_f_iteration.cpp:279-285) */
      double gdt[3];
      cf_rot(rt, 1, dt, gdt);
      for (int i = 0; i < 3; ++i) {
        const int64_t l = 3 * q + i, o = orig_dof_h[e + n_el * l];
        if (o < 0) continue;
        delta_trac[o] = -gdt[i];
        act[o] = a[i] == 2 ? (dof_h[e + n_el * l] >= 0) : a[i];
      }
    }
  }
  /* new handle: the active original DoF renumbered in their original order
   * (the reference's dof - dd_dof_dec where that is well defined) */
  int64_t* num = (int64_t*)malloc(sizeof(int64_t) * (size_t)n0);
  int64_t nu = 0;
  for (int64_t c = 0; c < n0; ++c) num[c] = act[c] ? nu++ : -1;
  for (int64_t k = 0; k < full; ++k) dof_h[k] = orig_dof_h[k] >= 0 ? num[orig_dof_h[k]] : -1;
  *n_dof = nu;
  int rc = 0;
  for (int64_t t = 0; t <= nu; ++t) incr[t] = 0.0;
  if (nu > 0) {
    /* truncated system (mod_3dbem_system_vc) and its FP64 LU solve */
    const int64_t ldm = nu + 1;
    double* M = (double*)malloc(sizeof(double) * (size_t)(ldm * ldm));
    int32_t* piv = (int32_t*)malloc(sizeof(int32_t) * (size_t)ldm);
    hfpo_mod_system_vc(A, lda, n0, orig_dof_h, dof_h, n_el, 18, nu, delta_trac, n0, delta_v, M, incr);
    if (hfpo_getrf(ldm, M, ldm, piv, 1) != 0)
      rc = 2;
    else
      hfpo_getrs(ldm, 1, M, ldm, piv, incr, ldm);
    free(M);
    free(piv);
  } else {
    rc = 1;
  }
  if (rc == 2) {
    free(ddv); free(et); free(act); free(num);
    return 2;
  }
  const double p_new = p_old + incr[nu];
  /* DD, pressure update (c_f_iteration.cpp:310-428) */
  for (int64_t e = 0; e < n_el; ++e) {
    double vert[9], rt[9], del[6][3];
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k) vert[k + 3 * j] = nods[k + nods_rows * conn[j + conn_rows * e]];
    el_rt(vert, rt);
    for (int q = 0; q < 6; ++q) {
      const int64_t g = 6 * e + q;
      double dn[3];
      for (int i = 0; i < 3; ++i) {
        const int64_t l = 3 * q + i, o = orig_dof_h[e + n_el * l], d = dof_h[e + n_el * l];
        dn[i] = dd[g + n_nod * i] + ((o >= 0 && d >= 0) ? incr[d] : 0.0);
      }
      if (!is_dd_local) cf_rot(rt, 0, dn, dn);
      if (dn[2] < 0.0) dn[2] = 0.0; /* no overlap */
      for (int i = 0; i < 3; ++i) del[q][i] = dn[i];
      if (!is_dd_local) cf_rot(rt, 1, dn, dn);
      for (int i = 0; i < 3; ++i)
        if (orig_dof_h[e + n_el * (3 * q + i)] >= 0) dd[g + n_nod * i] = dn[i];
    }
    for (int q = 0; q < 6; ++q) {
      const int64_t g = 6 * e + q;
      /* opening at the collocation point, element frame (the reference
       * rotates this already-local value once more, c_f_iteration.cpp:404-409) */
      double op = 0.0;
      for (int j = 0; j < 6; ++j) op += del[j][2] * sf[q][j];
      double st = op / par[6 * mat_id[g]];
      if (st >= 1.0 || mr_open[g] >= 1.0) st = 1.0;
      if (st >= 1.0 || (st > 0.0 && mr_open[g] >= 1.0)) pp[g] = p_new;
    }
  }
  double res = 0.0;
  for (int64_t t = 0; t <= nu; ++t) res += fabs(incr[t]);
  for (int64_t c = 0; c < n0; ++c) res += fabs(delta_trac[c]);
  *pressure = p_new;
  *residual = res;
  free(ddv); free(et); free(act); free(num);
  return rc;
}
