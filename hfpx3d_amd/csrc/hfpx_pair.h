// hfpx_pair.h -- per (point, source element) influence of a 6-node quadratic
// displacement-discontinuity element, device side.
//
// Replaces hfp3d::make_local_3dbem_submatrix
// (src/Elasticity/elasticity_kernel_integration.cpp:34-205, coefficient
// tables src/Elasticity/h_potential.cpp:34-1233) followed by the traction /
// rotation steps of hfp3d::make_3dbem_matrix_s
// (src/Solvers/system_assembly.cpp:121-199) and of
// hfp3d::make_3dbem_stress_f_s (src/Solvers/system_assembly.cpp:334-417).
//
// Mathematics (derived for this code, not taken from the coefficient tables):
// in the source element's frame (element in the plane x3 = 0), the stress at
// X due to the DD component D_i distributed with density P(xi) is
//     K * sum_a c_a d^2 Phi[P] / dX^a  +  K * x3 * sum_b c_b d^3 Phi[P] / dX^b,
// K = G / (16 pi (1 - nu)), Phi[P](X) = int_T P(xi) / |X - xi| dA, with the
// small integer/nu coefficient table below (obtained symbolically from
// Kelvin's solution).  Integration by parts in the plane turns each needed
// derivative of Phi into line integrals over the three edges of
//     M[j][l] = int a^j (a^2 + c^2)^(-l-1/2) da,   j <= 4, l <= 2,
// where a is the arc coordinate measured from the foot of the perpendicular
// from the projected point, c^2 = s^2 + x3^2 and s is the signed distance of
// the projected point to the edge line, plus two scalar potentials: the
// solid-angle term W1 = dPhi[1]/dx3 and Phi[1].  Complex notation
// (w = xi - z) groups the twelve independent derivatives into
//     A = Phi_11 + Phi_22,  B = Phi_11 - Phi_22 + 2i Phi_12,  C = Phi_13 + i Phi_23,
//     D = (Phi_111 + Phi_122) + i (Phi_112 + Phi_222),
//     F = (Phi_111 - 3 Phi_122) + i (3 Phi_112 - Phi_222),
//     G = Phi_113 + Phi_223,  H = Phi_113 - Phi_223 + 2i Phi_123.
// The in-plane hypersingular case (x3 = 0) is the limit of the same formulas
// with the solid-angle term at its principal value (zero), which is the
// finite-part value of the reference's in-plane branch.
#pragma once
#include <hip/hip_runtime.h>

namespace hfpx {

template <typename T>
struct cx {
  T re, im;
};
template <typename T>
__device__ __forceinline__ cx<T> mk(T r, T i) { return cx<T>{r, i}; }
template <typename T>
__device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <typename T>
__device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <typename T>
__device__ __forceinline__ cx<T> operator*(cx<T> a, cx<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T>
__device__ __forceinline__ cx<T> operator*(T s, cx<T> a) { return {s * a.re, s * a.im}; }
template <typename T>
__device__ __forceinline__ cx<T> conj(cx<T> a) { return {a.re, -a.im}; }
// i * a
template <typename T>
__device__ __forceinline__ cx<T> mul_i(cx<T> a) { return {-a.im, a.re}; }

// Source-element data, identical for every point the element acts on.
template <typename T>
struct SrcElem {
  T R[9];      // rows e1, e2, e3 of the element frame (row-major)
  T v0[3];     // vertex 0 (frame origin)
  T t[6];      // vertex coordinates in the frame (re, im) x 3
  T th[6];     // unit edge tangents (re, im) x 3, edge m from vertex m to m+1
  T len[3];    // edge lengths
  T l0[3];     // barycentric L_k(w) = l0_k + 2 Re(l1_k w)
  T l1[6];     // (re, im) x 3
  // Shape functions of a non-uniform edge partition (vertex weights w):
  //   N_v = L_v^2 - sw[2v] L_v L_{v+1} - sw[2v+1] L_v L_{v+2},
  //   sw[2v] = w_v / w_{v+1}, sw[2v+1] = w_v / w_{v+2};
  //   N_{3+v} = sw[6+v] L_{v+1} L_{v+2}, sw[6+v] = (w_a + w_b)^2 / (w_a w_b).
  // Equal weights give 1 and 4, the uniform element; only the NU variant of
  // pair_influence reads them.
  T sw[9];
};

// Edge moments M[j][l] (only the entries used below), differences of the
// two end values in forms that keep relative accuracy for small c and for
// ends of equal sign.
template <typename T>
struct EdgeMom {
  T m00, m10, m01, m11, m21, m31, m02, m12, m22, m32, m42;
};

template <typename T>
__device__ __forceinline__ void edge_moments(T am, T an, T c2, EdgeMom<T>& M) {
  const T rm = sqrt(am * am + c2), rn = sqrt(an * an + c2);
  T lam, f01, f02;
  if (am >= T(0)) {
    const T pm = rm + am, pn = rn + an;
    lam = log(pn / pm);
    const T gm1 = -T(1) / (rm * pm), gn1 = -T(1) / (rn * pn);
    const T gm2 = -(T(2) * rm + am) / (T(3) * rm * rm * rm * pm * pm);
    const T gn2 = -(T(2) * rn + an) / (T(3) * rn * rn * rn * pn * pn);
    f01 = gn1 - gm1;
    f02 = gn2 - gm2;
  } else if (an <= T(0)) {
    const T qm = rm - am, qn = rn - an;
    lam = log(qm / qn);
    const T gm1 = T(1) / (rm * qm), gn1 = T(1) / (rn * qn);
    const T gm2 = (T(2) * rm - am) / (T(3) * rm * rm * rm * qm * qm);
    const T gn2 = (T(2) * rn - an) / (T(3) * rn * rn * rn * qn * qn);
    f01 = gn1 - gm1;
    f02 = gn2 - gm2;
  } else {
    const T qm = rm - am, pn = rn + an;
    lam = log(pn * qm / c2);
    const T gm1 = T(1) / (rm * qm), gn1 = -T(1) / (rn * pn);
    const T gm2 = (T(2) * rm - am) / (T(3) * rm * rm * rm * qm * qm);
    const T gn2 = -(T(2) * rn + an) / (T(3) * rn * rn * rn * pn * pn);
    const T ic2 = T(1) / c2;
    f01 = T(2) * ic2 + gn1 - gm1;
    f02 = (T(4) / T(3)) * ic2 * ic2 + gn2 - gm2;
  }
  const T dr = (an - am) * (an + am) / (rn + rm);
  const T irr = T(1) / (rn * rm);
  const T dinv = -dr * irr;
  const T dinv3 = -dr * (rn * rn + rn * rm + rm * rm) * irr * irr * irr;
  M.m00 = lam;
  M.m10 = dr;
  M.m01 = f01;
  M.m11 = -dinv;
  M.m21 = lam - c2 * f01;
  M.m31 = dr + c2 * dinv;
  M.m02 = f02;
  M.m12 = -dinv3 * (T(1) / T(3));
  M.m22 = f01 - c2 * f02;
  M.m32 = -dinv + c2 * dinv3 * (T(1) / T(3));
  M.m42 = lam - T(2) * c2 * f01 + c2 * c2 * f02;
}

template <typename T>
__device__ __forceinline__ T solid_end(T a, T s, T ax3) {
  const T r = sqrt(a * a + s * s + ax3 * ax3);
  return atan(a * s * (a * a + s * s) / ((r + ax3) * (s * s * r + ax3 * a * a)));
}

// Adds s L_a L_b to the re-centred density P = p00 + 2 Re(p10 w) +
// 2 Re(p20 w^2) + p11 |w|^2, with L_k(w) = lam_k + 2 Re(g_k w).
template <typename T>
__device__ __forceinline__ void add_bary_product(T s, T la, cx<T> ga, T lb, cx<T> gb, T& p00,
                                                 cx<T>& p10, cx<T>& p20, T& p11) {
  p00 += s * la * lb;
  p10 = p10 + s * (la * gb + lb * ga);
  p20 = p20 + s * (ga * gb);
  p11 += T(2) * s * (ga.re * gb.re + ga.im * gb.im);
}

// Stress influence at a point given in the source frame (z = y0 + i y1,
// x3 = y2).  Calls sink(node, S_local[6][3]) once per node with
// S_local[c][k] = stress component c (S11 S22 S33 S12 S13 S23) due to unit
// DD component k (source frame) at the node.
// This core takes any quadratic density: dens(j, p00, p10, p20, p11) gives
// shape function j re-centred at the projected point z, as
// P = p00 + 2 Re(p10 w) + 2 Re(p20 w^2) + p11 |w|^2 (w = xi - z).  It reads
// only the edge data of E (t, th, len).
template <typename T, typename Dens, typename Sink>
__device__ __forceinline__ void pair_influence_dens(const SrcElem<T>& E, T y0, T y1, T x3, T shear_m, T nu,
                                                    Dens&& dens, Sink&& sink) {
  const T ax3 = fabs(x3);
  const T sgn3 = x3 > T(0) ? T(1) : (x3 < T(0) ? T(-1) : T(0));
  EdgeMom<T> em[3];
  T es[3];
  cx<T> enu[3];
  T w1 = T(0), phi1 = T(0);
  cx<T> enu_r = mk<T>(0, 0), enub_r = mk<T>(0, 0);
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    const cx<T> th = mk<T>(E.th[2 * m], E.th[2 * m + 1]);
    const cx<T> nv = mk<T>(th.im, -th.re);  // -i * th
    const T dx = E.t[2 * m] - y0, dy = E.t[2 * m + 1] - y1;
    const T s = dx * nv.re + dy * nv.im;
    const T am = dx * th.re + dy * th.im;
    const T an = am + E.len[m];
    const T c2 = s * s + x3 * x3;
    edge_moments<T>(am, an, c2, em[m]);
    es[m] = s;
    enu[m] = nv;
    if (s != T(0)) w1 -= sgn3 * (solid_end<T>(an, s, ax3) - solid_end<T>(am, s, ax3));
    phi1 += s * em[m].m00;
    enu_r = enu_r + em[m].m00 * nv;
    enub_r = enub_r + em[m].m00 * conj(nv);
  }
  phi1 += x3 * w1;

  const T K = shear_m / (T(16) * T(3.14159265358979323846) * (T(1) - nu));
#pragma unroll 1
  for (int j = 0; j < 6; ++j) {
    // density re-centred at z: P = p00 + 2 Re(p10 w) + 2 Re(p20 w^2) + p11 |w|^2
    T p00, p11;
    cx<T> p10, p20;
    dens(j, p00, p10, p20, p11);
    const cx<T> p01 = conj(p10), p02 = conj(p20);
    T A = T(0), G = T(0);
    cx<T> B = mk<T>(0, 0), C = mk<T>(0, 0), D = mk<T>(0, 0), F = mk<T>(0, 0), H = mk<T>(0, 0);
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      const T s = es[m];
      const cx<T> nv = enu[m], nb = conj(nv), n2 = nv * nv, nb2 = conj(n2);
      const EdgeMom<T>& M = em[m];
      // real density along the edge: P(a) = P0 + P1 a + P2 a^2
      const cx<T> p10n = p10 * nv, p20n2 = p20 * n2;
      const T P0 = p00 + T(2) * s * p10n.re + T(2) * s * s * p20n2.re + p11 * s * s;
      const T P1 = -T(2) * p10n.im - T(4) * s * p20n2.im;
      const T P2 = -T(2) * p20n2.re + p11;
      // dP/d(wb) along the edge: Q0 + Q1 a
      const cx<T> Q0 = p01 + (T(2) * s) * (p02 * nb) + (p11 * s) * nv;
      const cx<T> Q1 = mul_i((T(-2)) * (p02 * nb) + p11 * nv);
      // real moment combinations of a^k P
      const T J10 = P0 * M.m01 + P1 * M.m11 + P2 * M.m21;
      const T J11 = P0 * M.m11 + P1 * M.m21 + P2 * M.m31;
      const T J20 = P0 * M.m02 + P1 * M.m12 + P2 * M.m22;
      const T J21 = P0 * M.m12 + P1 * M.m22 + P2 * M.m32;
      const T J22 = P0 * M.m22 + P1 * M.m32 + P2 * M.m42;
      const cx<T> IQ0 = M.m00 * Q0 + M.m10 * Q1;
      const cx<T> IQ1 = M.m01 * Q0 + M.m11 * Q1;
      const cx<T> IaQ1 = M.m11 * Q0 + M.m21 * Q1;
      // A += -s J10 + i J11 (imaginary part vanishes for real P) - 2 nb IQ0
      A += -s * J10 - T(2) * (nb * IQ0).re;
      B = B - (n2 * mk<T>(s * J10, J11)) - T(2) * (nv * IQ0);
      C = C + (x3 * J10) * nv;
      G += T(3) * x3 * s * J20 + T(2) * x3 * (nb * IQ1).re;
      H = H + (T(3) * x3) * (n2 * mk<T>(s * J20, J21)) + (T(2) * x3) * (nv * IQ1);
      const cx<T> smQ = s * IQ1 - mul_i(IaQ1);
      const cx<T> spQ = s * IQ1 + mul_i(IaQ1);
      const T m00 = M.m00;
      D = D + (T(2) * J10 - T(3) * (s * s * J20 + J22)) * nv - T(2) * smQ -
          (T(8) * m00) * (p02 * nb);
      F = F - T(3) * ((n2 * nv) * mk<T>(s * s * J20 - J22, T(2) * s * J21)) -
          T(2) * (n2 * spQ) - (T(8) * m00) * (p02 * nv);
    }
    A += T(4) * p11 * phi1;
    B = B + (T(8) * phi1) * p02;
    C = C + T(2) * (w1 * p01 + x3 * (p11 * enu_r + T(2) * (p02 * enub_r)));
    G += T(4) * p11 * w1;
    H = H + (T(8) * w1) * p02;

    // stress combinations S0 = S11+S22, S2 = S11-S22+2i S12, S1 = S13+i S23
    T S[6][3];
    {
      // DD component 1
      const T S0 = K * (-T(8) * (T(1) + nu) * C.re + T(4) * x3 * D.re);
      const cx<T> S2 = K * ((-T(8) * (T(1) - nu)) * C + (T(2) * x3) * (D + F));
      const T S33 = -T(4) * K * x3 * D.re;
      const cx<T> S1 = K * mk<T>(T(2) * (T(2) - nu) * A + T(2) * nu * B.re + T(2) * x3 * (G + H.re),
                                 T(2) * nu * B.im + T(2) * x3 * H.im);
      S[0][0] = T(0.5) * (S0 + S2.re);
      S[1][0] = T(0.5) * (S0 - S2.re);
      S[2][0] = S33;
      S[3][0] = T(0.5) * S2.im;
      S[4][0] = S1.re;
      S[5][0] = S1.im;
    }
    {
      // DD component 2
      const T S0 = K * (-T(8) * (T(1) + nu) * C.im + T(4) * x3 * D.im);
      const cx<T> S2 = K * mul_i((-T(8) * (T(1) - nu)) * C + (T(2) * x3) * (D - F));
      const T S33 = -T(4) * K * x3 * D.im;
      // i K (2(2-nu) A - 2 nu B + 2 x3 (G - H))
      const cx<T> in = mk<T>(T(2) * (T(2) - nu) * A - T(2) * nu * B.re + T(2) * x3 * (G - H.re),
                             -T(2) * nu * B.im - T(2) * x3 * H.im);
      const cx<T> S1 = K * mul_i(in);
      S[0][1] = T(0.5) * (S0 + S2.re);
      S[1][1] = T(0.5) * (S0 - S2.re);
      S[2][1] = S33;
      S[3][1] = T(0.5) * S2.im;
      S[4][1] = S1.re;
      S[5][1] = S1.im;
    }
    {
      // DD component 3 (normal)
      const T S0 = T(4) * K * ((T(1) + T(2) * nu) * A + x3 * G);
      const cx<T> S2 = (T(4) * K) * ((T(1) - T(2) * nu) * B + x3 * H);
      const T S33 = T(4) * K * (A - x3 * G);
      const cx<T> S1 = (-T(4) * K * x3) * D;
      S[0][2] = T(0.5) * (S0 + S2.re);
      S[1][2] = T(0.5) * (S0 - S2.re);
      S[2][2] = S33;
      S[3][2] = T(0.5) * S2.im;
      S[4][2] = S1.re;
      S[5][2] = S1.im;
    }
    sink(j, S);
  }
}

// The 6-node quadratic element of E: uniform (mid-edge nodes) or, with NU,
// the non-uniform edge partition of SrcElem::sw, through the barycentric
// coordinates L_k(w) = l0_k + 2 Re(l1_k w) at the projected point.
template <typename T, bool NU = false, typename Sink>
__device__ __forceinline__ void pair_influence(const SrcElem<T>& E, T y0, T y1, T x3,
                                               T shear_m, T nu, Sink&& sink) {
  auto dens = [&](int j, T& p00, cx<T>& p10, cx<T>& p20, T& p11) {
    // barycentric coordinates at the projected point (loop-invariant: the
    // compiler evaluates them once, after the edge terms)
    T lam[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) lam[k] = E.l0[k] + T(2) * (E.l1[2 * k] * y0 - E.l1[2 * k + 1] * y1);
    if constexpr (NU) {
      p00 = T(0);
      p11 = T(0);
      p10 = mk<T>(0, 0);
      p20 = mk<T>(0, 0);
      const int v = j < 3 ? j : j - 3;
      const int a = (v + 1) % 3, b = (v + 2) % 3;
      const cx<T> gv = mk<T>(E.l1[2 * v], E.l1[2 * v + 1]);
      const cx<T> ga = mk<T>(E.l1[2 * a], E.l1[2 * a + 1]);
      const cx<T> gb = mk<T>(E.l1[2 * b], E.l1[2 * b + 1]);
      if (j < 3) {
        add_bary_product<T>(T(1), lam[v], gv, lam[v], gv, p00, p10, p20, p11);
        add_bary_product<T>(-E.sw[2 * v], lam[v], gv, lam[a], ga, p00, p10, p20, p11);
        add_bary_product<T>(-E.sw[2 * v + 1], lam[v], gv, lam[b], gb, p00, p10, p20, p11);
      } else {
        add_bary_product<T>(E.sw[6 + v], lam[a], ga, lam[b], gb, p00, p10, p20, p11);
      }
    } else if (j < 3) {
      const T l = lam[j];
      const cx<T> g = mk<T>(E.l1[2 * j], E.l1[2 * j + 1]);
      p00 = T(2) * l * l - l;
      p10 = (T(4) * l - T(1)) * g;
      p20 = T(2) * (g * g);
      p11 = T(4) * (g.re * g.re + g.im * g.im);
    } else {
      const int a = (j - 3 + 1) % 3, b = (j - 3 + 2) % 3;
      const cx<T> ga = mk<T>(E.l1[2 * a], E.l1[2 * a + 1]);
      const cx<T> gb = mk<T>(E.l1[2 * b], E.l1[2 * b + 1]);
      p00 = T(4) * lam[a] * lam[b];
      p10 = T(4) * (lam[a] * gb + lam[b] * ga);
      p20 = T(4) * (ga * gb);
      p11 = T(8) * (ga.re * gb.re + ga.im * gb.im);
    }
  };
  pair_influence_dens<T>(E, y0, y1, x3, shear_m, nu, dens, sink);
}

}  // namespace hfpx
