"""Deciding, pair by pair, which of two matrices is right where the oracle's
assembled matrix and the reference's differ (test infrastructure).

A differing entry comes from one (collocation point, source element) pair.
For each such pair this module gives
  * its geometry in the source element's frame: the distance c of the point
    to the nearest edge LINE and the position t of its foot along that edge
    (t outside [0, 1]: on the extension), both relative to the element size;
  * an independent value of the pair's 6 x 18 stress block by brute-force
    quadrature of Kelvin's tensor (kelvin_quad, float64, graded towards the
    element for close points);
  * the three traction rows of the block (S33, S13, S23 for a point in the
    element's plane or above it) in 30-digit arithmetic (mpmath), by
    Gauss-Legendre quadrature of the same kernel written out component by
    component, for points well away from the element where the rule
    converges to that precision.
"""
import numpy as np

import kelvin_quad
import oracle_lib

O = oracle_lib.oracle()

# rows of the 6 x 18 block that make the traction on a plane parallel to
# the element (S33, S13, S23): continuous across the element's plane
TRACTION_ROWS = [2, 4, 5]


def suspect_pairs(A, Ar, n_el, tol):
    """(te, k, se) with max over the pair's 3 x 18 entries of |A - Ar| above
    tol * max|Ar| (identity DoF map: row 18 te + 3 k + comp, col 18 se + ...),
    worst first, and the matching discrepancies."""
    mx = np.abs(Ar).max()
    P = np.abs(A - Ar).reshape(n_el, 6, 3, n_el, 18).max(axis=(2, 4)) / mx
    idx = np.argwhere(P > tol)
    d = P[tuple(idx.T)]
    order = np.argsort(-d)
    return idx[order], d[order]


def pair_frame(nods, conn, te, k, se, beta=0.125):
    """Source vertices, collocation point, and (h, z, tau, sfm) of the pair."""
    vs, vt = nods[:3, conn[:3, se]], nods[:3, conn[:3, te]]
    sfm, rt, tau, _ = O.el_geom(vs, beta)
    _, _, _, cpt = O.el_geom(vt, beta)
    x = cpt[:, k]
    y = rt @ (x - vs[:, 0])
    return vs, x, -y[2], y[0] + 1j * y[1], tau, sfm


def edge_line_position(tau, h, z):
    """(c / L, t, d / L): distance of the point to the nearest edge line,
    the foot's position along that edge (0, 1 = its vertices), and the
    distance to the nearest vertex, relative to the longest edge L."""
    L = max(abs(tau[(m + 1) % 3] - tau[m]) for m in range(3))
    best = None
    for m in range(3):
        a, b = tau[m], tau[(m + 1) % 3]
        th = (b - a) / abs(b - a)
        w = (z - a) * np.conj(th)
        c = np.hypot(w.imag, h)
        if best is None or c < best[0]:
            best = (c, w.real / abs(b - a))
    dv = min(np.hypot(abs(z - t), h) for t in tau)
    return best[0] / L, best[1], dv / L


def on_edge_extension(tau, h, z, max_angle=3e-2):
    """True when the point lies in the thin cone around the extension of a
    source edge line beyond one of its vertices (angle below max_angle
    radians seen from the vertex), the region where the reference's per-pair
    formulas lose accuracy.  Also returns (c / L, t, angle)."""
    c, t, _ = edge_line_position(tau, h, z)
    beyond = -t if t < 0 else t - 1.0
    if beyond <= 0:
        return False, (c, t, np.inf)
    ang = c / beyond
    return ang < max_angle, (c, t, ang)


def quad_block(vs, x, G, nu, d_rel):
    """Kelvin quadrature of the pair's block (float64): a plain collapsed
    Gauss rule away from the element, graded cells close to it."""
    if d_rel >= 1.0:
        return kelvin_quad.local_submatrix(vs, x, G, nu, n=40)
    depth = int(np.ceil(np.log2(1.0 / max(d_rel, 1e-6)))) + 4
    return kelvin_quad.local_submatrix(vs, x, G, nu, n=16, depth=depth)


def mp_traction_rows(vs, x, G, nu, dps=30, n=30):
    """Traction rows (S33, S13, S23) x 18 of the pair's block in the source
    frame, in dps-digit arithmetic: collapsed Gauss-Legendre rule of order n
    on the element (accurate to that precision for a point several element
    sizes away), Kelvin's DD kernel

        H_pqi = -C_pqmn C_i3kl d_n d_l G_mk,
        G_mk = ((3 - 4 nu) delta_mk / R + r_m r_k / R^3) / (16 pi G (1 - nu)),

    contracted with the isotropic stiffness term by term.  Returns an
    mpmath matrix (3 x 18) and the float64 version."""
    import mpmath as mp
    with mp.workdps(dps):
        mu, nu_ = mp.mpf(G), mp.mpf(nu)
        lam = 2 * mu * nu_ / (1 - 2 * nu_)
        v = [[mp.mpf(float(vs[i, j])) for j in range(3)] for i in range(3)]  # v[comp][vertex]
        X = [mp.mpf(float(c)) for c in x]
        # element frame
        a1 = [v[i][1] - v[i][0] for i in range(3)]
        a2 = [v[i][2] - v[i][0] for i in range(3)]
        n1 = mp.sqrt(sum(c * c for c in a1))
        e1 = [c / n1 for c in a1]
        e3 = [e1[1] * a2[2] - e1[2] * a2[1], e1[2] * a2[0] - e1[0] * a2[2], e1[0] * a2[1] - e1[1] * a2[0]]
        n3 = mp.sqrt(sum(c * c for c in e3))
        e3 = [c / n3 for c in e3]
        e2 = [e3[1] * e1[2] - e3[2] * e1[1], e3[2] * e1[0] - e3[0] * e1[2], e3[0] * e1[1] - e3[1] * e1[0]]
        Rm = [e1, e2, e3]
        P = [[sum(Rm[r][i] * (v[i][j] - v[i][0]) for i in range(3)) for j in range(3)] for r in range(2)]
        xl = [sum(Rm[r][i] * (X[i] - v[i][0]) for i in range(3)) for r in range(3)]
        area2 = abs((P[0][1] - P[0][0]) * (P[1][2] - P[1][0]) - (P[1][1] - P[1][0]) * (P[0][2] - P[0][0]))
        gx, gw = _mp_gauss_legendre(mp, n)
        cK = 1 / (16 * mp.pi * mu * (1 - nu_))
        out = mp.zeros(3, 18)
        rows = [(2, 2), (0, 2), (1, 2)]
        for iu in range(n):
            u = (gx[iu] + 1) / 2
            for iv in range(n):
                vv = (gx[iv] + 1) / 2 * (1 - u)
                wq = gw[iu] * gw[iv] / 4 * (1 - u) * area2
                L1, L2 = u, vv
                L0 = 1 - L1 - L2
                Lb = [L0, L1, L2]
                N = [Lb[q] * (2 * Lb[q] - 1) for q in range(3)] + [4 * Lb[(q + 1) % 3] * Lb[(q + 2) % 3]
                                                                   for q in range(3)]
                xi = [L0 * P[0][0] + L1 * P[0][1] + L2 * P[0][2], L0 * P[1][0] + L1 * P[1][1] + L2 * P[1][2], 0]
                r = [xl[0] - xi[0], xl[1] - xi[1], xl[2] - xi[2]]
                R2 = r[0] ** 2 + r[1] ** 2 + r[2] ** 2
                R = mp.sqrt(R2)
                iR3, iR5, iR7 = 1 / (R * R2), 1 / (R * R2 * R2), 1 / (R * R2 * R2 * R2)

                def dd(m, k, nn, l):  # d_n d_l G_mk / cK
                    dmk = 1 if m == k else 0
                    t1 = (3 - 4 * nu_) * dmk * (3 * r[nn] * r[l] * iR5 - (1 if nn == l else 0) * iR3)
                    t2 = ((1 if m == l else 0) * (1 if k == nn else 0) + (1 if m == nn else 0) * (1 if k == l else 0)) * iR3
                    t3 = -3 * ((1 if m == l else 0) * r[k] * r[nn] + (1 if k == l else 0) * r[m] * r[nn] +
                               (1 if m == nn else 0) * r[k] * r[l] + (1 if k == nn else 0) * r[m] * r[l] +
                               (1 if l == nn else 0) * r[m] * r[k]) * iR5
                    t4 = 15 * r[m] * r[k] * r[nn] * r[l] * iR7
                    return t1 + t2 + t3 + t4

                D = {}
                for m in range(3):
                    for k in range(3):
                        for nn in range(3):
                            for l in range(nn, 3):
                                D[m, k, nn, l] = D[m, k, l, nn] = dd(m, k, nn, l) * cK
                for i in range(3):
                    # W_mn = C_i3kl d_n d_l G_mk
                    W = [[lam * (1 if i == 2 else 0) * sum(D[m, kk, nn, kk] for kk in range(3)) +
                          mu * (D[m, i, nn, 2] + D[m, 2, nn, i]) for nn in range(3)] for m in range(3)]
                    trW = W[0][0] + W[1][1] + W[2][2]
                    for c, (p, q) in enumerate(rows):
                        H = -(lam * (1 if p == q else 0) * trW + mu * (W[p][q] + W[q][p]))
                        for j in range(6):
                            out[c, 3 * j + i] += wq * N[j] * H
        f = np.array([[float(out[c, j]) for j in range(18)] for c in range(3)])
    return out, f


def _mp_gauss_legendre(mp, n):
    """Nodes and weights of the n-point Gauss-Legendre rule on [-1, 1] at the
    working precision (Newton on P_n from the float64 nodes)."""
    x0, _ = np.polynomial.legendre.leggauss(n)
    xs, ws = [], []
    for x in x0:
        x = mp.mpf(float(x))
        for _ in range(100):
            p0, p1 = mp.mpf(1), x
            for k in range(2, n + 1):
                p0, p1 = p1, ((2 * k - 1) * x * p1 - (k - 1) * p0) / k
            dp = n * (x * p1 - p0) / (x * x - 1)
            dx = p1 / dp
            x -= dx
            if abs(dx) < mp.mpf(10) ** (-mp.mp.dps - 5):
                break
        p0, p1 = mp.mpf(1), x
        for k in range(2, n + 1):
            p0, p1 = p1, ((2 * k - 1) * x * p1 - (k - 1) * p0) / k
        dp = n * (x * p1 - p0) / (x * x - 1)
        xs.append(x)
        ws.append(2 / ((1 - x * x) * dp * dp))
    return xs, ws
