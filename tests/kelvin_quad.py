"""Independent check of the per-pair kernel: brute-force quadrature of the
DD -> stress kernel built directly from Kelvin's displacement tensor,

    H_pqi(r) = - C_pqmn C_i3kl d_n d_l G_mk(r),
    G_mk = ((3 - 4 nu) delta_mk / |r| + r_m r_k / |r|^3) / (16 pi mu (1 - nu)),

with r = X - xi, integrated over the element against the quadratic shape
functions (barycentric form) with a collapsed (Duffy) Gauss-Legendre rule.
Valid for points well away from the element, where the integrand is smooth.
Shares no code or coefficient table with the oracle or the HIP path."""
import numpy as np


def _c_tensor(mu, nu):
    lam = 2 * mu * nu / (1 - 2 * nu)
    d = np.eye(3)
    return (lam * np.einsum("ij,kl->ijkl", d, d) +
            mu * (np.einsum("ik,jl->ijkl", d, d) + np.einsum("il,jk->ijkl", d, d)))


def _ddG(r, mu, nu):
    """d_n d_l G_mk at points r (N x 3) -> (N, m, k, n, l)."""
    R = np.linalg.norm(r, axis=1)
    d = np.eye(3)
    R3, R5, R7 = R ** 3, R ** 5, R ** 7
    # (3 - 4 nu) delta_mk d_n d_l (1/R)
    t1 = (3 - 4 * nu) * np.einsum("mk,pnl->pmknl", d,
                                  3 * np.einsum("pn,pl->pnl", r, r) / R5[:, None, None] -
                                  d[None] / R3[:, None, None])
    # d_n d_l (r_m r_k / R^3)
    a = (np.einsum("ml,kn->mknl", d, d) + np.einsum("mn,kl->mknl", d, d))[None] / R3[:, None, None, None, None]
    b = -3 * (np.einsum("ml,pk,pn->pmknl", d, r, r) + np.einsum("kl,pm,pn->pmknl", d, r, r) +
              np.einsum("mn,pk,pl->pmknl", d, r, r) + np.einsum("kn,pm,pl->pmknl", d, r, r) +
              np.einsum("ln,pm,pk->pmknl", d, r, r)) / R5[:, None, None, None, None]
    c = 15 * np.einsum("pm,pk,pn,pl->pmknl", r, r, r, r) / R7[:, None, None, None, None]
    return (t1 + a + b + c) / (16 * np.pi * mu * (1 - nu))


def kernel_H(r, mu, nu):
    """H_pqi at points r: (N, 3, 3, 3) (stress pq due to DD component i)."""
    C = _c_tensor(mu, nu)
    dd = _ddG(r, mu, nu)
    return -np.einsum("pqmn,ikl,xmknl->xpqi", C, C[:, 2], dd)


def tri_rule(n):
    """Collapsed Gauss-Legendre rule on the unit triangle: (u, v, w)."""
    x, w = np.polynomial.legendre.leggauss(n)
    x = 0.5 * (x + 1)
    w = 0.5 * w
    U, V = np.meshgrid(x, x, indexing="ij")
    WU, WV = np.meshgrid(w, w, indexing="ij")
    u = U.ravel()
    v = (V * (1 - U)).ravel()
    wt = (WU * WV * (1 - U)).ravel()
    return u, v, wt


def _subdivide(tris, levels):
    for _ in range(levels):
        out = []
        for a, b, c in tris:
            ab, bc, ca = (a + b) / 2, (b + c) / 2, (c + a) / 2
            out += [(a, ab, ca), (ab, b, bc), (ca, bc, c), (ab, bc, ca)]
        tris = out
    return tris


def _graded(tris, x, depth):
    """Refine sub-triangles whose distance to x is below twice their size."""
    out = []
    for t in tris:
        a, b, c = t
        size = max(np.linalg.norm(a - b), np.linalg.norm(b - c), np.linalg.norm(c - a))
        cen = (a + b + c) / 3
        if depth > 0 and np.linalg.norm(cen - x) < 2.0 * size:
            out += _graded(_subdivide([t], 1), x, depth - 1)
        else:
            out.append(t)
    return out


def local_submatrix(vert, X, mu, nu, n=40, depth=0):
    """6 x 18 stress influence in the element frame (rows S11 S22 S33 S12 S13
    S23, cols 3*node + DD component in the element frame) by quadrature;
    depth > 0 grades the integration cells towards the point (for points
    near the element but outside it)."""
    if depth > 0:
        return _local_graded(vert, X, mu, nu, n, depth)
    v0, v1, v2 = vert[:, 0], vert[:, 1], vert[:, 2]
    e1 = (v1 - v0) / np.linalg.norm(v1 - v0)
    e3 = np.cross(e1, v2 - v0)
    e3 /= np.linalg.norm(e3)
    e2 = np.cross(e3, e1)
    R = np.array([e1, e2, e3])
    P = (R @ (vert - v0[:, None]))          # vertices in the frame (z = 0)
    x = R @ (X - v0)
    u, v, wt = tri_rule(n)
    L1, L2 = u, v
    L0 = 1 - u - v
    xi = (np.outer(L0, P[:, 0]) + np.outer(L1, P[:, 1]) + np.outer(L2, P[:, 2]))
    area = 0.5 * abs((P[0, 1] - P[0, 0]) * (P[1, 2] - P[1, 0]) - (P[1, 1] - P[1, 0]) * (P[0, 2] - P[0, 0]))
    Lb = [L0, L1, L2]
    N = [Lb[k] * (2 * Lb[k] - 1) for k in range(3)] + [4 * Lb[(k + 1) % 3] * Lb[(k + 2) % 3] for k in range(3)]
    H = kernel_H(x[None, :] - xi, mu, nu)
    comp = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    out = np.zeros((6, 18))
    for j in range(6):
        wj = 2 * area * wt * N[j]
        for c, (p, q) in enumerate(comp):
            out[c, 3 * j:3 * j + 3] = np.einsum("x,xi->i", wj, H[:, p, q, :])
    return out


def _local_graded(vert, X, mu, nu, n, depth):
    v0, v1, v2 = vert[:, 0], vert[:, 1], vert[:, 2]
    e1 = (v1 - v0) / np.linalg.norm(v1 - v0)
    e3 = np.cross(e1, v2 - v0)
    e3 /= np.linalg.norm(e3)
    e2 = np.cross(e3, e1)
    R = np.array([e1, e2, e3])
    P = R @ (vert - v0[:, None])
    x = R @ (X - v0)
    area2 = (P[0, 1] - P[0, 0]) * (P[1, 2] - P[1, 0]) - (P[1, 1] - P[1, 0]) * (P[0, 2] - P[0, 0])
    cells = _graded([(P[:2, 0], P[:2, 1], P[:2, 2])], x[:2], depth)
    u, v, wt = tri_rule(n)
    pts, wts = [], []
    for a, b, c in cells:
        pts.append(np.outer(1 - u - v, a) + np.outer(u, b) + np.outer(v, c))
        ar = abs((b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0]))
        wts.append(wt * ar)
    xi2 = np.concatenate(pts)
    w = np.concatenate(wts)
    # barycentric coordinates of the integration points
    T = np.array([[P[0, 1] - P[0, 0], P[0, 2] - P[0, 0]], [P[1, 1] - P[1, 0], P[1, 2] - P[1, 0]]])
    lam = np.linalg.solve(T, (xi2 - P[:2, 0]).T)
    L1, L2 = lam
    L0 = 1 - L1 - L2
    Lb = [L0, L1, L2]
    N = [Lb[k] * (2 * Lb[k] - 1) for k in range(3)] + [4 * Lb[(k + 1) % 3] * Lb[(k + 2) % 3] for k in range(3)]
    xi = np.zeros((len(w), 3))
    xi[:, :2] = xi2
    H = kernel_H(x[None, :] - xi, mu, nu)
    comp = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    out = np.zeros((6, 18))
    for j in range(6):
        wj = w * N[j]
        for c, (p, q) in enumerate(comp):
            out[c, 3 * j:3 * j + 3] = np.einsum("x,xi->i", wj, H[:, p, q, :])
    return out
