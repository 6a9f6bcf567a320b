"""Synthetic crack meshes (concentric-ring triangulations).

A disk is meshed as a centre node plus K rings of nodes; ring k holds n_k
nodes at radius k/K and the band between rings k-1 and k is zipped into
n_{k-1} + n_k triangles (n_1 triangles around the centre).  The total is
2 * sum(n_k) - n_K, so the element count can be hit exactly by adjusting the
outermost ring.  Nodes use the reference's 6-row layout
(x, y, z, surface id = 1, tip flag, 0) and elements the 5-row layout
(v0, v1, v2, surface ids 1, 1), 0-based
(Mesh_Files/*.npy, src/Core/surface_mesh_utilities.cpp:38-74).
"""
import numpy as np


def _ring_counts(n_el):
    if n_el < 3:
        raise ValueError("need at least 3 elements")
    best = None
    for K in range(1, 400):
        # n_k ~ a k with a chosen so the total is close to n_el
        a = n_el / (K * K)
        n = [max(3, int(round(a * k))) for k in range(1, K + 1)]
        total = 2 * sum(n) - n[-1]
        nK = n[-1] + (n_el - total)
        if nK < max(3, n[-2] if K > 1 else 3):
            continue
        n[-1] = nK
        # prefer ring spacing ~ arc spacing (equilateral-ish triangles)
        h = 1.0 / K
        arc = 2 * np.pi / n[-1]
        score = abs(np.log(arc / h))
        if best is None or score < best[0]:
            best = (score, n)
    return best[1]


def disk(n_el):
    """Unit disk in z = 0 with exactly n_el triangles (counter-clockwise)."""
    counts = _ring_counts(n_el)
    K = len(counts)
    pts = [(0.0, 0.0)]
    rings = [[0]]
    for k, nk in enumerate(counts, start=1):
        r = k / K
        # golden-angle phase per ring: no exact alignments between rings, so
        # no collocation point sits on the extension of another element's edge
        phase = np.mod(k * 2.399963229728653, 2 * np.pi / nk)
        ids = []
        for i in range(nk):
            t = phase + 2 * np.pi * i / nk
            ids.append(len(pts))
            pts.append((r * np.cos(t), r * np.sin(t)))
        rings.append(ids)
    pts = np.array(pts)
    ang = np.arctan2(pts[:, 1], pts[:, 0])
    tris = []
    for k in range(1, K + 1):
        inner, outer = rings[k - 1], rings[k]
        if k == 1:
            for i in range(len(outer)):
                tris.append((0, outer[i], outer[(i + 1) % len(outer)]))
            continue
        # zip the two rings by angle
        ni, no = len(inner), len(outer)
        a_in = np.mod(ang[inner] - ang[outer[0]], 2 * np.pi)
        s_in = int(np.argmin(a_in))
        inner = inner[s_in:] + inner[:s_in]
        a_in = np.mod(ang[inner] - ang[outer[0]], 2 * np.pi)
        a_in = np.where(a_in < 0, a_in + 2 * np.pi, a_in)
        a_out = np.mod(ang[outer] - ang[outer[0]], 2 * np.pi)
        i = j = 0
        while i < ni or j < no:
            ci, cj = inner[i % ni], outer[j % no]
            ai_next = a_in[i + 1] if i + 1 < ni else 2 * np.pi + a_in[0]
            ao_next = a_out[j + 1] if j + 1 < no else 2 * np.pi
            if j < no and (i >= ni or ao_next <= ai_next):
                tris.append((ci, cj, outer[(j + 1) % no]))
                j += 1
            else:
                tris.append((ci, cj, inner[(i + 1) % ni]))
                i += 1
    tris = np.array(tris, dtype=np.int64)
    # orient counter-clockwise (normal +z)
    p = pts[tris]
    cr = (p[:, 1, 0] - p[:, 0, 0]) * (p[:, 2, 1] - p[:, 0, 1]) - (p[:, 1, 1] - p[:, 0, 1]) * (p[:, 2, 0] - p[:, 0, 0])
    flip = cr < 0
    tris[flip] = tris[flip][:, [0, 2, 1]]
    assert len(tris) == n_el, (len(tris), n_el)
    nods = np.zeros((6, len(pts)))
    nods[0], nods[1] = pts[:, 0], pts[:, 1]
    nods[3] = 1.0
    nods[4, rings[-1]] = 1.0
    conn = np.ones((5, n_el), dtype=np.int64)
    conn[:3] = tris.T
    return np.asfortranarray(nods), np.asfortranarray(conn)


def rotation(axis, angle):
    axis = np.asarray(axis, float)
    axis = axis / np.linalg.norm(axis)
    x, y, z = axis
    c, s = np.cos(angle), np.sin(angle)
    C = 1 - c
    return np.array([[c + x * x * C, x * y * C - z * s, x * z * C + y * s],
                     [y * x * C + z * s, c + y * y * C, y * z * C - x * s],
                     [z * x * C - y * s, z * y * C + x * s, c + z * z * C]])


def ellipse(n_el, a=1.0, b=0.5, R=None):
    """Elliptical crack with semi-axes a, b (x, y), optionally rotated by R."""
    nods, conn = disk(n_el)
    nods[0] *= a
    nods[1] *= b
    if R is not None:
        nods[:3] = R @ nods[:3]
    return nods, conn


def wavy_disk(n_el, amp=0.1):
    """Non-planar crack z = amp sin(2 pi x) sin(2 pi y) over the unit disk."""
    nods, conn = disk(n_el)
    nods[2] = amp * np.sin(2 * np.pi * nods[0]) * np.sin(2 * np.pi * nods[1])
    return nods, conn


def random_rotation(rng):
    """Uniform rotation from a normalised Gaussian quaternion."""
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def batch_cracks(n_cracks=256, seed=2, n_min=40, n_max=500):
    """Independent planar elliptical cracks of uneven size (draw order:
    element counts, then per crack semi-axes (2), quaternion (4), nu, p)."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(n_min, n_max + 1, size=n_cracks)
    out = []
    for n in sizes:
        a, b = rng.uniform(0.5, 2.0, size=2)
        R = random_rotation(rng)
        nu = rng.uniform(0.2, 0.4)
        p = rng.uniform(0.5, 2.0)
        nods, conn = ellipse(int(n), a, b, R)
        out.append(dict(nods=nods, conn=conn, nu=nu, p=p, a=a, b=b, R=R))
    return out
