"""The spherical-shell blending map (geometry/IcosahedralShellMap.hpp) and the blended P1 Laplacian
(P1ElementwiseBlendingLaplaceOperator = P1ElementwiseOperator< p1_diffusion_blending_q3 >), restated from their
definitions for the tests (test infrastructure only), independently of the oracle's C++.

Map of a macro-cell: two of its vertices lie on one ray from the origin. In a cell with one vertex on one sphere and
three on the other, the lone vertex is `ref`, its partner on the ray is `ray` and the other two are `thr` (the larger
local index) and `for`; a cell with two vertices on each sphere takes the classification of the cell across its
first face (local order (0,1,2), (0,1,3), (0,2,3), (1,2,3)) that does not hold both ray vertices. For x in the cell,
lam(x) = |det(x - ray, thr - ray, for - ray) / det(ref - ray, thr - ray, for - ray)| and
F(x) = x (r_ray + lam (r_ref - r_ray)) / |x|.
DF is derived here from that formula: with s = |x|, r = r_ray + lam (r_ref - r_ray),
DF = (r / s) I + x (x) ((r_ref - r_ray) grad lam / s - r x / s^3).
Element matrix of an affine micro-tetrahedron X (rows = vertices) with the 5-point degree-3 rule (centroid weight
-2/15, points with barycentric coordinates (1/2, 1/6, 1/6, 1/6) weight 3/40):
a_ij = |det J| sum_q w_q |det DF_q| (DF_q^-T g_i) . (DF_q^-T g_j), g_i the affine gradients of the barycentric
coordinates, DF_q evaluated at the affine quadrature point.
"""
import numpy as np

LOCAL_FACES = [(0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)]


def _ray_pair(V):
    r = np.linalg.norm(V, axis=1)
    for k in range(4):
        for j in range(k + 1, 4):
            c = np.cross(V[k], V[j])
            if np.all(np.abs(c) / (r[k] * r[j]) <= 1e-14):
                return k, j
    raise ValueError("no two vertices on one ray")


def _kind(V):
    r = np.linalg.norm(V, axis=1)
    eps = (r.max() - r.min()) * 1e-3
    n_outer = int(np.sum(r.max() - r < eps))
    return {1: "outwards", 2: "skew", 3: "inwards"}[n_outer]


def classify(X, cells):
    """Per cell the map parameters (ref, ray, thr, for, r_ref, r_ray) from the macro-vertex coordinates X and the
    cells' vertex lists (local order)."""
    faces = {}
    for c, cv in enumerate(cells):
        for f in LOCAL_FACES:
            faces.setdefault(tuple(sorted(int(cv[i]) for i in f)), []).append(c)

    def of(c, depth=0):
        cv = cells[c]
        V = X[cv]
        k, j = _ray_pair(V)
        kind = _kind(V)
        if kind == "skew":
            assert depth < 4
            ray_ids = {int(cv[k]), int(cv[j])}
            for f in LOCAL_FACES:
                ids = tuple(sorted(int(cv[i]) for i in f))
                if ray_ids <= set(ids):
                    continue
                other = [o for o in faces[ids] if o != c]
                if other:
                    return of(other[0], depth + 1)
            raise ValueError("no non-skew neighbour")
        r = np.linalg.norm(V, axis=1)
        ref, ray = k, j
        if kind == "outwards" and r[ref] < r[ray]:
            ref, ray = ray, ref
        if kind == "inwards" and r[ref] > r[ray]:
            ref, ray = ray, ref
        rest = [i for i in range(4) if i not in (ref, ray)]
        thr, fou = rest[1], rest[0]
        return (V[ref], V[ray], V[thr], V[fou], r[ref], r[ray])

    return [of(c) for c in range(len(cells))]


def _lam(p, x):
    ref, ray, thr, fou, _, _ = p
    n = np.cross(thr - ray, fou - ray)
    vt = np.dot(ref - ray, n)
    v = (x - ray) @ n / vt
    return np.abs(v), np.sign(v)[:, None] * n[None, :] / vt


def F(p, x):
    x = np.atleast_2d(x)
    lam, _ = _lam(p, x)
    r = p[5] + lam * (p[4] - p[5])
    return x * (r / np.linalg.norm(x, axis=1))[:, None]


def DF(p, x):
    """(n, 3, 3) Jacobians at the rows of x"""
    x = np.atleast_2d(x)
    lam, glam = _lam(p, x)
    s = np.linalg.norm(x, axis=1)
    r = p[5] + lam * (p[4] - p[5])
    out = (r / s)[:, None, None] * np.eye(3)[None]
    v = (p[4] - p[5]) * glam / s[:, None] - (r / s ** 3)[:, None] * x
    return out + x[:, :, None] * v[:, None, :]


QUAD_LAM = np.array([[0.25] * 4] + [[0.5 if i == q else 1.0 / 6 for i in range(4)] for q in range(4)])
QUAD_W = np.array([-2.0 / 15] + [3.0 / 40] * 4)


def element_matrix(p, Xe):
    J = (Xe[1:] - Xe[0]).T
    Ji = np.linalg.inv(J)
    g = np.vstack([-Ji.sum(axis=0), Ji])  # rows: gradients of the barycentric coordinates
    A = np.zeros((4, 4))
    for lam, w in zip(QUAD_LAM, QUAD_W):
        xq = lam @ Xe
        D = DF(p, xq)[0]
        Dit = np.linalg.inv(D).T
        G = g @ Dit.T  # rows: DF^-T g_i
        A += w * abs(np.linalg.det(D)) * (G @ G.T)
    return abs(np.linalg.det(J)) * A
