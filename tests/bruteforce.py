"""Independent brute-force model of the P1 finite-element problem, used to pin the oracle.

Nothing here shares code with the oracle or the HIP library. The refined mesh is built
element by element from the Gmsh macro mesh, using the six micro-tetrahedron types of a
uniformly refined macro-cell (reference: src/hyteg/volumedofspace/CellDoFIndexing.hpp:155-198),
micro-vertices are identified globally by their exact coordinates, and the global P1
stiffness matrix is assembled with scipy from the closed-form element matrix
K_T = vol(T) * G G^T (G = gradients of the barycentric coordinates). Linear interpolation
between levels is assembled from the coarse micro-edges (a fine vertex is either a coarse
vertex or the midpoint of a coarse micro-edge).
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

# lattice offsets of the four vertices of the six micro-cell types (WHITE_UP, BLUE_UP,
# GREEN_UP, WHITE_DOWN, BLUE_DOWN, GREEN_DOWN)
MICRO_TETS = np.array([
    [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)],
    [(1, 0, 0), (1, 1, 0), (0, 1, 0), (1, 0, 1)],
    [(1, 0, 0), (0, 1, 0), (1, 0, 1), (0, 0, 1)],
    [(1, 1, 0), (1, 1, 1), (0, 1, 1), (1, 0, 1)],
    [(1, 0, 1), (0, 1, 1), (0, 0, 1), (0, 1, 0)],
    [(0, 1, 0), (1, 1, 0), (1, 0, 1), (0, 1, 1)],
], dtype=np.int64)


def read_gmsh(path):
    """Nodes (dict id -> xyz) and tetrahedra (vertex id lists, Gmsh order) of a v2.2 file."""
    with open(path) as fh:
        tok = fh.read().split()
    i = tok.index("$Nodes") + 1
    nn = int(tok[i]); i += 1
    nodes = {}
    for _ in range(nn):
        nodes[int(tok[i])] = np.array([float(t) for t in tok[i + 1:i + 4]])
        i += 4
    i = tok.index("$Elements") + 1
    ne = int(tok[i]); i += 1
    tets = []
    nper = {15: 1, 1: 2, 2: 3, 4: 4}
    for _ in range(ne):
        typ, ntag = int(tok[i + 1]), int(tok[i + 2])
        i += 3 + ntag
        vs = [int(t) for t in tok[i:i + nper[typ]]]
        i += nper[typ]
        if typ == 4:
            tets.append(vs)
    return nodes, tets


def _lattice(N):
    pts = [(x, y, z) for z in range(N + 1) for y in range(N + 1 - z) for x in range(N + 1 - y - z)]
    return np.array(pts, dtype=np.int64)


class RefinedMesh:
    """Uniformly refined macro mesh at level L with globally numbered micro-vertices. `path` is a Gmsh file or a
    (nodes, tets) pair: nodes maps a vertex id to its coordinates, tets lists the cells' vertex ids in local order."""

    def __init__(self, path, L, scale_key=None):
        nodes, tets = read_gmsh(path) if isinstance(path, str) else path
        self.N = N = 1 << L
        self.L = L
        keyscale = 4 * N if scale_key in (None, "tree") else scale_key
        allkeys = []
        allxyz = []
        elems = []
        off = 0
        w = N + 1
        lat = _lattice(N)
        for cv in tets:
            V = np.array([nodes[v] for v in cv])
            xyz = V[0] + (lat @ (V[1:] - V[0])) / N
            allkeys.append(np.rint(xyz * keyscale).astype(np.int64))
            allxyz.append(xyz)
            # micro-tets: for every type, every lattice base with all four vertices inside
            for t in MICRO_TETS:
                verts = lat[:, None, :] + t[None, :, :]
                ok = np.all(verts.sum(axis=2) <= N, axis=1)
                vv = verts[ok]
                x, y, z = vv[..., 0], vv[..., 1], vv[..., 2]
                s3 = lambda q: q * (q + 1) * (q + 2) // 6
                ids = s3(w) - s3(w - z) + y * (w - z + 1) - y * (y + 1) // 2 + x
                elems.append(ids + off)
            off += len(lat)
        keys = np.concatenate(allkeys)
        if scale_key == "tree":
            # coordinates that are not dyadic (curved meshes): identify micro-vertices that coincide up to
            # rounding by a radius search instead of exact keys
            from scipy.sparse.csgraph import connected_components
            from scipy.spatial import cKDTree
            pts = np.concatenate(allxyz)
            pairs = cKDTree(pts).query_pairs(r=1e-10, output_type="ndarray")
            adj = sp.coo_matrix((np.ones(len(pairs)), (pairs[:, 0], pairs[:, 1])), shape=(len(pts), len(pts)))
            _, lab = connected_components(adj, directed=False)
            _, first, inv = np.unique(lab, return_index=True, return_inverse=True)
            uniq = keys[first]
        else:
            uniq, first, inv = np.unique(keys, axis=0, return_index=True, return_inverse=True)
        inv = inv.reshape(-1)
        self.keys = uniq
        self.xyz = np.concatenate(allxyz)[first]  # coordinates of the first cell that produced the vertex
        self.keyscale = keyscale
        self.tets = inv[np.concatenate(elems)]
        self.n = len(uniq)
        # boundary micro-vertices: vertices of micro-faces that belong to one micro-tet only
        faces = np.sort(np.concatenate([self.tets[:, [0, 1, 2]], self.tets[:, [0, 1, 3]],
                                        self.tets[:, [0, 2, 3]], self.tets[:, [1, 2, 3]]]), axis=1)
        uf, cnt = np.unique(faces, axis=0, return_counts=True)
        self.boundary = np.zeros(self.n, dtype=bool)
        self.boundary[uf[cnt == 1].reshape(-1)] = True

    def index_of(self, xyz):
        """Global indices of the given coordinates (nearest micro-vertex, which must coincide)."""
        from scipy.spatial import cKDTree
        if not hasattr(self, "_tree"):
            self._tree = cKDTree(self.xyz)
        d, i = self._tree.query(np.asarray(xyz).reshape(-1, 3))
        assert d.max() < 1e-9
        return i

    def key_index(self, xyz):
        """Global indices of the given coordinates."""
        k = np.rint(np.asarray(xyz) * self.keyscale).astype(np.int64)
        lookup = {tuple(r): i for i, r in enumerate(self.keys)}
        return np.array([lookup[tuple(r)] for r in k], dtype=np.int64)

    def stiffness(self, kfun=None):
        X = self.xyz[self.tets]                     # (ne, 4, 3)
        J = (X[:, 1:, :] - X[:, :1, :])             # rows = edge vectors x_k - x_0
        # lambda_k(x) = e_k . J^{-T} (x - x_0): the gradient of lambda_k is row k of J^{-1} transposed
        Jinv = np.linalg.inv(J)
        g = np.empty((len(X), 4, 3))
        g[:, 1:, :] = Jinv.transpose(0, 2, 1)
        g[:, 0, :] = -g[:, 1:, :].sum(axis=1)
        vol = np.abs(np.linalg.det(J)) / 6.0
        if kfun is not None:
            vol = vol * kfun(X)
        K = vol[:, None, None] * np.einsum("eik,ejk->eij", g, g)
        rows = np.repeat(self.tets, 4, axis=1).reshape(-1)
        cols = np.tile(self.tets, (1, 4)).reshape(-1)
        return sp.csr_matrix((K.reshape(-1), (rows, cols)), shape=(self.n, self.n))


def prolongation(coarse, fine):
    """Linear interpolation matrix (fine.n x coarse.n)."""
    # map coarse vertices into fine numbering via coordinates
    cf = fine.key_index(coarse.xyz)
    rows, cols, vals = list(cf), list(range(coarse.n)), [1.0] * coarse.n
    edges = set()
    for t in coarse.tets:
        for a in range(4):
            for b in range(a + 1, 4):
                i, j = int(t[a]), int(t[b])
                edges.add((min(i, j), max(i, j)))
    edges = np.array(sorted(edges))
    mids = 0.5 * (coarse.xyz[edges[:, 0]] + coarse.xyz[edges[:, 1]])
    mf = fine.key_index(mids)
    for (i, j), f in zip(edges, mf):
        rows += [f, f]
        cols += [i, j]
        vals += [0.5, 0.5]
    P = sp.csr_matrix((vals, (rows, cols)), shape=(fine.n, coarse.n))
    assert np.allclose(np.asarray(P.sum(axis=1)).ravel(), 1.0)
    return P


class GMGModel:
    """V-cycle on inner DoFs with global sparse matrices (zero Dirichlet boundary)."""

    def __init__(self, path, minL, maxL):
        self.minL, self.maxL = minL, maxL
        self.mesh = {L: RefinedMesh(path, L) for L in range(minL, maxL + 1)}
        self.inner = {L: np.flatnonzero(~m.boundary) for L, m in self.mesh.items()}
        self.A = {}
        self.P = {}
        for L, m in self.mesh.items():
            Af = m.stiffness()
            ii = self.inner[L]
            self.A[L] = Af[ii][:, ii].tocsr()
        for L in range(minL + 1, maxL + 1):
            P = prolongation(self.mesh[L - 1], self.mesh[L])
            self.P[L] = P[self.inner[L]][:, self.inner[L - 1]].tocsr()

    def jacobi(self, L, x, b, omega):
        A = self.A[L]
        return x + omega * (b - A @ x) / A.diagonal()

    def chebyshev(self, L, x, b, coeffs):
        A = self.A[L]
        dinv = 1.0 / A.diagonal()
        t = dinv * (b - A @ x)
        x = x + coeffs[0] * t
        for c in coeffs[1:]:
            t = dinv * (A @ t)
            x = x + c * t
        return x

    def vcycle(self, L, x, b, smooth, pre, post):
        if L == self.minL:
            if self.A[L].shape[0] == 0:
                return x
            return spla.spsolve(self.A[L].tocsc(), b) if self.A[L].shape[0] > 1 else b / self.A[L].toarray()[0]
        for _ in range(pre):
            x = smooth(L, x, b)
        r = b - self.A[L] @ x
        bc = self.P[L].T @ r
        xc = self.vcycle(L - 1, np.zeros_like(bc), bc, smooth, pre, post)
        x = x + self.P[L] @ xc
        for _ in range(post):
            x = smooth(L, x, b)
        return x


# ---------------------------------------------------------------- P2 (quadratic) model
# local P2 nodes: vertices 0..3, then edges (2,3) (1,3) (1,2) (0,3) (0,2) (0,1)
P2_EDGES = [(2, 3), (1, 3), (1, 2), (0, 3), (0, 2), (0, 1)]


def p2_element_exact(X):
    """P2 stiffness matrices of tetrahedra X (ne, 4, 3) by exact integration of the barycentric monomials:
    grad phi_i = sum_k lambda_k a_ik + b_i, int lambda_k lambda_l = vol (1 + delta_kl) / 20, int lambda_k = vol / 4."""
    J = X[:, 1:, :] - X[:, :1, :]
    Jinv = np.linalg.inv(J)
    g = np.empty((len(X), 4, 3))
    g[:, 1:, :] = Jinv.transpose(0, 2, 1)
    g[:, 0, :] = -g[:, 1:, :].sum(axis=1)
    vol = np.abs(np.linalg.det(J)) / 6.0
    ne = len(X)
    a = np.zeros((ne, 10, 4, 3))
    b = np.zeros((ne, 10, 3))
    for v in range(4):
        a[:, v, v, :] = 4.0 * g[:, v, :]
        b[:, v, :] = -g[:, v, :]
    for e, (p, q) in enumerate(P2_EDGES):
        a[:, 4 + e, p, :] = 4.0 * g[:, q, :]
        a[:, 4 + e, q, :] = 4.0 * g[:, p, :]
    mll = (np.ones((4, 4)) + np.eye(4)) / 20.0
    K = np.einsum("eikc,ejlc,kl->eij", a, a, mll)
    K += np.einsum("eikc,ejc->eij", a, b) / 4.0 + np.einsum("eic,ejkc->eij", b, a) / 4.0
    K += np.einsum("eic,ejc->eij", b, b)
    return vol[:, None, None] * K


def tet_gauss_rule(n):
    """Collapsed-coordinate Gauss rule on the reference tetrahedron (n^3 points; the collapse multiplies the
    integrand by (1 - a)^2 (1 - b), so polynomials of total degree 2n - 3 are integrated exactly):
    barycentric coordinates (npts, 4) and weights summing to 1/6."""
    g, w = np.polynomial.legendre.leggauss(n)
    g, w = (g + 1) / 2, w / 2
    lam, wts = [], []
    for a, wa in zip(g, w):
        for b, wb in zip(g, w):
            for c, wc in zip(g, w):
                p = (a, b * (1 - a), c * (1 - a) * (1 - b))
                lam.append((1 - sum(p), p[0], p[1], p[2]))
                wts.append(wa * wb * wc * (1 - a) ** 2 * (1 - b))
    return np.array(lam), np.array(wts)


def p2_element_varcoef(X, kfun, lam, wts):
    """P2 matrices of -div(k grad) on tetrahedra X (ne, 4, 3) with the quadrature rule (lam, wts):
    K_e = sum_q w_q |det J_e| k(x_q) grad phi(lam_q) grad phi(lam_q)^T, k evaluated at the physical points."""
    J = X[:, 1:, :] - X[:, :1, :]
    g = np.empty((len(X), 4, 3))
    g[:, 1:, :] = np.linalg.inv(J).transpose(0, 2, 1)
    g[:, 0, :] = -g[:, 1:, :].sum(axis=1)
    adet = np.abs(np.linalg.det(J))
    K = np.zeros((len(X), 10, 10))
    for lq, wq in zip(lam, wts):
        # gradients of the ten P2 basis functions at lq: vertices (4 l_v - 1) g_v, edges 4 (l_a g_b + l_b g_a)
        d = [(4 * lq[v] - 1) * g[:, v, :] for v in range(4)]
        d += [4 * (lq[a] * g[:, b, :] + lq[b] * g[:, a, :]) for a, b in P2_EDGES]
        d = np.stack(d, axis=1)                                    # (ne, 10, 3)
        xq = np.einsum("v,evc->ec", lq, X)                         # physical point of every element
        K += (wq * adet * kfun(xq))[:, None, None] * np.einsum("eic,ejc->eij", d, d)
    return K


class P2Model:
    """P2 of level L on the lattice of level L+1 (global numbering by exact coordinates)."""

    def __init__(self, path, L, scale_key=None):
        self.coarse = RefinedMesh(path, L, scale_key)
        self.lat = RefinedMesh(path, L + 1, scale_key)
        self.n = self.lat.n
        self.boundary = self.lat.boundary
        X = self.coarse.xyz[self.coarse.tets]
        nodes = np.concatenate([X, np.stack([(X[:, p] + X[:, q]) / 2 for p, q in P2_EDGES], axis=1)], axis=1)
        self.X = X
        if scale_key == "tree":
            self.nodes = self.lat.index_of(nodes).reshape(-1, 10)
            return
        lookup = {tuple(r): i for i, r in enumerate(self.lat.keys)}
        k = np.rint(nodes.reshape(-1, 3) * self.lat.keyscale).astype(np.int64)
        self.nodes = np.array([lookup[tuple(r)] for r in k], dtype=np.int64).reshape(-1, 10)

    def stiffness(self, kfun=None, rule=None):
        """Global P2 matrix of -Laplace, or of -div(k grad) with kfun (points (n, 3) -> values) under `rule`
        = (barycentric points, weights)."""
        K = p2_element_exact(self.X) if kfun is None else p2_element_varcoef(self.X, kfun, *rule)
        rows = np.repeat(self.nodes, 10, axis=1).reshape(-1)
        cols = np.tile(self.nodes, (1, 10)).reshape(-1)
        return sp.csr_matrix((K.reshape(-1), (rows, cols)), shape=(self.n, self.n))


def p2_basis(lam):
    """P2 basis values (npts, 10) at barycentric coordinates lam (npts, 4)."""
    out = [lam[:, v] * (2 * lam[:, v] - 1) for v in range(4)]
    out += [4 * lam[:, p] * lam[:, q] for p, q in P2_EDGES]
    return np.stack(out, axis=1)


def prolongation_p2(coarse, fine):
    """Quadratic interpolation matrix (fine.n x coarse.n) from P2Model coarse (level L) to fine (level L+1)."""
    pts4 = np.array([(i, j, k, 4 - i - j - k) for i in range(5) for j in range(5 - i) for k in range(5 - i - j)], float)
    lam = np.stack([pts4[:, 3], pts4[:, 0], pts4[:, 1], pts4[:, 2]], axis=1) / 4.0  # lambda_0..3
    phi = p2_basis(lam)                                                             # (35, 10)
    X = coarse.X                                                                    # (ne, 4, 3)
    P = np.einsum("pk,ekc->epc", lam, X)                                            # points (ne, 35, 3)
    lookup = {tuple(r): i for i, r in enumerate(fine.lat.keys)}
    k = np.rint(P.reshape(-1, 3) * fine.lat.keyscale).astype(np.int64)
    fidx = np.array([lookup[tuple(r)] for r in k], dtype=np.int64).reshape(len(X), 35)
    rows = np.repeat(fidx[:, :, None], 10, axis=2).reshape(-1)
    cols = np.repeat(coarse.nodes[:, None, :], 35, axis=1).reshape(-1)
    vals = np.broadcast_to(phi[None], (len(X), 35, 10)).reshape(-1)
    keep = vals != 0.0
    rc = np.stack([rows[keep], cols[keep]], axis=1)
    uniq, first = np.unique(rc, axis=0, return_index=True)
    Pm = sp.csr_matrix((vals[keep][first], (uniq[:, 0], uniq[:, 1])), shape=(fine.n, coarse.n))
    assert np.allclose(np.asarray(Pm.sum(axis=1)).ravel(), 1.0)
    return Pm


class P2GMGModel(GMGModel):
    """V-cycle for P2 on inner DoFs with global sparse matrices (zero Dirichlet boundary)."""

    def __init__(self, path, minL, maxL):
        self.minL, self.maxL = minL, maxL
        self.mesh = {L: P2Model(path, L) for L in range(minL, maxL + 1)}
        self.inner = {L: np.flatnonzero(~m.boundary) for L, m in self.mesh.items()}
        self.A, self.P = {}, {}
        for L, m in self.mesh.items():
            ii = self.inner[L]
            self.A[L] = m.stiffness()[ii][:, ii].tocsr()
        for L in range(minL + 1, maxL + 1):
            P = prolongation_p2(self.mesh[L - 1], self.mesh[L])
            self.P[L] = P[self.inner[L]][:, self.inner[L - 1]].tocsr()
