"""Multi-colour SOR restated for the tests (test infrastructure only).

The library orders SOR by colour (SORSmoothable::smooth_sor, P1Operator.hpp:343-423 visits primitives
lexicographically, which has no parallel counterpart): the macro-vertices get colours c(v) (vertex_colouring),
and a lattice point with barycentric lattice coordinates b_i over the vertices v_i of a cell containing it gets
the colour (sum_i c(v_i) b_i) mod m. Stencil neighbours differ by e_i - e_j or, along a cell's inner diagonal
(x, y, z) +- (1, -1, 1), by e_1 - e_2 + e_3 - e_0, so the colouring is proper when vertex colours differ along
macro-edges and c_1 + c_3 != c_0 + c_2 (mod m) in every cell.
One sweep visits the colours 0..m-1 (backwards: m-1..0); a colour pass sets x_p += relax d_p (b - A x)_p at
every point of that colour whose flag matches, from the residual before the pass (points of one colour are not
coupled, so this is Gauss-Seidel).
"""
import numpy as np


def vertex_colouring(nv, cells):
    """(m, colour of every vertex): the smallest m >= 4 for which the greedy choice in vertex order (smallest
    colour differing from every vertex sharing a cell, with c1 + c3 != c0 + c2 mod m in every cell whose four
    colours are set) succeeds."""
    nb = [set() for _ in range(nv)]
    vc = [[] for _ in range(nv)]
    for ci, c in enumerate(cells):
        for a in c:
            vc[a].append(ci)
            for b in c:
                if a != b:
                    nb[a].add(int(b))
    for m in range(4, 65):
        col = np.full(nv, -1, dtype=np.int64)
        ok = True
        for v in range(nv):
            for k in range(m):
                if any(col[u] == k for u in nb[v]):
                    continue
                col[v] = k
                good = True
                for ci in vc[v]:
                    c = cells[ci]
                    if (col[c] >= 0).all() and (col[c[1]] + col[c[3]] - col[c[0]] - col[c[2]]) % m == 0:
                        good = False
                if good:
                    break
                col[v] = -1
            if col[v] < 0:
                ok = False
                break
        if ok:
            return m, col
    raise RuntimeError("no colouring with at most 64 colours")


def point_colours(xyz, X, cells, vcol, m, N):
    """Colour of every point (rows of xyz) from a containing cell; checks that every containing cell agrees."""
    col = np.full(len(xyz), -1, dtype=np.int64)
    for c in cells:
        V = X[c]
        lam3 = np.linalg.solve((V[1:] - V[0]).T, (xyz - V[0]).T).T
        lam = np.column_stack([1.0 - lam3.sum(axis=1), lam3])
        inside = (lam > -1e-9).all(axis=1)
        b = np.rint(N * lam[inside]).astype(np.int64)
        k = (b @ vcol[c]) % m
        prev = col[inside]
        assert np.all((prev < 0) | (prev == k)), "cells disagree on a point's colour"
        col[inside] = k
    assert np.all(col >= 0), "point outside every cell"
    return col


def sor_sweep(apply, dinv, x, b, relax, colours, m, mask, backwards=False):
    """One multi-colour SOR sweep; apply(x) returns A x, mask marks the points the flag selects."""
    x = np.array(x, dtype=np.float64)
    order = range(m - 1, -1, -1) if backwards else range(m)
    for k in order:
        r = b - apply(x)
        sel = mask & (colours == k)
        x[sel] += relax * dinv[sel] * r[sel]
    return x


# ------------------------------------------------------------------ SOR in the reference's order
def canonical_primitives(om, L):
    """For every DoF in the canonical order (macro-vertices, then the interior points of every edge, face and cell,
    each primitive in storage order): the set of global macro-vertex ids of the primitive it belongs to."""
    N = 1 << L
    edges, faces = om.prims()
    cells = om.cells()
    nv = int(om.info()[0])
    out = [frozenset([v]) for v in range(nv)]
    for e in edges:
        out += [frozenset(int(v) for v in e)] * (N - 1)
    nf = (N - 1) * (N - 2) // 2 if N >= 3 else 0
    for f in faces:
        out += [frozenset(int(v) for v in f)] * nf
    nc = (N - 1) * (N - 2) * (N - 3) // 6 if N >= 4 else 0
    for c in cells:
        out += [frozenset(int(v) for v in c)] * nc
    assert len(out) == om.ndofs(L)
    return out


def ordered_sor(A, x, b, relax, prims, update, backwards=False, closure_rule=True):
    """One SOR sweep in the order of SORSmoothable::smooth_sor (P1Operator.hpp:343-416), written on the global
    matrix A (canonical order): points are visited in canonical order (backwards: reversed), which visits vertices,
    edges, faces and cells in turn and every primitive lexicographically. A point p of primitive P (vertex set
    prims[p]) reads neighbour q's current value when q's primitive lies in P's closure (prims[q] a subset of prims[p])
    or has a higher dimension, and the value before the sweep otherwise (the reference reads those from ghost
    layers filled before the phase). Points with update[p] False keep their values. closure_rule=False reads every
    neighbour's current value (plain Gauss-Seidel in the same order).
    x_p <- (1 - relax) x_p + relax (b_p - sum_{q != p} a_pq x_q) / a_pp."""
    A = A.tocsr()
    old = np.array(x, dtype=np.float64)
    x = old.copy()
    order = range(len(x) - 1, -1, -1) if backwards else range(len(x))
    for p in order:
        if not update[p]:
            continue
        s, c = 0.0, 0.0
        for k in range(A.indptr[p], A.indptr[p + 1]):
            q, a = A.indices[k], A.data[k]
            if q == p:
                c = a
                continue
            cur = not closure_rule or len(prims[q]) > len(prims[p]) or prims[q] <= prims[p]
            s += a * (x[q] if cur else old[q])
        x[p] = (1.0 - relax) * x[p] + relax * (b[p] - s) / c
    return x
