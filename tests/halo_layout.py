"""Expected contents of the reference's per-primitive vertex-DoF memory with ghost layers (test model).

Built from the layout's definition alone (p1functionspace/VertexDoFMemory.hpp:44-64 sizes, the ghost points the
communicate chain of VertexDoFPackInfo.hpp delivers): for every element, the barycentric position of the lattice
point it holds, as a point in space. A linear function's value there is then known exactly, which pins any
implementation of the layout that fills the memory from a function's values.
"""
import numpy as np


def expected_points(xyz_vertices, cells, edges, faces, L):
    """List of 3-D points, one per memory element (vertices, edges, faces in storage order)."""
    N = 2 ** L
    X = np.asarray(xyz_vertices, dtype=float)
    pts = []

    def bary(verts, weights):
        return sum(w * X[v] for v, w in zip(verts, weights)) / N

    for v in range(len(X)):
        pts.append(X[v])
        for e in range(len(edges)):
            if v in edges[e]:
                other = edges[e][1] if edges[e][0] == v else edges[e][0]
                pts.append(bary([v, other], [N - 1, 1]))
    for e, (a, b) in enumerate(edges):
        for i in range(N + 1):
            pts.append(bary([a, b], [N - i, i]))
        for f, t in enumerate(faces):
            if a in t and b in t:
                op = [v for v in t if v != a and v != b][0]
                for i in range(N):
                    pts.append(bary([a, b, op], [N - 1 - i, i, 1]))
        for c, cv in enumerate(cells):
            if a in cv and b in cv:
                o = [v for v in cv if v != a and v != b]
                for i in range(N - 1):
                    pts.append(bary([a, b, o[0], o[1]], [N - 2 - i, i, 1, 1]))
    for f, t in enumerate(faces):
        for y in range(N + 1):
            for x in range(N + 1 - y):
                pts.append(bary(t, [N - x - y, x, y]))
        for c, cv in enumerate(cells):
            if all(v in cv for v in t):
                fourth = [v for v in cv if v not in t][0]
                for y in range(N):
                    for x in range(N - y):
                        pts.append(bary(list(t) + [fourth], [N - 1 - x - y, x, y, 1]))
    return np.array(pts)
