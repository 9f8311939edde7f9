"""The oracle's reference-layout memory with ghost layers (VertexDoFMemory.hpp:44-64 filled by the communicate chain
of VertexDoFPackInfo.hpp), pinned geometrically: for a linear function u, element k of the memory must hold u at the
lattice point the layout definition assigns to k (tests/halo_layout.py), and the memory has the reference's size."""
import numpy as np
import pytest

import oracle_lib as ol
from halo_layout import expected_points

A = np.array([0.7, -1.3, 2.1])


def u(p):
    return 0.25 + p @ A


@pytest.mark.parametrize("name", ["cube_24el.msh", "cube_kuhn_6el.msh", "pyramid_tilted_4el.msh",
                                  "three_tets_with_two_joint_faces.msh"])
@pytest.mark.parametrize("L", [0, 1, 2, 3])
def test_oracle_halo_memory_holds_the_expected_points(name, L):
    m = ol.Mesh(name)
    nv, ne, nf, nc = (int(v) for v in m.info()[:4])
    edges, faces = m.prims()
    cells = m.cells()
    xyz0, _ = m.coords(0)
    pts = expected_points(xyz0[:nv], [list(c) for c in cells], [tuple(e) for e in edges], [tuple(f) for f in faces], L)
    # sizes of VertexDoFMemory.hpp:44-64
    N, w = 2 ** L, 2 ** L + 1
    vdeg = np.bincount(edges.reshape(-1), minlength=nv)
    efaces = [sum(1 for t in faces if e[0] in t and e[1] in t) for e in edges]
    ecells = [sum(1 for c in cells if e[0] in c and e[1] in c) for e in edges]
    fcells = [sum(1 for c in cells if all(v in c for v in t)) for t in faces]
    size = (int(np.sum(1 + vdeg)) + sum(w + a * (w - 1) + b * (w - 2) for a, b in zip(efaces, ecells))
            + sum(w * (w + 1) // 2 + k * N * (N + 1) // 2 for k in fcells))
    assert len(pts) == size
    f = ol.Function(m, L, L)
    xyz, _ = m.coords(L)
    f.set(L, u(xyz))
    out = np.zeros(size)
    ol.lib().or_halo_pack(f.h, L, out)
    assert np.abs(out - u(pts)).max() < 1e-12
