"""A copy of meshes/cube_24el.msh whose boundary elements carry physical tags (test data builder).

The file's own line and triangle elements all have physical tag 0; the copy gives the triangles of the x = 0 side
tag 1 (Dirichlet under create0123BC), x = 1 tag 2 (Neumann), y = 0 tag 3 (free-slip) and keeps 0 on the other sides;
the twelve cube edges get tag 1 if they lie in x = 0, else 0.
"""
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FACE_TAG = {1: 1, 2: 2, 3: 3}  # elementary entity of a side triangle -> physical tag


def write_tagged_cube(path):
    lines = open(os.path.join(ROOT, "meshes", "cube_24el.msh")).read().split("\n")
    i = lines.index("$Nodes")
    n = int(lines[i + 1])
    xyz = {int(l.split()[0]): tuple(float(v) for v in l.split()[1:4]) for l in lines[i + 2:i + 2 + n]}
    j = lines.index("$Elements")
    m = int(lines[j + 1])
    out = list(lines)
    tagged_edges = 0
    for k in range(j + 2, j + 2 + m):
        f = lines[k].split()
        if f[1] == "2":
            f[3] = str(FACE_TAG.get(int(f[4]), 0))
        elif f[1] == "1":
            a, b = int(f[5]), int(f[6])
            on_x0 = xyz[a][0] == 0.0 and xyz[b][0] == 0.0
            f[3] = "1" if on_x0 else "0"
            tagged_edges += on_x0
        out[k] = " ".join(f)
    with open(path, "w") as fh:
        fh.write("\n".join(out))
    return tagged_edges
