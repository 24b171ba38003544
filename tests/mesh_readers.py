"""Minimal readers for the reference's test meshes (test infrastructure).

gmsh 2.2 ascii (triangles + tetrahedra) and UGRID ascii (tri/quad boundary
faces, tet/pyramid/prism/hex volumes), both 1-based on disk.
"""
import numpy as np

TRI, QUAD, TET, PYRAMID, PRISM, HEX = 0, 1, 2, 3, 4, 5


def read_gmsh2(path):
    lines = open(path).read().split("\n")
    i = lines.index("$Nodes")
    nn = int(lines[i + 1])
    xyz = np.array([[float(v) for v in lines[i + 2 + k].split()[1:4]] for k in range(nn)])
    j = lines.index("$Elements")
    ne = int(lines[j + 1])
    elems, bfaces = [], []
    for k in range(ne):
        t = [int(v) for v in lines[j + 2 + k].split()]
        etype, ntag = t[1], t[2]
        tags = t[3:3 + ntag]
        nodes = [v - 1 for v in t[3 + ntag:]]
        if etype == 2:
            bfaces.append((TRI, nodes, tags[0] if tags else 1))
        elif etype == 4:
            elems.append((TET, nodes))
        else:
            raise ValueError(f"gmsh element type {etype} not handled")
    return xyz, elems, bfaces


def read_ugrid(path):
    tok = open(path).read().split()
    p = 0

    def take(n, f=int):
        nonlocal p
        out = [f(v) for v in tok[p:p + n]]
        p += n
        return out

    nn, ntri, nquad, ntet, npyr, npri, nhex = take(7)
    xyz = np.array(take(3 * nn, float)).reshape(nn, 3)
    tris = np.array(take(3 * ntri)).reshape(ntri, 3) - 1
    quads = np.array(take(4 * nquad)).reshape(nquad, 4) - 1
    tags = take(ntri + nquad)
    bfaces = [(TRI, list(t), tags[k]) for k, t in enumerate(tris)]
    bfaces += [(QUAD, list(q), tags[ntri + k]) for k, q in enumerate(quads)]
    elems = []
    for n, et, nv in [(ntet, TET, 4), (npyr, PYRAMID, 5), (npri, PRISM, 6), (nhex, HEX, 8)]:
        conn = np.array(take(nv * n)).reshape(n, nv) - 1
        if et == PYRAMID:
            # UGRID pyramid numbering -> base 0..3 + apex 4 (ucs/mesh.tcc:6750-6758)
            conn = conn[:, [0, 3, 4, 1, 2]]
        elems += [(et, list(c)) for c in conn]
    return xyz, elems, bfaces
