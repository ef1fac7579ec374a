"""Mesh input of the reference's particle set-ups (module Mesh): the Gmsh v2
reader (Mesh.jl:634-714), the Q4 rectangle mesher (Mesh.jl:523-567) and the
Q4 Jacobian that orients CPDI domains (Mesh.jl:84-112).  Host-side: they
produce the corner coordinates and connectivity that jump_add_solid takes
for CPDI-Q4 solids (Solid.jl:162-222, Solid2D(fileName, mat))."""
import numpy as np


def loadGMSH(path):
    """Mesh.jl:634-714.  Returns (coords, elems): coords (3, nNodes) in file
    order (node ids are taken to be 1..nNodes, as the reference does), elems
    (nElem, 4) or (nElem, 8) 1-based node numbers of the accepted element
    types 3 (4-node quad), 4 (4-node tetrahedron) and 5 (8-node hexahedron),
    in file order; other types (lines, points) are skipped.

    Kept from the reference: a section is found by a line that occurs
    inside "$Nodes" / "$Elements" (occursin(line, tag)); the node list of an
    element starts at its sixth field (two tags assumed)."""
    with open(path) as f:
        lines = f.read().splitlines()
    n_nodes = n_elems = 0
    for i, ln in enumerate(lines):
        if ln in "$Nodes":
            n_nodes = int(lines[i + 1])
        if ln in "$Elements":
            n_elems = int(lines[i + 1])
    coords = np.empty((3, n_nodes))
    elems = []
    for i, ln in enumerate(lines):
        if ln in "$Nodes":
            for k in range(n_nodes):
                w = lines[i + 2 + k].split()
                coords[:, k] = [float(w[1]), float(w[2]), float(w[3])]
        if ln in "$Elements":
            for k in range(n_elems):
                w = [int(v) for v in lines[i + 2 + k].split()]
                etype = w[1]
                if etype in (3, 4):
                    elems.append(w[5:9])
                elif etype == 5:
                    elems.append(w[5:13])
    width = max((len(e) for e in elems), default=4)
    if any(len(e) != width for e in elems):  # the reference's vcat fails alike
        raise ValueError(f"{path}: elements with different node counts")
    return coords, np.array(elems, dtype=np.int64).reshape(-1, width)


def quad4_detJ(xi, eta, corners):
    """lagrange_basis!(N, ::Quad4, xieta, coords) (Mesh.jl:84-112): the
    Jacobian determinant of the bilinear map at (xi, eta); corners (4, 2)."""
    d = [(-0.25 * (1 - eta), -0.25 * (1 - xi)), (0.25 * (1 - eta), -0.25 * (1 + xi)),
         (0.25 * (1 + eta), 0.25 * (1 + xi)), (-0.25 * (1 + eta), 0.25 * (1 - xi))]
    c = [(float(p[0]), float(p[1])) for p in corners]
    J11 = d[0][0] * c[0][0] + d[1][0] * c[1][0] + d[2][0] * c[2][0] + d[3][0] * c[3][0]
    J12 = d[0][0] * c[0][1] + d[1][0] * c[1][1] + d[2][0] * c[2][1] + d[3][0] * c[3][1]
    J21 = d[0][1] * c[0][0] + d[1][1] * c[1][0] + d[2][1] * c[2][0] + d[3][1] * c[3][0]
    J22 = d[0][1] * c[0][1] + d[1][1] * c[1][1] + d[2][1] * c[2][1] + d[3][1] * c[3][1]
    return J11 * J22 - J12 * J21


def _linrange(a, b, n):
    """Julia LinRange(a, b, n): ((d-j)/d) a + (j/d) b, d = n-1."""
    d = n - 1
    return [a if d == 0 else ((d - j) / d) * a + (j / d) * b for j in range(n)]


def square_node_array(pt1, pt2, pt3, pt4, numnod_u, numnod_v):
    """Mesh.jl:523-542: nodes of the bilinear patch pt1..pt4, (2, nu*nv),
    u fastest.  The node is dot(x_pts, N) with N the Q4 functions
    0.25 (1+-xi)(1+-eta); the dot product is summed in corner order (the
    reference calls BLAS ddot, whose order for four terms is the library's)."""
    xs = [pt1[0], pt2[0], pt3[0], pt4[0]]
    ys = [pt1[1], pt2[1], pt3[1], pt4[1]]
    X = np.zeros((2, numnod_u * numnod_v))
    for r, eta in enumerate(_linrange(-1.0, 1.0, numnod_v)):
        for c, xi in enumerate(_linrange(-1.0, 1.0, numnod_u)):
            N = [0.25 * ((1 - xi) * (1 - eta)), 0.25 * ((1 + xi) * (1 - eta)), 0.25 * ((1 + xi) * (1 + eta)),
                 0.25 * ((1 - xi) * (1 + eta))]
            X[0, r * numnod_u + c] = ((xs[0] * N[0] + xs[1] * N[1]) + xs[2] * N[2]) + xs[3] * N[3]
            X[1, r * numnod_u + c] = ((ys[0] * N[0] + ys[1] * N[1]) + ys[2] * N[2]) + ys[3] * N[3]
    return X


def make_elem(num_u, num_v):
    """Mesh.jl:545-563: Q4 connectivity (1-based), counter-clockwise from
    the lower-left node, row by row."""
    nnx = num_u + 1
    pattern = np.array([1, 2, nnx + 2, nnx + 1], dtype=np.int64)
    out = np.zeros((num_u * num_v, 4), dtype=np.int64)
    e = 0
    inc = 0
    for row in range(1, num_v + 1):
        for _ in range(num_u):
            out[e] = pattern + inc
            inc += 1
            e += 1
        inc = row * nnx
    return out


def createMeshForRectangle(pt1, pt2, pt3, pt4, elemX, elemY):
    """Mesh.jl:565-569: (nodes (2, n), elems (elemX*elemY, 4) 1-based)."""
    return square_node_array(pt1, pt2, pt3, pt4, elemX + 1, elemY + 1), make_elem(elemX, elemY)


def orient_q4(nodes2, elems):
    """Solid.jl:189-200: an element whose Jacobian at its centre is negative
    has its second and fourth corners swapped (in place); nodes2 (2, n),
    elems 1-based."""
    for e in range(elems.shape[0]):
        c = nodes2[:, elems[e] - 1].T
        if quad4_detJ(0.0, 0.0, c) < 0.0:
            elems[e, 1], elems[e, 3] = elems[e, 3], elems[e, 1]
    return elems
