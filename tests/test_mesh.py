"""The mesh input of CPDI set-ups (jump_amd/mesh.py): the Gmsh v2 reader
(Mesh.jl:634-714) on a hand-written file and on the reference's disk.msh,
ring.msh and tube.msh (tests/golden/, input files of Main_TwoDisksCPDI.jl,
Main_TwoRings_CPDI.jl and Main_LateralCompressionTube_CPDI.jl), the Q4
orientation fix and areas of Solid2D(fileName, mat) (Solid.jl:162-222), and
the rectangle mesher (Mesh.jl:523-569) against closed forms."""
import math
import os

import numpy as np
import pytest

from jump_amd import material, mesh, solid

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

MSH = """$MeshFormat
2.2 0 8
$EndMeshFormat
$Nodes
{nn}
1 0 0 0
2 1 0 0
3 1 1 0
4 0 1 0
5 2 0 0
6 2 1 0
7 0 0 1
8 1 0 1
9 1 1 1
10 0 1 1
$EndNodes
$Elements
{ne}
{els}$EndElements
"""


def _msh(els):
    return MSH.format(nn=10, ne=len(els), els="".join(e + "\n" for e in els))


def test_gmsh_reader_on_hand_written_files(tmp_path):
    p = tmp_path / "q.msh"
    p.write_text(_msh(["1 15 2 0 1 1", "2 1 2 1 1 1 2", "3 3 2 2 2 1 2 3 4", "4 3 2 2 2 5 6 3 2",
                       "5 4 2 3 3 1 2 4 7"]))
    coords, elems = mesh.loadGMSH(str(p))
    assert coords.shape == (3, 10)
    assert np.array_equal(coords[:, 4], [2.0, 0.0, 0.0]) and np.array_equal(coords[:, 9], [0.0, 1.0, 1.0])
    # the point (15) and the line (1) are skipped; quads and the tet keep file order
    assert elems.tolist() == [[1, 2, 3, 4], [5, 6, 3, 2], [1, 2, 4, 7]]
    p = tmp_path / "h.msh"
    p.write_text(_msh(["1 5 2 4 4 1 2 3 4 7 8 9 10"]))
    _, elems = mesh.loadGMSH(str(p))
    assert elems.tolist() == [[1, 2, 3, 4, 7, 8, 9, 10]]


def _area(c):
    x, y = c[:, 0], c[:, 1]
    return 0.5 * (x[0] * y[1] - x[1] * y[0] + x[1] * y[2] - x[2] * y[1] + x[2] * y[3] - x[3] * y[2]
                  + x[3] * y[0] - x[0] * y[3])


def test_quad4_jacobian_and_orientation():
    ccw = np.array([[0.0, 0.0], [2.0, 0.0], [3.0, 1.0], [1.0, 1.0]])  # parallelogram, area 2
    assert mesh.quad4_detJ(0.0, 0.0, ccw) == 0.5  # area / 4 for the bilinear map
    assert mesh.quad4_detJ(0.0, 0.0, ccw[::-1]) == -0.5
    nodes = np.array([[0.0, 1.0, 1.0, 0.0], [0.0, 0.0, 1.0, 1.0]])
    elems = np.array([[1, 2, 3, 4], [1, 4, 3, 2]])  # second one clockwise
    mesh.orient_q4(nodes, elems)
    assert elems.tolist() == [[1, 2, 3, 4], [1, 2, 3, 4]]


def test_disk_mesh_solid():
    m = material.ElasticMaterial(1000.0, 0.3, 1000.0, 0, 0, 1000)
    s = solid.Solid2D(os.path.join(GOLD, "disk.msh"), m)
    coords, elems = mesh.loadGMSH(os.path.join(GOLD, "disk.msh"))
    assert coords.shape == (3, 450) and elems.shape == (417, 4)
    assert s.pos.shape == (417, 2) and s.elems.shape == (417, 4) and s.elems.min() == 0
    areas = np.array([_area(s.nodes[e]) for e in s.elems])
    assert np.all(areas > 0.0)
    assert np.array_equal(s.volume, areas) and np.array_equal(s.mass, areas * 1000.0)
    assert np.array_equal(s.pos, s.nodes[s.elems].mean(axis=1))
    total = areas.sum()  # a polygon inscribed in the r = 0.2 circle
    assert 0.99 * math.pi * 0.04 < total < math.pi * 0.04
    assert np.all(np.abs((s.pos * s.mass[:, None]).sum(0) / s.mass.sum()) < 1e-3)
    solid.move_cpdi(s, (0.45, 0.45))
    assert np.array_equal(s.pos, s.nodes[s.elems].mean(axis=1)) or \
        np.max(np.abs(s.pos - s.nodes[s.elems].mean(axis=1))) < 1e-15


@pytest.mark.parametrize("name,centre,ri,ro,nodes,nel", [
    ("ring.msh", (0.0, 0.0), 30.0, 40.0, 3108, 2884),           # Main_TwoRings_CPDI.jl
    ("tube.msh", (50.0, 30.0), 47.9 / 2, 47.9 / 2 + 1.48, 8679, 7898),  # Main_LateralCompressionTube_CPDI.jl
])
def test_annulus_meshes(name, centre, ri, ro, nodes, nel):
    """The reference's ring and tube meshes: every Q4 positive after the
    orientation fix, every element node on or between the two circles (the
    files also hold the arcs' centre points, which no element uses), and the
    areas summing to the annulus less the inscribed polygons' deficit."""
    m = material.ElasticMaterial(1000.0, 0.3, 1000.0, 0, 0, 1000)
    s = solid.Solid2D(os.path.join(GOLD, name), m)
    coords, elems = mesh.loadGMSH(os.path.join(GOLD, name))
    assert coords.shape == (3, nodes) and elems.shape == (nel, 4)
    areas = np.array([_area(s.nodes[e]) for e in s.elems])
    assert np.all(areas > 0.0) and np.array_equal(s.volume, areas)
    r = np.hypot(*(s.nodes[np.unique(s.elems)] - np.asarray(centre)).T)
    assert r.min() > ri - 1e-9 * ro and r.max() < ro + 1e-9 * ro
    ann = math.pi * (ro * ro - ri * ri)
    assert ann * (1.0 - 1e-5) < areas.sum() <= ann * (1.0 + 1e-12)


def test_rectangle_mesher():
    X, E = mesh.createMeshForRectangle((0.0, 0.0), (2.0, 0.0), (2.0, 1.0), (0.0, 1.0), 4, 2)
    assert X.shape == (2, 15) and E.shape == (8, 4)
    xs, ys = np.meshgrid(np.linspace(0, 2, 5), np.linspace(0, 1, 3))
    assert np.array_equal(X[0], xs.ravel()) and np.array_equal(X[1], ys.ravel())
    assert E[0].tolist() == [1, 2, 7, 6] and E[4].tolist() == [6, 7, 12, 11]  # [1 2 nnx+2 nnx+1] + inc
    areas = [_area(X[:, e - 1].T) for e in E]
    assert np.allclose(areas, 0.25) and math.isclose(sum(areas), 2.0)
    # a skewed patch: nodes are the bilinear map of the (xi, eta) lattice
    X2, _ = mesh.createMeshForRectangle((0.0, 0.0), (4.0, 1.0), (5.0, 3.0), (-1.0, 2.0), 3, 3)
    assert np.allclose(X2[:, 0], [0.0, 0.0]) and np.allclose(X2[:, 3], [4.0, 1.0])
    assert np.allclose(X2[:, 15], [5.0, 3.0]) and np.allclose(X2[:, 12], [-1.0, 2.0])
