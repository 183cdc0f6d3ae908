"""Quadtree mesh producer (MeshSetUp, GenerateQuadtreeMesh with QrefineR,
funMarkEdge, funMarkInside) and the initial-guess transfer.

The reference's own result file holds the mesh its driver built for the
circular-hole frames (tests/golden/circular_hole_000_001.npz: 1434 nodes, 938
elements with their hanging mid-nodes, the 1174 nodes marked near the hole and
the border): the mesh produced here from the same mask must equal it entry for
entry.  QrefineR is also checked on small meshes (node counts, area, hanging
nodes), and the transfer on a linear field (the thin-plate RBF reproduces it).
"""
import os

import numpy as np

from staq_dic_amd import mesh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _fixture():
    return np.load(os.path.join(ROOT, "tests", "golden", "circular_hole_000_001.npz"))


def test_uniform_mesh_is_the_references_first_nodes():
    z = _fixture()
    xn, yn = mesh.uniform_nodes(400, 400, float(z["winsize"]), float(z["winstepsize"]))
    c0, e0 = mesh.MeshSetUp(xn, yn)
    assert c0.shape == (529, 2) and e0.shape == (484, 4)
    np.testing.assert_array_equal(c0, z["coordinatesFEM"][:529])


def test_quadtree_mesh_equals_the_references():
    z = _fixture()
    xn, yn = mesh.uniform_nodes(400, 400, 32, 16)
    c0, e0 = mesh.MeshSetUp(xn, yn)
    c, E, irr, mark = mesh.GenerateQuadtreeMesh(c0, e0, z["ImgRefMask"], 16, 32, elementMinSize=2)
    np.testing.assert_array_equal(c, z["coordinatesFEM"])
    np.testing.assert_array_equal(E, z["elementsFEM"])
    np.testing.assert_array_equal(mark, z["markCoordHoleEdge"])
    np.testing.assert_array_equal(mark, z["dirichlet"])
    assert irr.shape == (352, 3)


def _area(c, E):
    p = c[E[:, :4] - 1]
    x, y = p[..., 0], p[..., 1]
    return 0.5 * np.abs(np.sum(x * np.roll(y, -1, 1) - np.roll(x, -1, 1) * y, axis=1))


def test_qrefiner_small_meshes():
    c0, e0 = mesh.MeshSetUp(np.array([0.0, 8, 16]), np.array([0.0, 8, 16]))
    # every element: 16 children, 25 nodes, no hanging node, area kept
    c, E, irr = mesh.QrefineR(c0, e0, np.zeros((0, 3)), np.arange(1, 5))
    assert c.shape[0] == 25 and E.shape[0] == 16 and irr.shape[0] == 0
    assert abs(_area(c, E).sum() - 256.0) < 1e-12
    assert len(np.unique(c, axis=0)) == 25
    # one element: its two edges shared with neighbours carry hanging nodes
    c, E, irr = mesh.QrefineR(c0, e0, np.zeros((0, 3)), np.array([1]))
    assert E.shape[0] == 7 and irr.shape[0] == 2
    assert abs(_area(c, E).sum() - 256.0) < 1e-12
    for a, b, m in irr:
        np.testing.assert_array_equal(c[m - 1], 0.5 * (c[a - 1] + c[b - 1]))
    # refining a neighbour of a hanging node first resolves it (no irregular left there)
    c2, E2, irr2 = mesh.QrefineR(c, E, irr, np.arange(5, 8))
    assert abs(_area(c2, E2).sum() - 256.0) < 1e-12
    assert len(np.unique(c2, axis=0)) == c2.shape[0]


def test_transfer_reproduces_a_linear_field():
    z = _fixture()
    xn, yn = mesh.uniform_nodes(400, 400, 32, 16)
    c0, e0 = mesh.MeshSetUp(xn, yn)
    c, E, irr, mark = mesh.GenerateQuadtreeMesh(c0, e0, z["ImgRefMask"], 16, 32)
    U0 = np.empty(2 * len(c0))
    U0[0::2] = 0.3 + 0.01 * c0[:, 0] - 0.002 * c0[:, 1]
    U0[1::2] = -1.0 + 0.004 * c0[:, 0] + 0.003 * c0[:, 1]
    U = mesh.transfer_initial_guess(U0, c0, c, z["ImgRefMask"])
    used = np.unique(E[:, :4])  # nodes of elements outside the hole
    ref_u = 0.3 + 0.01 * c[used - 1, 0] - 0.002 * c[used - 1, 1]
    ref_v = -1.0 + 0.004 * c[used - 1, 0] + 0.003 * c[used - 1, 1]
    np.testing.assert_allclose(U[2 * (used - 1)], ref_u, atol=1e-8)
    np.testing.assert_allclose(U[2 * (used - 1) + 1], ref_v, atol=1e-8)
