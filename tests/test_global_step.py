"""Host FP64 global step (Subpb2Quadtree) against closed forms and invariants."""
import numpy as np
import pytest

from staq_dic_amd.global_step import Subpb2Quadtree, assemble_subpb2


def _square(h=3.0):
    coords = np.array([[0.0, 0.0], [h, 0.0], [h, h], [0.0, h]])
    el = np.array([[1, 2, 3, 4, 0, 0, 0, 0]])
    return coords, el


@pytest.mark.parametrize("order", [2, 3])
def test_element_mass_matrix_closed_form(order):
    """beta = 0: A is mu times the bilinear mass matrix h^2/36 [4 2 1 2; ...]."""
    h = 3.0
    coords, el = _square(h)
    U = np.zeros(8)
    F = np.zeros(16)
    A, _ = assemble_subpb2(coords, el, order, 0.0, 2.0, U, F)
    M = h * h / 36 * np.array([[4, 2, 1, 2], [2, 4, 2, 1], [1, 2, 4, 2], [2, 1, 2, 4]])
    Ad = A.toarray()[:8, :8]
    np.testing.assert_allclose(Ad[0::2, 0::2], 2.0 * M, atol=1e-12)
    np.testing.assert_allclose(Ad[1::2, 1::2], 2.0 * M, atol=1e-12)
    np.testing.assert_allclose(Ad[0::2, 1::2], 0.0, atol=1e-14)


def test_element_stiffness_matrix_closed_form():
    """mu = 0: A is beta times the bilinear Laplacian 1/6 [4 -1 -2 -1; ...],
    the same for any square size."""
    for h in (1.0, 5.0):
        coords, el = _square(h)
        A, _ = assemble_subpb2(coords, el, 2, 1.5, 0.0, np.zeros(8), np.zeros(16))
        K = np.array([[4, -1, -2, -1], [-1, 4, -1, -2], [-2, -1, 4, -1], [-1, -2, -1, 4]]) / 6
        np.testing.assert_allclose(A.toarray()[:8, :8][0::2, 0::2], 1.5 * K, atol=1e-12)


def test_rhs_uses_the_test_nodes_own_F():
    """The F term is beta * (dN_a/dx F11(a) + dN_a/dy F12(a)) with node a's own
    F (Subpb2Quadtree.m:262): with F nonzero at node 1 only, only node 1's rows
    see it."""
    coords, el = _square(2.0)
    F = np.zeros(16)
    F[0:4] = [0.3, -0.1, 0.2, 0.05]
    _, b = assemble_subpb2(coords, el, 2, 1.0, 0.0, np.zeros(8), F)
    # integral of dN1/dx and dN1/dy over the square [0,2]^2: -h/2 each
    np.testing.assert_allclose(b[0], -1.0 * 0.3 - 1.0 * 0.2, atol=1e-12)
    np.testing.assert_allclose(b[1], -1.0 * -0.1 - 1.0 * 0.05, atol=1e-12)
    np.testing.assert_allclose(b[2:8], 0.0, atol=1e-14)


@pytest.mark.parametrize("order", [2, 3, 4])
def test_patch_linear_field_on_reference_mesh(ref_frames, order):
    """A linear displacement with its constant gradient is reproduced exactly
    on the reference's quadtree mesh (938 elements, 200 hanging mid-nodes)."""
    mesh = ref_frames["mesh"]
    c = mesh["coordinatesFEM"]
    N = c.shape[0]
    B = np.array([[0.01, 0.003], [-0.002, 0.02]])
    a0 = np.array([1.5, -0.7])
    U = np.stack([a0[0] + c @ B[0], a0[1] + c @ B[1]], axis=1).ravel()
    F = np.tile([B[0, 0], B[1, 0], B[0, 1], B[1, 1]], N)
    mu = 1e-3
    Uh = Subpb2Quadtree(mesh, order, 0.256 * mu, mu, U, F, 0 * F, 0 * U, 0, 16)
    inv = np.unique(mesh["elementsFEM"])[1:]
    idx = np.stack([2 * inv - 2, 2 * inv - 1], axis=1).ravel()
    np.testing.assert_allclose(Uh[idx], U[idx], atol=1e-10)


def test_beta_zero_returns_U_and_duals_are_ignored(ref_frames):
    """beta = 0 leaves only the mass term: Uhat = U on every involved node;
    the duals are multiplied by zero in the reference (:11-12)."""
    mesh = dict(ref_frames["mesh"])
    mesh["dirichlet"] = np.zeros(0, dtype=np.int64)
    rng = np.random.default_rng(0)
    N = mesh["coordinatesFEM"].shape[0]
    U = rng.normal(size=2 * N)
    F = rng.normal(size=4 * N)
    Uh = Subpb2Quadtree(mesh, 2, 0.0, 1e-3, U, F, rng.normal(size=4 * N), rng.normal(size=2 * N),
                        0, 16)
    inv = np.unique(mesh["elementsFEM"])[1:]
    idx = np.stack([2 * inv - 2, 2 * inv - 1], axis=1).ravel()
    np.testing.assert_allclose(Uh[idx], U[idx], atol=1e-9)


def test_dirichlet_nodes_hold_U(ref_frames):
    mesh = ref_frames["mesh"]
    N = mesh["coordinatesFEM"].shape[0]
    rng = np.random.default_rng(1)
    U = rng.normal(size=2 * N)
    F = rng.normal(scale=0.01, size=4 * N)
    Uh = Subpb2Quadtree(mesh, 2, 0.3, 1e-3, U, F, 0 * F, 0 * U, 0, 16)
    d = np.unique(mesh["dirichlet"])
    np.testing.assert_array_equal(Uh[2 * d - 2], U[2 * d - 2])
    np.testing.assert_array_equal(Uh[2 * d - 1], U[2 * d - 1])
    A, _ = assemble_subpb2(mesh["coordinatesFEM"], mesh["elementsFEM"], 2, 0.3, 1e-3, U, F)
    assert abs(A - A.T).max() < 1e-12  # symmetric
