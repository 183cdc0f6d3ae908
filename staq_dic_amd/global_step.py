"""Global step of the augmented-Lagrangian loop on the host, in FP64.

``Subpb2Quadtree`` (func_quadtree/Subpb2Quadtree.m:1-354) minimises
beta ||grad(Uhat) - F||^2 + mu ||Uhat - U||^2 (+ alpha ||grad(Uhat - U)||^2)
over the quadtree finite-element mesh: 4-node quadrilaterals whose edges may
carry a hanging mid-node (element columns 5-8, 0 = none), Gauss quadrature of
order 2-5, Dirichlet nodes held at U, one sparse solve.  It is the consumer of
the local steps' (U, F) and stays on the host, as in the reference.

The restatement keeps the reference's discretisation exactly, including:
  * the duals enter multiplied by zero (v = 0*vdual, W = 0*udual, :11-12), so
    U - v = U and F - W = F;
  * the F term of the right-hand side uses the nodal F of the test function's
    own node: beta * diag(DN' * F(tempIndexF)') (:259-262);
  * the involved-DOF list drops the smallest node index of
    unique(elementsFEM(:, 1:8)) on the assumption that it is the 0 of a missing
    mid-node (:273-276);
  * missing mid-nodes point at a padding node N+1 whose DOFs stay zero (:128-136).
Element matrices are built for all elements and Gauss points at once with
NumPy and assembled with scipy.sparse; the solve is scipy's sparse LU.  An AL
loop calls the global step and the nodal strain many times on one mesh, so
the parts that depend on the mesh and the weights alone (the factorised
matrix, the natural-neighbour map of the nodal strain) are kept in a small
cache keyed by their inputs' bytes; every call still assembles its own
right-hand side.
"""
from __future__ import annotations

import hashlib
from collections import OrderedDict

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

__all__ = ["Subpb2Quadtree", "gauss_points", "assemble_subpb2", "gauss_point_strain",
           "funGlobalNodalStrainQuadtree"]

_CACHE: "OrderedDict[bytes, object]" = OrderedDict()
_CACHE_SIZE = 8


def _digest(*parts):
    h = hashlib.blake2b(digest_size=20)
    for p in parts:
        if isinstance(p, np.ndarray):
            h.update(str((p.dtype.str, p.shape)).encode())
            h.update(np.ascontiguousarray(p).tobytes())
        else:
            h.update(repr(p).encode())
    return h.digest()


def _cached(key, build):
    if key in _CACHE:
        _CACHE.move_to_end(key)
        return _CACHE[key]
    val = build()
    _CACHE[key] = val
    while len(_CACHE) > _CACHE_SIZE:
        _CACHE.popitem(last=False)
    return val


def gauss_points(order):
    """Points and weights of Subpb2Quadtree.m:20-43 (the reference's own
    rounded constants for orders 4 and 5)."""
    if order == 2:
        return np.array([-1 / np.sqrt(3), 1 / np.sqrt(3)]), np.array([1.0, 1.0])
    if order == 3:
        return np.array([0.0, np.sqrt(3 / 5), -np.sqrt(3 / 5)]), np.array([8 / 9, 5 / 9, 5 / 9])
    if order == 4:
        return (np.array([0.339981, -0.339981, 0.861136, -0.861136]),
                np.array([0.652145, 0.652145, 0.347855, 0.347855]))
    if order == 5:
        return (np.array([0.0, 0.538469, -0.538469, 0.90618, -0.90618]),
                np.array([0.568889, 0.478629, 0.478629, 0.236927, 0.236927]))
    raise ValueError("GaussPtOrder must be 2, 3, 4 or 5")


def _shape(ksi, eta, dl):
    """N1..N8 and their ksi / eta derivatives (Subpb2Quadtree.m:147-157 and
    the funDN* helpers :355-410) for per-element mid-node flags dl (E x 4)."""
    d5, d6, d7, d8 = dl[:, 0], dl[:, 1], dl[:, 2], dl[:, 3]
    ak, ae = abs(ksi), abs(eta)
    sk, se = np.sign(-ksi), np.sign(-eta)
    N5 = d5 * 0.5 * (1 + ksi) * (1 - ae)
    N6 = d6 * 0.5 * (1 + eta) * (1 - ak)
    N7 = d7 * 0.5 * (1 - ksi) * (1 - ae)
    N8 = d8 * 0.5 * (1 - eta) * (1 - ak)
    N1 = (1 - ksi) * (1 - eta) * 0.25 - 0.5 * (N7 + N8)
    N2 = (1 + ksi) * (1 - eta) * 0.25 - 0.5 * (N8 + N5)
    N3 = (1 + ksi) * (1 + eta) * 0.25 - 0.5 * (N5 + N6)
    N4 = (1 - ksi) * (1 + eta) * 0.25 - 0.5 * (N6 + N7)
    k5, e5 = d5 * 0.5 * (1 - ae), d5 * 0.5 * (1 + ksi) * se
    k6, e6 = d6 * 0.5 * (1 + eta) * sk, d6 * 0.5 * (1 - ak)
    k7, e7 = d7 * 0.5 * (-1) * (1 - ae), d7 * 0.5 * (1 - ksi) * se
    k8, e8 = d8 * 0.5 * (1 - eta) * sk, d8 * 0.5 * (-1) * (1 - ak)
    k1, e1 = -0.25 * (1 - eta) - 0.5 * (k7 + k8), -0.25 * (1 - ksi) - 0.5 * (e7 + e8)
    k2, e2 = 0.25 * (1 - eta) - 0.5 * (k8 + k5), -0.25 * (1 + ksi) - 0.5 * (e8 + e5)
    k3, e3 = 0.25 * (1 + eta) - 0.5 * (k5 + k6), 0.25 * (1 + ksi) - 0.5 * (e5 + e6)
    k4, e4 = -0.25 * (1 + eta) - 0.5 * (k6 + k7), 0.25 * (1 - ksi) - 0.5 * (e6 + e7)
    N = np.stack([N1, N2, N3, N4, N5, N6, N7, N8], axis=1)
    Dk = np.stack([k1, k2, k3, k4, k5, k6, k7, k8], axis=1)
    De = np.stack([e1, e2, e3, e4, e5, e6, e7, e8], axis=1)
    return N, Dk, De


def assemble_subpb2(coordinatesFEM, elementsFEM, GaussPtOrder, beta, mu, U, F, alpha=0.0,
                    need_A=True):
    """Global matrix A and right-hand side b of Subpb2Quadtree.m:70-264 (size
    2N+8, the last 8 entries being the padding DOFs); A is None when
    need_A is false."""
    coords = np.asarray(coordinatesFEM, dtype=np.float64)
    el = np.asarray(elementsFEM, dtype=np.int64)
    N = coords.shape[0]
    E = el.shape[0]
    n_dof = 2 * N + 8
    Up = np.concatenate([np.asarray(U, dtype=np.float64).ravel(), np.zeros(8)])
    Fp = np.concatenate([np.asarray(F, dtype=np.float64).ravel(), np.zeros(16)])
    dl = (el[:, 4:8] != 0).astype(np.float64)
    node = np.where(el > 0, el, N + 1)  # missing mid-nodes -> padding node N+1
    xy = np.zeros((E, 8, 2))
    has = el > 0
    xy[has] = coords[el[has] - 1]
    # DOFs (0-based) [u1 v1 ... u8 v8] and the nodal F of each node
    du = 2 * (node - 1)
    dof = np.stack([du, du + 1], axis=2).reshape(E, 16)
    f0 = 4 * (node - 1)
    Fx = np.stack([Fp[f0], Fp[f0 + 1]], axis=2)      # F11 = u_x, F21 = v_x
    Fy = np.stack([Fp[f0 + 2], Fp[f0 + 3]], axis=2)  # F12 = u_y, F22 = v_y
    Ue = Up[dof].reshape(E, 8, 2)

    pts, wts = gauss_points(GaussPtOrder)
    ksiA, etaA = np.meshgrid(pts, pts, indexing="ij")
    wkA, weA = np.meshgrid(wts, wts, indexing="ij")
    Ae = np.zeros((E, 8, 8, 2))  # per component block (u-u and v-v are equal)
    be = np.zeros((E, 8, 2))
    for ksi, eta, wk, we in zip(ksiA.ravel(order="F"), etaA.ravel(order="F"),
                                wkA.ravel(order="F"), weA.ravel(order="F")):
        Nn, Dk, De = _shape(ksi, eta, dl)
        J11 = np.einsum("ea,ea->e", Dk, xy[:, :, 0])
        J12 = np.einsum("ea,ea->e", Dk, xy[:, :, 1])
        J21 = np.einsum("ea,ea->e", De, xy[:, :, 0])
        J22 = np.einsum("ea,ea->e", De, xy[:, :, 1])
        det = J11 * J22 - J12 * J21
        i11, i12, i21, i22 = J22 / det, -J12 / det, -J21 / det, J11 / det
        Dx = i11[:, None] * Dk + i12[:, None] * De
        Dy = i21[:, None] * Dk + i22[:, None] * De
        s = (det * wk * we)[:, None, None]
        K = Dx[:, :, None] * Dx[:, None, :] + Dy[:, :, None] * Dy[:, None, :]
        M = Nn[:, :, None] * Nn[:, None, :]
        if need_A:
            blk = s * ((beta + alpha) * K + mu * M)
            Ae[..., 0] += blk
            Ae[..., 1] += blk
        for c in range(2):
            rhs = beta * (Dx * Fx[:, :, c] + Dy * Fy[:, :, c])
            rhs = rhs + mu * np.einsum("eab,eb->ea", M, Ue[:, :, c])
            rhs = rhs + alpha * np.einsum("eab,eb->ea", K, Ue[:, :, c])
            be[:, :, c] += s[:, :, 0] * rhs
    b = np.zeros(n_dof)
    np.add.at(b, dof.ravel(), be.reshape(E, 16).ravel())
    if not need_A:
        return None, b
    # interleave the two component blocks into 16 x 16 element matrices
    A16 = np.zeros((E, 8, 2, 8, 2))
    A16[:, :, 0, :, 0] = Ae[..., 0]
    A16[:, :, 1, :, 1] = Ae[..., 1]
    A16 = A16.reshape(E, 16, 16)
    rows = np.repeat(dof[:, :, None], 16, axis=2).ravel()
    cols = np.repeat(dof[:, None, :], 16, axis=1).ravel()
    A = sp.coo_matrix((A16.ravel(), (rows, cols)), shape=(n_dof, n_dof)).tocsr()
    return A, b


def Subpb2Quadtree(DICmesh, GaussPtOrder, beta, mu, U, F, udual, vdual, alpha, winstepsize,
                   waitBarDisplayOrNot=1):
    """[Uhat] = Subpb2Quadtree(DICmesh,GaussPtOrder,beta,mu,U,F,udual,vdual,
    alpha,winstepsize,waitBarDisplayOrNot).  DICmesh holds 'coordinatesFEM'
    (N x 2), 'elementsFEM' (E x 8, 1-based, 0 = no mid-node), 'dirichlet'
    (node list) and optionally 'neumann' (K x 4: node1, node2, nx, ny)."""
    coords = np.asarray(DICmesh["coordinatesFEM"], dtype=np.float64)
    el = np.asarray(DICmesh["elementsFEM"], dtype=np.int64)
    N = coords.shape[0]
    n_dof = 2 * N + 8
    U = np.asarray(U, dtype=np.float64).ravel()
    F = np.asarray(F, dtype=np.float64).ravel()
    dirichlet = np.unique(np.asarray(DICmesh.get("dirichlet", []), dtype=np.int64).ravel())

    def build():
        A, _ = assemble_subpb2(coords, el, GaussPtOrder, beta, mu, np.zeros(2 * N),
                               np.zeros(4 * N), alpha)
        # DOFs of the nodes the elements use, without the smallest index (:273-276)
        involved = np.unique(el[:, :8])[1:]
        uidx = np.stack([2 * involved - 2, 2 * involved - 1], axis=1).ravel()
        dir_dof = np.concatenate([2 * dirichlet - 1, 2 * dirichlet - 2]) if dirichlet.size else \
            np.zeros(0, dtype=np.int64)
        free = np.setdiff1d(uidx, dir_dof)
        lu = spla.splu(A[free][:, free].tocsc()) if free.size else None
        return A, free, lu

    A, free, lu = _cached(_digest("subpb2", coords, el, dirichlet, GaussPtOrder, float(beta),
                                  float(mu), float(alpha)), build)
    _, b = assemble_subpb2(coords, el, GaussPtOrder, beta, mu, U, F, alpha, need_A=False)

    # Neumann edges with traction -F / winstepsize (:279-287)
    neumann = np.asarray(DICmesh.get("neumann", np.zeros((0, 4))), dtype=np.float64)
    if neumann.size:
        Fp = np.concatenate([F, np.zeros(16)])
        bc = -Fp / winstepsize
        for row in neumann.reshape(-1, neumann.shape[-1]):
            n1, n2 = int(row[0]), int(row[1])
            ln = 0.5 * np.linalg.norm(coords[n1 - 1] - coords[n2 - 1])
            for n in (n1, n2):
                b[2 * n - 2] += ln * (bc[4 * n - 4] * row[2] + bc[4 * n - 2] * row[3])
            for n in (n1, n2):
                b[2 * n - 1] += ln * (bc[4 * n - 3] * row[2] + bc[4 * n - 1] * row[3])

    Uhat = np.zeros(n_dof)
    if dirichlet.size:
        Uhat[2 * dirichlet - 1] = U[2 * dirichlet - 1]
        Uhat[2 * dirichlet - 2] = U[2 * dirichlet - 2]
    b = b - A @ Uhat
    if free.size:
        Uhat[free] = lu.solve(b[free])
    return Uhat[:2 * N]


def gauss_point_strain(coordinatesFEM, elementsFEM, U, GaussPtOrder):
    """StrainGaussPt and CoordGaussPt of funGlobalNodalStrainQuadtree.m:
    43-253: the displacement gradient [u_x, u_y, v_x, v_y] of the finite-
    element field U at the GaussPtOrder^2 Gauss points of every element.

    As the reference does (:35), the mid-node columns of elementsFEM are
    cleared first, so every element is a plain bilinear quadrilateral.  The
    element's (ksi, eta) are the bilinear maps that take the corners 1..4 to
    ksi = (-1, 1, 1, -1), eta = (-1, -1, 1, 1) (:106-113); the Gauss points
    lie on the box between corners 1 and 3 (:128-133), row
    o^2 (e - 1) + o (i - 1) + j holding x point i and y point j (:202-216)."""
    coords = np.asarray(coordinatesFEM, dtype=np.float64)
    el = np.asarray(elementsFEM, dtype=np.int64)[:, :4]
    E = el.shape[0]
    U = np.asarray(U, dtype=np.float64).ravel()
    pts, _ = gauss_points(GaussPtOrder)
    o = pts.size
    xy = coords[el - 1]  # (E, 4, 2)
    Lm = np.stack([xy[..., 0] * xy[..., 1], xy[..., 0], xy[..., 1], np.ones((E, 4))], axis=2)
    lk = np.linalg.solve(Lm, np.broadcast_to(np.array([-1.0, 1.0, 1.0, -1.0])[:, None], (E, 4, 1)))[..., 0]
    le = np.linalg.solve(Lm, np.broadcast_to(np.array([-1.0, -1.0, 1.0, 1.0])[:, None], (E, 4, 1)))[..., 0]
    p1, p3 = xy[:, 0], xy[:, 2]
    gx = pts[None, :] * 0.5 * (p3[:, :1] - p1[:, :1]) + 0.5 * (p3[:, :1] + p1[:, :1])  # (E, o)
    gy = pts[None, :] * 0.5 * (p3[:, 1:] - p1[:, 1:]) + 0.5 * (p3[:, 1:] + p1[:, 1:])
    Ue = np.stack([U[2 * el - 2], U[2 * el - 1]], axis=2)  # (E, 4, 2)
    zero = np.zeros((E, 4))
    strain = np.zeros((E, o, o, 4))
    cgp = np.zeros((E, o, o, 2))
    for i in range(o):
        for j in range(o):
            X, Y = gx[:, i], gy[:, j]
            ksi = lk[:, 0] * X * Y + lk[:, 1] * X + lk[:, 2] * Y + lk[:, 3]
            eta = le[:, 0] * X * Y + le[:, 1] * X + le[:, 2] * Y + le[:, 3]
            _, Dk, De = _shape(ksi, eta, zero)
            Dk, De = Dk[:, :4], De[:, :4]
            J11 = np.einsum("ea,ea->e", Dk, xy[..., 0])
            J12 = np.einsum("ea,ea->e", Dk, xy[..., 1])
            J21 = np.einsum("ea,ea->e", De, xy[..., 0])
            J22 = np.einsum("ea,ea->e", De, xy[..., 1])
            det = J11 * J22 - J12 * J21
            Dx = (J22 / det)[:, None] * Dk + (-J12 / det)[:, None] * De
            Dy = (-J21 / det)[:, None] * Dk + (J11 / det)[:, None] * De
            strain[:, i, j, 0] = np.einsum("ea,ea->e", Dx, Ue[..., 0])
            strain[:, i, j, 1] = np.einsum("ea,ea->e", Dy, Ue[..., 0])
            strain[:, i, j, 2] = np.einsum("ea,ea->e", Dx, Ue[..., 1])
            strain[:, i, j, 3] = np.einsum("ea,ea->e", Dy, Ue[..., 1])
            cgp[:, i, j, 0] = X
            cgp[:, i, j, 1] = Y
    return strain.reshape(E * o * o, 4), cgp.reshape(E * o * o, 2)


def funGlobalNodalStrainQuadtree(DICmesh, U, GaussPtOrder, waitBarDisplayOrNot=1):
    """[StrainNodalPt, StrainGaussPt, CoordGaussPt] =
    funGlobalNodalStrainQuadtree(DICmesh, U, GaussPtOrder, waitBarDisplayOrNot)
    (func_quadtree/funGlobalNodalStrainQuadtree.m:1-334).

    The Gauss-point gradients (``gauss_point_strain``) are carried to the
    nodes by natural-neighbour interpolation with nearest-point extrapolation,
    one column at a time (:285-292): nodal column k interpolates Gauss column
    k.  StrainNodalPt is 4N, node-major; its four entries per node are the
    Gauss columns in their order, [u_x, u_y, v_x, v_y], although the reference
    names them F11, F21, F12, F22 (:294) - so entries 2 and 3 are in the
    opposite order to the F of the local steps ([u_x, v_x, u_y, v_y]); kept
    as the reference returns it.  The smoothing branch (:296-330) is not
    reached: the reference fixes smoothness = 0 (:40)."""
    from .natural import natural_plan
    coords = np.asarray(DICmesh["coordinatesFEM"], dtype=np.float64)
    sg, cg = gauss_point_strain(coords, DICmesh["elementsFEM"], U, GaussPtOrder)
    W = _cached(_digest("natural", cg, coords), lambda: natural_plan(cg, coords))
    nodal = W @ sg
    return nodal.reshape(-1), sg, cg
