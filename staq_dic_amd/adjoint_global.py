"""The augmented-Lagrangian global step fed by the adjoint ZNSSD gradient.

The reference's global step (Subpb2Quadtree, func_quadtree/Subpb2Quadtree.m:
70-339; its loop main_ALDIC_Quadtree_inc.m:335-428) sees the images only
through the local steps' results: it minimises
    beta ||grad(Uhat) - F||^2 + mu ||Uhat - U||^2
with (U, F) from the subsets' IC-GN.  This variant keeps the regulariser and
replaces the displacement data term by the image match itself,
    J(Uhat) = sum_i C_i(F_i, Uhat_i) + (beta / 2) ||grad(Uhat) - F||^2,
where C_i is the ZNSSD mismatch of node i's subset warped by [F_i, Uhat_i]
(the cost of funICGNQuadtree.m:362-371 written as a function of P) and F the
local step's deformation gradient.  One Gauss-Newton step on J:
    (D + beta K) dU = -(g + beta (K Uhat - b_F)),
with g_i = dC_i/du_i from the adjoint kernel (staq_znssd_bwd: the
vector-Jacobian product through the bicubic warp), D the per-node 2 x 2
Gauss-Newton curvature of C_i, 2 H_uu / bottomf^2 (H_uu the translation block
of the subset Hessian the local step returns, bottomf^2 = sum (f - fbar)^2
over the window), K the finite-element stiffness and b_F the F term of the
reference's right-hand side (global_step.assemble_subpb2).  FP64 on the host,
like the reference's global step; the cost and its gradient come from the GPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from ._lib import d, ip, lp
from .global_step import _cached, _digest, assemble_subpb2

__all__ = ["znssd_cost_grad", "znssd_curvature", "AdjointGlobalStep"]


def znssd_cost_grad(ctx, coords, winsize, F, U, slot=None):
    """Per-node ZNSSD cost (N,), status (N,) and gradient dC/du (N, 2) of the
    frame pair resident in ``ctx`` for P_i = [F_i, U_i] (staq_znssd_fwd and
    staq_znssd_bwd through the C ABI)."""
    coords = np.asarray(coords, dtype=np.float64)
    N = coords.shape[0]
    P = np.concatenate([np.asarray(F, dtype=np.float64).reshape(N, 4),
                        np.asarray(U, dtype=np.float64).reshape(N, 2)], axis=1)
    P = np.ascontiguousarray(P)
    cf = np.asfortranarray(coords)
    w = np.ascontiguousarray(np.broadcast_to(winsize, (N,)), dtype=np.int64)
    sl = None if slot is None else np.ascontiguousarray(slot, dtype=np.int32)
    cost = np.zeros(N)
    st = np.zeros(N, dtype=np.int32)
    ctx.check(ctx.L.staq_znssd_fwd(ctx.h, N, d(P), d(cf), lp(w), None if sl is None else ip(sl),
                                   d(cost), ip(st)), "staq_znssd_fwd")
    gP = np.zeros(6 * N)
    gU = np.zeros(2 * N)
    ctx.check(ctx.L.staq_znssd_bwd(ctx.h, N, d(P), d(cf), lp(w), None if sl is None else ip(sl),
                                   None, d(gP), d(gU)), "staq_znssd_bwd")
    return cost, st, gU.reshape(N, 2)


def znssd_curvature(Hpar, f, coords, winsize):
    """Per-node 2 x 2 Gauss-Newton curvature of the ZNSSD in u:
    2 H_uu / bottomf^2, with H_uu the (u, v) block of the subset Hessian in
    the packed lower-triangle layout of funICGNQuadtree.m:448 (entries 18, 19,
    20) and bottomf^2 = sum (f - fbar)^2 over the window (funICGNQuadtree.m:
    269-270; f as the frame holds it, windows at the rounded node)."""
    Hpar = np.asarray(Hpar, dtype=np.float64)
    coords = np.asarray(coords, dtype=np.float64)
    N = coords.shape[0]
    wins = np.broadcast_to(np.asarray(winsize, dtype=np.int64), (N,))
    f = np.asarray(f, dtype=np.float64)
    bf2 = np.zeros(N)
    for w in np.unique(wins):
        sel = np.flatnonzero(wins == w)
        h = int(w) // 2
        x = np.round(coords[sel, 0]).astype(np.int64) - 1
        y = np.round(coords[sel, 1]).astype(np.int64) - 1
        o = np.arange(-h, h + 1)
        win = f[x[:, None, None] + o[None, :, None], y[:, None, None] + o[None, None, :]]
        win = win.reshape(len(sel), -1)
        bf2[sel] = np.sum((win - win.mean(axis=1, keepdims=True)) ** 2, axis=1)
    D = np.zeros((N, 2, 2))
    D[:, 0, 0] = Hpar[:, 18]
    D[:, 0, 1] = D[:, 1, 0] = Hpar[:, 19]
    D[:, 1, 1] = Hpar[:, 20]
    return 2.0 * D / bf2[:, None, None]


def AdjointGlobalStep(DICmesh, ctx, winsize, U, F, beta, D, n_iter=1, GaussPtOrder=2):
    """n_iter Gauss-Newton steps of J(Uhat) = sum_i C_i(F_i, Uhat_i) +
    (beta/2) ||grad(Uhat) - F||^2 from Uhat = U over the mesh's nodes (the
    subsets are centred on them).  D: per-node curvature (znssd_curvature).
    Returns Uhat (2N) and the total cost sum_i C_i before each step and after
    the last one.  Nodes whose warped window leaves the image (status 1) keep
    their displacement."""
    coords = np.asarray(DICmesh["coordinatesFEM"], dtype=np.float64)
    el = np.asarray(DICmesh["elementsFEM"], dtype=np.int64)
    if el.shape[1] < 8:
        el = np.concatenate([el, np.zeros((el.shape[0], 8 - el.shape[1]), dtype=np.int64)], axis=1)
    N = coords.shape[0]
    F = np.asarray(F, dtype=np.float64).reshape(N, 4)
    Uh = np.asarray(U, dtype=np.float64).reshape(N, 2).copy()
    # the regulariser's operator and its F term (global_step.assemble_subpb2
    # with mu = 0); the element Jacobians of MeshSetUp's corner order have a
    # negative determinant, which flips A and b together: take K positive.
    # K depends on the mesh alone and is kept per mesh; b_F is assembled here.
    # The F term of a node's DOF takes that node's own F only
    # (Subpb2Quadtree.m:259-262): b_F[u_a] = cx_a F11_a + cy_a F12_a and
    # b_F[v_a] = cx_a F21_a + cy_a F22_a, with cx, cy the node's sums of the
    # weighted shape-function derivatives, kept per mesh with K.
    def build():
        z2 = np.zeros(2 * N)
        K0, _ = assemble_subpb2(coords, el, GaussPtOrder, 1.0, 0.0, z2, np.zeros(4 * N))
        sign = -1.0 if K0.diagonal()[:2 * N].sum() < 0 else 1.0
        e11 = np.zeros((N, 4))
        e11[:, 0] = 1.0
        e12 = np.zeros((N, 4))
        e12[:, 2] = 1.0
        _, bx = assemble_subpb2(coords, el, GaussPtOrder, 1.0, 0.0, z2, e11.ravel(), need_A=False)
        _, by = assemble_subpb2(coords, el, GaussPtOrder, 1.0, 0.0, z2, e12.ravel(), need_A=False)
        return (sign * K0)[:2 * N, :2 * N].tocsr(), sign * bx[0:2 * N:2], sign * by[0:2 * N:2]

    K, cx, cy = _cached(_digest("adjoint K", coords, el, GaussPtOrder), build)
    bF = np.empty(2 * N)
    bF[0::2] = cx * F[:, 0] + cy * F[:, 2]
    bF[1::2] = cx * F[:, 1] + cy * F[:, 3]
    involved = np.unique(el[:, :4]) - 1
    free = np.sort(np.concatenate([2 * involved, 2 * involved + 1]))
    rows = np.repeat(np.arange(2 * N).reshape(N, 2), 2, axis=1).ravel()
    cols = np.tile(np.arange(2 * N).reshape(N, 2), (1, 2)).ravel()
    Dm = sp.coo_matrix((np.asarray(D, dtype=np.float64).reshape(N, 4).ravel(), (rows, cols)),
                       shape=(2 * N, 2 * N)).tocsr()
    costs = []
    for _ in range(n_iter):
        cost, st, g = znssd_cost_grad(ctx, coords, winsize, F, Uh)
        costs.append(float(cost[st == 0].sum()))
        g = np.where((st == 0)[:, None], g, 0.0)
        u = Uh.ravel()
        rhs = -(g.ravel() + beta * (K @ u - bF))
        A = (Dm + beta * K)[free][:, free].tocsc()
        du = np.zeros(2 * N)
        # A is symmetric: a minimum-degree ordering of A + A^T fills in less
        du[free] = spla.spsolve(A, rhs[free], permc_spec="MMD_AT_PLUS_A")
        Uh = (u + du).reshape(N, 2)
    cost, st, _ = znssd_cost_grad(ctx, coords, winsize, F, Uh)
    costs.append(float(cost[st == 0].sum()))
    return Uh.ravel(), costs
