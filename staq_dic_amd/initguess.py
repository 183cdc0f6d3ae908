"""Initial guess on the uniform mesh, on the host (FP64): the drivers' step
between the integer search and the quadtree mesh
(main_ALDIC_Quadtree_v2.m:100-141).

1. the search runs from g to f (the arguments swapped, :104) and its result is
   moved onto f's grid: u_f = -u_g at x_f = x_g + u_g (:105);
2. a thin-plate RBF through the finite points is evaluated on the uniform
   nodes (:113-119; ``refill.thinplate_fit``);
3. ``regularizeNd`` with smoothness 1e-3 smooths it on those same nodes
   (:124-125; func/regularizeNd.m);
4. ``MeshSetUp`` builds the uniform mesh and ``Init`` fills NaNs with the
   springs in-painting and lays the field out as U0 (:128-131; func/Init.m);
5. nodes on a hole of f's mask, or whose displaced integer position lies on a
   hole of g's mask, get NaN (:133-143); a position outside g is skipped, as
   the reference's try/catch skips it.

``regularizeNd`` is restated for every interpolation method of the reference
('linear', 'nearest', 'cubic').  Its direct solvers ('normal', '\\') and its
iterative ones ('lsqr', 'pcg', 'symmlq', stopped at 1e-11 of the data range)
all target the same least-squares minimiser; here every solver is the direct
solve of the normal equations.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from .mesh import MeshSetUp, uniform_nodes
from .outliers import inpaint_nans
from .refill import thinplate_eval, thinplate_fit

__all__ = ["regularizeNd", "Init", "mask_nan_nodes", "uniform_initial_guess"]


def _histc(x, grid):
    """histc's bin (1-based): grid[k-1] <= x < grid[k], the last node in bin
    n; regularizeNd then moves bin n to n-1 (regularizeNd.m, each method)."""
    k = np.searchsorted(grid, x, side="right")
    return k


def _second_diff_weights(g):
    """secondDerivativeWeights of regularizeNd.m: the three-point second
    derivative on a non-uniform grid, one row per interior node."""
    x1, x2, x3 = g[:-2], g[1:-1], g[2:]
    return 2.0 / np.stack([(x1 - x3) * (x1 - x2), (x2 - x1) * (x2 - x3), (x3 - x1) * (x3 - x2)], axis=1)


def regularizeNd(x, y, xGrid, smoothness=0.01, interpMethod="linear", solver="normal"):
    """yGrid = regularizeNd(x, y, xGrid, smoothness, interpMethod, solver).

    x (n, d) scattered points, y (n,) values, xGrid a list of d increasing
    grid vectors; returns the grid values (n_1 x ... x n_d, MATLAB's order).
    The fit minimises ||A_fid yGrid - y||^2 + sum_i ||L_i yGrid||^2 with A_fid
    the interpolation of the grid at the points and L_i the second derivative
    along dimension i scaled by smoothness_i * sqrt(n / rows(L_i)) *
    (range of grid i)^2 (regularizeNd.m)."""
    x = np.atleast_2d(np.asarray(x, dtype=np.float64))
    y = np.asarray(y, dtype=np.float64).ravel()
    grids = [np.asarray(g, dtype=np.float64).ravel() for g in xGrid]
    d = len(grids)
    if x.shape[1] != d:
        raise ValueError("Dimensionality mismatch: columns of x and cells of xGrid differ")
    if x.shape[0] != y.size:
        raise ValueError("x and y must have the same number of rows")
    sm = np.broadcast_to(np.asarray(smoothness, dtype=np.float64), (d,)).copy()
    if (sm < 0).any():
        raise ValueError("smoothness must be positive in all components")
    method = interpMethod.lower()
    if method not in ("linear", "nearest", "cubic"):
        raise ValueError(f"{interpMethod} is not a possible interpolation method")
    if solver not in ("\\", "normal", "lsqr", "pcg", "symmlq"):
        raise ValueError(f"{solver} is not an acceptable solver")
    nG = np.array([g.size for g in grids])
    for g in grids:
        if (np.diff(g) <= 0).any():
            raise ValueError("All grid points must be monotonically increasing")
    if (nG < (4 if method == "cubic" else 3)).any():
        raise ValueError("Not enough grid points in each dimension")
    gmin = np.array([g[0] for g in grids])
    gmax = np.array([g[-1] for g in grids])
    if (x < gmin).any() or (x > gmax).any():
        raise ValueError("All x points must be within the range of the grid vectors")
    n = x.shape[0]
    total = int(np.prod(nG))
    stride = np.concatenate([[1], np.cumprod(nG)[:-1]])

    # fidelity rows: each point's interpolation weights on its cell's nodes
    per_dim = []  # (indices (n, k), weights (n, k)) per dimension, 0-based
    for i, g in enumerate(grids):
        xi = x[:, i]
        k = _histc(xi, g)
        if method == "cubic":
            k0 = np.minimum(np.maximum(k - 1, 1), nG[i] - 3) - 1
            t = [g[k0 + q] for q in range(4)]
            a = [xi - tq for tq in t]
            w = np.stack([
                a[1] / (t[0] - t[1]) * a[2] / (t[0] - t[2]) * a[3] / (t[0] - t[3]),
                -a[0] / (t[0] - t[1]) * a[2] / (t[1] - t[2]) * a[3] / (t[1] - t[3]),
                a[0] / (t[0] - t[2]) * a[1] / (t[1] - t[2]) * a[3] / (t[2] - t[3]),
                -a[0] / (t[0] - t[3]) * a[1] / (t[1] - t[3]) * a[2] / (t[2] - t[3])], axis=1)
            idx = k0[:, None] + np.arange(4)[None, :]
        else:
            k = np.where(k == nG[i], nG[i] - 1, k) - 1
            frac = np.clip((xi - g[k]) / (g[k + 1] - g[k]), 0.0, 1.0)
            if method == "nearest":
                idx = (k + np.floor(frac + 0.5).astype(np.int64))[:, None]
                w = np.ones((n, 1))
            else:
                idx = np.stack([k, k + 1], axis=1)
                w = np.stack([1.0 - frac, frac], axis=1)
        per_dim.append((idx, w))
    col = np.zeros((n, 1), dtype=np.int64)
    wt = np.ones((n, 1))
    for i, (idx, w) in enumerate(per_dim):
        col = (col[:, :, None] + stride[i] * idx[:, None, :]).reshape(n, -1)
        wt = (wt[:, :, None] * w[:, None, :]).reshape(n, -1)
    rows = np.repeat(np.arange(n), col.shape[1])
    A_fid = sp.csr_matrix((wt.ravel(), (rows, col.ravel())), shape=(n, total))

    # smoothness rows: the second difference along each dimension
    blocks = [A_fid]
    allnodes = np.arange(total).reshape(tuple(nG[::-1])).transpose()  # MATLAB order
    for i, g in enumerate(grids):
        if sm[i] == 0:
            continue
        sl = [slice(None)] * d
        trip = []
        for q in range(3):
            sl[i] = slice(q, nG[i] - 2 + q)
            trip.append(allnodes[tuple(sl)])
        m_i = trip[0].size
        wts = _second_diff_weights(g)  # (nG[i] - 2, 3)
        shape = [1] * d
        shape[i] = nG[i] - 2
        wfull = [np.broadcast_to(wts[:, q].reshape(shape), trip[0].shape) for q in range(3)]
        scale = sm[i] * np.sqrt(n / m_i) * (gmax[i] - gmin[i]) ** 2
        r = np.arange(m_i)
        L = sp.csr_matrix((scale * np.concatenate([wq.ravel(order="F") for wq in wfull]),
                           (np.tile(r, 3), np.concatenate([t.ravel(order="F") for t in trip]))),
                          shape=(m_i, total))
        blocks.append(L)
    A = sp.vstack(blocks).tocsr()
    rhs = np.concatenate([y, np.zeros(A.shape[0] - n)])
    yg = spla.spsolve((A.T @ A).tocsc(), A.T @ rhs)
    return yg.reshape(tuple(nG), order="F") if d > 1 else yg


def _std_omitnan(a, axis):
    k = np.isfinite(a).sum(axis=axis)
    mu = np.nanmean(a, axis=axis, keepdims=True)
    s = np.nansum((a - mu) ** 2, axis=axis) / np.maximum(k - 1, 1)
    return np.where(k > 1, np.sqrt(s), np.where(k == 1, 0.0, np.nan))


def _spring_pass(u, v, weighted8, nsig, clear8):
    """One neighbour-consistency pass of Init.m index 1: a node whose value
    differs from its neighbours' (weighted) average by more than nsig
    standard deviations of those neighbours gets NaN, and so do its 4 (or 8)
    neighbours; the border becomes NaN and everything is in-painted again."""
    uo, vo = u.copy(), v.copy()
    n, m = u.shape
    if n >= 3 and m >= 3:
        def stack(a):
            c = a[1:-1, 1:-1]
            e4 = [a[:-2, 1:-1], a[2:, 1:-1], a[1:-1, :-2], a[1:-1, 2:]]
            e8 = [a[:-2, :-2], a[2:, :-2], a[2:, 2:], a[:-2, 2:]]
            return c, e4, e8

        bad = np.zeros((n - 2, m - 2), dtype=bool)
        for a in (u, v):
            c, e4, e8 = stack(a)
            if weighted8:
                avg = (sum(e4) / 8 + sum(e8) / 16) / 0.75
                sd = _std_omitnan(np.stack(e4 + e8), 0)
            else:
                avg = np.nanmean(np.stack(e4), 0)
                sd = _std_omitnan(np.stack(e4), 0)
            err = np.abs(c - avg) - nsig * sd
            bad |= err > 0
        bi, bj = np.nonzero(bad)
        bi, bj = bi + 1, bj + 1
        offs = [(0, 0), (-1, 0), (1, 0), (0, -1), (0, 1)]
        if clear8:
            offs += [(-1, -1), (1, -1), (1, 1), (-1, 1)]
        for di, dj in offs:
            uo[bi + di, bj + dj] = np.nan
            vo[bi + di, bj + dj] = np.nan
    for a in (uo, vo):
        a[0, :] = np.nan
        a[-1, :] = np.nan
        a[:, 0] = np.nan
        a[:, -1] = np.nan
    return inpaint_nans(uo, 4), inpaint_nans(vo, 4)


def Init(u, v, Phi, x, y, index=0):
    """U0 = Init(u, v, Phi, x, y, index) (func/Init.m): NaNs of the gridded
    u, v (x index first, as ndgrid lays them out) in-painted with springs;
    index 1 adds three neighbour-consistency passes (3, 2 and 2 standard
    deviations).  U0 is 2N interleaved [u1 v1 u2 v2 ...] over the nodes in
    MeshSetUp's order (y index fastest).  Phi, x and y are accepted for the
    reference's signature; only the node count is taken from u."""
    uI = inpaint_nans(np.asarray(u, dtype=np.float64), 4)
    vI = inpaint_nans(np.asarray(v, dtype=np.float64), 4)
    if index == 1:
        uI, vI = _spring_pass(uI, vI, True, 3.0, True)
        uI, vI = _spring_pass(uI, vI, False, 2.0, False)
        uI, vI = _spring_pass(uI, vI, True, 2.0, True)
    U0 = np.empty(2 * uI.size)
    U0[0::2] = uI.ravel(order="C")  # transpose, then column-major: y index fastest
    U0[1::2] = vI.ravel(order="C")
    return U0


def mask_nan_nodes(X, Y, u, v, fMask=None, gMask=None):
    """Nodes of main_ALDIC_Quadtree_v2.m:133-143 that get NaN:
    ~fMask(x, y) || ~gMask(floor(x + u), floor(y + v)), evaluated inside a
    try/catch, so a node whose index would raise (not an integer, outside the
    image, NaN) is skipped; the g test runs only where f's passes.  All
    arrays nx x ny (ndgrid layout); a mask of None is all ones."""
    X = np.asarray(X, dtype=np.float64)
    Y = np.asarray(Y, dtype=np.float64)

    def lookup(mask, a, b):
        ok = np.isfinite(a) & np.isfinite(b)
        ok &= (a == np.floor(np.where(ok, a, 0))) & (b == np.floor(np.where(ok, b, 0)))
        if mask is None:
            return ok, np.ones(a.shape, dtype=bool)
        m = np.asarray(mask)
        ok &= (a >= 1) & (a <= m.shape[0]) & (b >= 1) & (b <= m.shape[1])
        val = np.zeros(a.shape, dtype=bool)
        ai = np.where(ok, a, 1).astype(np.int64) - 1
        bi = np.where(ok, b, 1).astype(np.int64) - 1
        val[ok] = m[ai[ok], bi[ok]] != 0
        return ok, val

    f_ok, f_val = lookup(fMask, X, Y)
    if gMask is None:
        return f_ok & ~f_val
    with np.errstate(invalid="ignore"):
        g_ok, g_val = lookup(gMask, np.floor(X + u), np.floor(Y + v))
    return f_ok & (~f_val | (g_ok & ~g_val))


def uniform_initial_guess(fNormalized, gNormalized, DICpara, fMask=None, gMask=None, ctx=None):
    """The initial guess of main_ALDIC_Quadtree_v2.m:100-141 on the uniform
    mesh.  Returns (DICpara, DICmesh, U0, (x0temp, y0temp, u, v), cc):
    DICmesh holds coordinatesFEM, elementsFEM (1-based), x0, y0 (ny x nx, as
    MeshSetUp transposes them), M, N; u, v are the regularised fields on the
    ndgrid nodes (nx x ny)."""
    from .api import IntegerSearch
    f = np.asarray(fNormalized)
    H, W = f.shape
    P = dict(DICpara)
    # the search from g to f, moved onto f (:104-105)
    P, xg, yg, ug, vg, cc = IntegerSearch(gNormalized, fNormalized, None, P, ctx=ctx)
    u_f, v_f = -ug, -vg
    xf, yf = xg + ug, yg + vg
    roi = P.get("gridxyROIRange", {"gridx": [1, H], "gridy": [1, W]})
    xnodes, ynodes = uniform_nodes(H, W, P["winsize"], P["winstepsize"],
                                   tuple(roi["gridx"]), tuple(roi["gridy"]))
    x0temp, y0temp = np.meshgrid(xnodes, ynodes, indexing="ij")
    ok = np.isfinite(u_f.ravel(order="F"))
    centers = np.stack([xf.ravel(order="F")[ok], yf.ravel(order="F")[ok]], axis=1)
    vals = np.stack([u_f.ravel(order="F")[ok], v_f.ravel(order="F")[ok]], axis=1)
    coeff = thinplate_fit(centers, vals)
    nodes = np.stack([x0temp.ravel(order="F"), y0temp.ravel(order="F")], axis=1)
    uv = thinplate_eval(centers, coeff, nodes)
    u = regularizeNd(nodes, uv[:, 0], [xnodes, ynodes], 1e-3)
    v = regularizeNd(nodes, uv[:, 1], [xnodes, ynodes], 1e-3)
    coords, elems = MeshSetUp(xnodes, ynodes)
    DICmesh = {"coordinatesFEM": coords, "elementsFEM": elems,
               "x0": x0temp.T, "y0": y0temp.T, "M": ynodes.size, "N": xnodes.size,
               "xnodes": xnodes, "ynodes": ynodes}
    U0 = Init(u, v, cc.get("max") if isinstance(cc, dict) else None, x0temp.T, y0temp.T, 0)
    fm = None if fMask is None else np.asarray(fMask)
    gm = None if gMask is None else np.asarray(gMask)
    hole = mask_nan_nodes(x0temp, y0temp, u, v, fm, gm)
    k = np.flatnonzero(hole.ravel(order="C"))  # node j + i ny of MeshSetUp's order
    U0[2 * k] = np.nan
    U0[2 * k + 1] = np.nan
    return P, DICmesh, U0, (x0temp, y0temp, u, v), cc
