"""Quadtree mesh and subset-list producer (host, FP64/integer): the uniform
mesh of MeshSetUp (func/MeshSetUp.m) and its refinement at the mask edge by
GenerateQuadtreeMesh (func_quadtree/GenerateQuadtreeMesh.m:18-203), with
funMarkEdge / funMarkInside and the red refinement with hanging nodes of
QrefineR (func_quadtree/refinement/QrefineR.m, provideGeometricData.m).

Node and element numbers follow the reference's (1-based in the returned
arrays, as DICmesh stores them): new edge nodes are numbered in the order of
the sorted unique edge list, element mid-nodes in element order, children in
the reference's corner order, so a mesh produced here equals the reference's
entry for entry on the same mask.  The initial-guess transfer onto the new
nodes (:152-185, thin-plate RBF per mask component) is
``transfer_initial_guess``.
"""
from __future__ import annotations

import numpy as np

__all__ = ["MeshSetUp", "funMarkEdge", "funMarkInside", "QrefineR", "GenerateQuadtreeMesh",
           "uniform_nodes", "transfer_initial_guess"]


def uniform_nodes(H, W, winsize, winstepsize, gridx=(1, None), gridy=(1, None)):
    """xnodes, ynodes of the uniform mesh (main_ALDIC_Quadtree_v2.m:108-111)."""
    gx1 = H if gridx[1] is None else gridx[1]
    gy1 = W if gridy[1] is None else gridy[1]
    x0 = max(4 + 0.5 * winsize, gridx[0])
    x1 = min(H - 0.5 * winsize - 3, gx1)
    y0 = max(4 + 0.5 * winsize, gridy[0])
    y1 = min(W - 0.5 * winsize - 3, gy1)
    xn = np.arange(x0, x1 + 1e-9, winstepsize)
    yn = np.arange(y0, y1 + 1e-9, winstepsize)
    return xn, yn


def MeshSetUp(xnodes, ynodes):
    """coordinatesFEM (N x 2) and elementsFEM (E x 4, 1-based) of the uniform
    mesh on ndgrid(xnodes, ynodes) (func/MeshSetUp.m:20-43): node k runs over
    y fastest; element corners (x_j, y_i), (x_j, y_i+1), (x_j+1, y_i+1),
    (x_j+1, y_i)."""
    xn = np.asarray(xnodes, dtype=float)
    yn = np.asarray(ynodes, dtype=float)
    M, N = yn.size, xn.size
    X, Y = np.meshgrid(xn, yn)  # M x N: X[i, j] = xn[j], Y[i, j] = yn[i]
    coords = np.stack([X.ravel(order="F"), Y.ravel(order="F")], axis=1)
    j, i = np.meshgrid(np.arange(1, N), np.arange(1, M))  # (M-1) x (N-1), i fastest
    i, j = i.ravel(order="F"), j.ravel(order="F")
    base = (j - 1) * M + i
    elems = np.stack([base, base + 1, base + M + 1, base + M], axis=1).astype(np.int64)
    return coords, elems


def _boxes(coords, elems4):
    c = coords[elems4[:, :4] - 1]
    xmin = c[:, :, 0].min(1).astype(np.int64)
    xmax = c[:, :, 0].max(1).astype(np.int64)
    ymin = c[:, :, 1].min(1).astype(np.int64)
    ymax = c[:, :, 1].max(1).astype(np.int64)
    return xmin, xmax, ymin, ymax


def funMarkEdge(coords, elems4, mask, elementMinSize):
    """Elements whose bounding box sees both values of the mask and whose size
    exceeds elementMinSize (func_quadtree/funMarkEdge.m:9-32); boolean."""
    xmin, xmax, ymin, ymax = _boxes(coords, elems4)
    m = np.asarray(mask)
    rng = np.zeros(len(xmin))
    for k in range(len(xmin)):
        blk = m[xmin[k] - 1:xmax[k], ymin[k] - 1:ymax[k]]
        rng[k] = float(blk.max()) - float(blk.min())
    size = np.minimum(np.abs(xmax - xmin), np.abs(ymax - ymin))
    return (rng > 0) & (size > elementMinSize)


def funMarkInside(coords, elems4, mask):
    """1-based indices of elements inside the hole (mask sum below half the
    box area) and outside it (func_quadtree/funMarkInside.m:8-31)."""
    xmin, xmax, ymin, ymax = _boxes(coords, elems4)
    m = np.asarray(mask, dtype=np.float64)
    inside = np.zeros(len(xmin), dtype=bool)
    for k in range(len(xmin)):
        s = m[xmin[k] - 1:xmax[k], ymin[k] - 1:ymax[k]].sum()
        inside[k] = s < 0.5 * (xmax[k] - xmin[k]) * (ymax[k] - ymin[k])
    idx = np.arange(1, len(xmin) + 1)
    return idx[inside], idx[~inside]


def _edges(irregular, elems):
    """provideGeometricData(irregular, elements): the sorted unique edge list
    and, per irregular triple and per element, its edges (1-based)."""
    K, E = irregular.shape[0], elems.shape[0]
    e3 = irregular[:, [0, 1, 2, 1, 2, 0]]
    e4 = elems[:, [0, 1, 2, 3, 1, 2, 3, 0]]
    ed = np.concatenate([np.concatenate([e3[:, :3].ravel(order="F")[:, None],
                                         e3[:, 3:].ravel(order="F")[:, None]], 1),
                         np.concatenate([e4[:, :4].ravel(order="F")[:, None],
                                         e4[:, 4:].ravel(order="F")[:, None]], 1)])
    ed = np.sort(ed, axis=1)
    edge2nodes, ie = np.unique(ed, axis=0, return_inverse=True)
    ie = ie.ravel() + 1
    irr2e = ie[:3 * K].reshape((K, 3), order="F")
    el2e = ie[3 * K:3 * K + 4 * E].reshape((E, 4), order="F")
    return edge2nodes, irr2e, el2e


def _find_cm(A):
    """MATLAB find on a matrix: (row, col, value) of nonzeros in column-major
    order, 0-based indices."""
    c, r = np.nonzero(A.T)
    return r, c, A[r, c]


def QrefineR(coords, elems, irregular, marked):
    """Red refinement of the marked quadrilaterals with hanging nodes
    (QrefineR.m:43-124): returns (coordinates, elements, irregular); elements
    and irregular are 1-based, irregular rows are [node, node, hanging node]."""
    coords = np.asarray(coords, dtype=float)
    elems = np.asarray(elems, dtype=np.int64)
    irregular = np.asarray(irregular, dtype=np.int64).reshape(-1, 3)
    marked = np.asarray(marked, dtype=np.int64)
    nE = elems.shape[0]
    edge2nodes, irr2e, el2e = _edges(irregular, elems)
    e2n = np.zeros(edge2nodes.shape[0] + 1)  # 1-based, entry 0 unused
    if marked.size:
        e2n[el2e[marked - 1].ravel()] = 1
    e2n[irr2e[:, 0]] = 1
    # closure (:57-70): an element with three marked edges (or one edge of an
    # irregular pair refined further) gets all four; until nothing changes
    more = True
    while more:
        me = e2n[el2e]
        s = np.abs(me).sum(1)
        kdx = np.nonzero((s < 4) & ((s > 2) | (me.min(1) < 0)))[0]
        r, c, _ = _find_cm((me[kdx] == 0).astype(np.int8))
        e2n[el2e[kdx[r], c]] = 1
        me3 = e2n[irr2e]
        flag = irr2e[np.any(me3[:, 1:] != 0, axis=1), 0]
        swap = np.nonzero(e2n[flag] != -1)[0]
        e2n[flag[swap]] = -1
        more = kdx.size > 0 or swap.size > 0
    e2n[irr2e[:, 0]] = -1
    pos = np.nonzero(e2n[1:] > 0)[0]  # 0-based edge indices
    n0 = coords.shape[0]
    e2n[pos + 1] = n0 + np.arange(1, pos.size + 1)
    coords = np.concatenate([coords, 0.5 * (coords[edge2nodes[pos, 0] - 1] +
                                            coords[edge2nodes[pos, 1] - 1])])
    e2n[irr2e[:, 0]] = irregular[:, 2]
    nn = e2n[el2e].astype(np.int64)  # nE x 4
    reftyp = (nn != 0) @ (2 ** np.arange(4))
    none = reftyp < 15
    red = reftyp == 15
    ridx = np.nonzero(red)[0]
    mid = np.zeros(nE, dtype=np.int64)
    mid[ridx] = coords.shape[0] + np.arange(1, ridx.size + 1)
    coords = np.concatenate([coords, 0.25 * (coords[elems[ridx, 0] - 1] + coords[elems[ridx, 1] - 1] +
                                              coords[elems[ridx, 2] - 1] + coords[elems[ridx, 3] - 1])])
    cnt = np.where(none, 1, 0) + np.where(red, 4, 0)
    start = np.concatenate([[0], np.cumsum(cnt)])  # 0-based row of each element's first child
    out = np.zeros((start[-1], 4), dtype=np.int64)
    out[start[:-1][none]] = elems[none]
    e, n, m = elems[red], nn[red], mid[red]
    s0 = start[:-1][red]
    out[s0] = np.stack([e[:, 0], n[:, 0], m, n[:, 3]], 1)
    out[s0 + 1] = np.stack([e[:, 1], n[:, 1], m, n[:, 0]], 1)
    out[s0 + 2] = np.stack([e[:, 2], n[:, 2], m, n[:, 1]], 1)
    out[s0 + 3] = np.stack([e[:, 3], n[:, 3], m, n[:, 2]], 1)
    # hanging nodes: on edges of partly refined elements, then on irregular half edges
    kdx = np.nonzero((reftyp > 0) & (reftyp < 15))[0]
    r, c, v = _find_cm(nn[kdx])
    edx = el2e[kdx[r], c]
    irr_new = np.concatenate([edge2nodes[edx - 1], v[:, None]], 1)
    nn3 = e2n[irr2e[:, 1:3]].astype(np.int64)
    kdx = np.nonzero(nn3.sum(1) != 0)[0]
    r, c, v = _find_cm(nn3[kdx])
    edx = irr2e[kdx[r], c + 1]
    irr_new = np.concatenate([irr_new, np.concatenate([edge2nodes[edx - 1], v[:, None]], 1)])
    return coords, out, irr_new.astype(np.int64)


def _ismember_row(row, A):
    """Lowest 0-based index of ``row`` among the rows of A, or -1."""
    hit = np.nonzero((A[:, 0] == row[0]) & (A[:, 1] == row[1]))[0]
    return int(hit[0]) if hit.size else -1


def GenerateQuadtreeMesh(coords0, elems0, mask, winstepsize, winsize, elementMinSize=2):
    """The mesh part of GenerateQuadtreeMesh.m:18-141: refine the uniform mesh
    at the mask edge down to elementMinSize, order each element's corners,
    attach hanging nodes as columns 5-8, drop elements inside the hole, and
    mark the nodes near the hole and the ROI border (markCoordHoleEdge, which
    the driver also uses as dirichlet).  Returns (coordinatesFEM,
    elementsFEM (E x 8), irregular, markCoordHoleEdge), 1-based."""
    coords = np.asarray(coords0, dtype=float)
    elems = np.asarray(elems0, dtype=np.int64)[:, :4]
    irregular = np.zeros((0, 3), dtype=np.int64)
    while True:
        mark = np.nonzero(funMarkEdge(coords, elems, mask, elementMinSize))[0] + 1
        coords, elems, irregular = QrefineR(coords, elems, irregular, mark)
        if mark.size == 0:
            break
    # corner order: the corner nearest the box's low corner first (:33-41)
    c = coords[elems - 1]
    for t in range(elems.shape[0]):
        cx, cy = c[t, :, 0], c[t, :, 1]
        key = (cx - np.sort(cx)[0]) ** 2 + 2 * (cy - np.sort(cy)[0]) ** 2
        o = np.argsort(key, kind="stable")
        elems[t] = elems[t, o[[0, 1, 3, 2]]]
    E = np.zeros((elems.shape[0], 8), dtype=np.int64)
    E[:, :4] = elems
    # hanging nodes into columns 5-8 (:43-61), in the reference's order of tests
    tests = [((0, 1), 7), ((1, 2), 4), ((2, 3), 5), ((3, 0), 6),
             ((1, 0), 7), ((2, 1), 4), ((3, 2), 5), ((0, 3), 6)]
    for t in range(irregular.shape[0]):
        row = irregular[t, :2]
        for (a, b), col in tests:
            k = _ismember_row(row, E[:, [a, b]])
            if k >= 0:
                E[k, col] = irregular[t, 2]
    _, outside = funMarkInside(coords, E, mask)
    E = E[outside - 1]
    # nodes near the hole edge and near the ROI border (:66-101)
    d = np.hypot(coords[E[:, 0] - 1, 0] - coords[E[:, 2] - 1, 0],
                 coords[E[:, 0] - 1, 1] - coords[E[:, 2] - 1, 1])
    step = float(np.max([winstepsize, 0 * winsize]))
    c0 = np.asarray(coords0, dtype=float)
    xMin, xMax, yMin, yMax = c0[:, 0].min(), c0[:, 0].max(), c0[:, 1].min(), c0[:, 1].max()
    sel = ((d < 0.99 * np.sqrt(2) * step) |
           (coords[E[:, 0] - 1, 0] < xMin + 1.01 * step) |
           (coords[E[:, 2] - 1, 0] > xMax - 1.01 * step) |
           (coords[E[:, 0] - 1, 1] < yMin + 1.01 * step) |
           (coords[E[:, 2] - 1, 1] > yMax - 1.01 * step))
    mark = np.unique(E[sel])
    mark = mark[mark != 0]
    for _ in range(2):
        near = np.isin(E, mark).any(axis=1)
        mark = np.unique(E[near])
        mark = mark[mark != 0]
    return coords, E, irregular, mark


def transfer_initial_guess(U0, coords0, coords, mask):
    """U0 of the uniform mesh onto the quadtree nodes (GenerateQuadtreeMesh.m:
    152-185): per 4-connected component of the slightly dilated mask, a
    thin-plate RBF through the uniform nodes with a finite U0 inside the
    component, evaluated at the quadtree nodes inside it; other nodes get 0."""
    from scipy import ndimage

    from .refill import thinplate_eval, thinplate_fit
    U0 = np.asarray(U0, dtype=float)
    coords0 = np.asarray(coords0, dtype=float)
    coords = np.asarray(coords, dtype=float)
    m = np.asarray(mask, dtype=float)
    dil = ndimage.gaussian_filter(m, 1.0, mode="nearest", truncate=2.0) > 0.01
    lab, n = ndimage.label(dil, structure=[[0, 1, 0], [1, 1, 1], [0, 1, 0]])
    out = np.zeros(2 * coords.shape[0])
    ok0 = np.isfinite(U0[0::2])
    l0 = lab[coords0[:, 0].astype(int) - 1, coords0[:, 1].astype(int) - 1]
    l1 = lab[np.round(coords[:, 0]).astype(int) - 1, np.round(coords[:, 1]).astype(int) - 1]
    for k in range(1, n + 1):
        src = np.nonzero(ok0 & (l0 == k))[0]
        dst = np.nonzero(l1 == k)[0]
        if src.size < 3 or dst.size == 0:
            continue
        vals = np.stack([U0[2 * src], U0[2 * src + 1]], axis=1)
        coeff = thinplate_fit(coords0[src], vals)
        val = thinplate_eval(coords0[src], coeff, coords[dst])
        out[2 * dst] = val[:, 0]
        out[2 * dst + 1] = val[:, 1]
    return out
