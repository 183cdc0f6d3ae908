"""Post-processing of the uniform-grid local steps on the host (FP64): the
plain twins' counterparts of refill.refill_bad.

LocalICGN (func/LocalICGN.m:117-292), after the bad subsets are set to NaN:
  1. for every 8-connected component of the reference mask, the nodes with a
     finite u whose rounded position lies in it are replaced by a local plane
     fit (funCompDefGrad2, func_quadtree/funCompDefGrad2.m:38-150): for each
     such node, the least-squares plane u = a + b dx + c dy (and the same for
     v) through the component's nodes closer than Rad = 1 + 2 winstepsize
     whose rounded positions lie in the 4-connected part of the mask window
     around the node that holds its centre; U takes a, F takes the slopes
     [u_x, v_x, u_y, v_y].  With 3 or fewer such neighbours, or a mask window
     that leaves the image, the node keeps its u and its F becomes NaN;
  2. rmoutliers(F_k, 'movmedian', 1 + winstepsize) on each F component in node
     order; a node flagged in any component loses all four (:218-233);
  3. the NaN nodes take the value of their nearest node with a finite value,
     U from the finite u, F from the finite F (scatteredInterpolant 'nearest',
     'nearest', :267-292).
funCompDefGrad2 also re-fills F by linear interpolation inside its node loop
(:160-176), from the F that is finite at that moment: one value after the
first node, more as the loop goes on.  What MATLAB's scatteredInterpolant
returns for such a fill from fewer than three points is not fixed by the
reference's files, so that in-loop fill is not restated: F stays NaN where the
fit had too few neighbours, and stage 3 fills it.  This piece is parity
unpinned.

Subpb1 (func/Subpb1.m:171-199), after its bad subsets are set to NaN: gridfit
(regularizer 'springs', smoothness 1, triangle interpolation) of u and of v
through the nodes with a finite u, on the grid unique(x) x unique(y), read
back at every node: a smoothing fit, so every node takes the fitted value,
not only the NaN ones.
"""
from __future__ import annotations

import numpy as np
from scipy import ndimage
from scipy.spatial import cKDTree

from .mgsearch import gridfit

__all__ = ["isoutlier_movmedian", "comp_def_grad2", "local_icgn_refill", "subpb1_gridfit"]

# scale of the MAD that makes it a consistent estimator of a normal
# distribution's standard deviation: -1 / (sqrt(2) erfcinv(3/2))
MAD_SCALE = 1.482602218505602


def isoutlier_movmedian(a, window):
    """isoutlier(A, 'movmedian', window) of a vector: true where an element is
    more than three local scaled MADs (MAD_SCALE x the median absolute
    deviation) from the local median.  The window holds ``window`` elements:
    centred for an odd length, one more before than after for an even one,
    shrunk at the ends.  NaN elements are never outliers and are left out of
    the local statistics."""
    a = np.asarray(a, dtype=np.float64).ravel()
    n = a.size
    k = int(window)
    before, after = (k - 1) // 2, k // 2
    if k % 2 == 0:
        before, after = k // 2, k // 2 - 1
    out = np.zeros(n, dtype=bool)
    for i in range(n):
        if not np.isfinite(a[i]):
            continue
        w = a[max(0, i - before):min(n, i + after + 1)]
        w = w[np.isfinite(w)]
        med = np.median(w)
        mad = MAD_SCALE * np.median(np.abs(w - med))
        out[i] = abs(a[i] - med) > 3.0 * mad
    return out


def comp_def_grad2(uv, coords, f_o_s, mask):
    """[~, U, F] = funCompDefGrad2(uv, coords, f_o_s, n_neighbors, mask) with
    n_neighbors above the node count (every node is a candidate neighbour):
    uv (n, 2), coords (n, 2) 1-based pixel positions; returns U (2n), F (4n)."""
    uv = np.asarray(uv, dtype=np.float64).reshape(-1, 2)
    c = np.asarray(coords, dtype=np.float64).reshape(-1, 2)
    n = c.shape[0]
    U = np.full(2 * n, np.nan)
    F = np.full(4 * n, np.nan)
    if n == 0:
        return U, F
    m = None if mask is None else np.asarray(mask) != 0
    tree = cKDTree(c)
    fos = float(f_o_s)
    h = int(round(fos))
    four = ndimage.generate_binary_structure(2, 1)
    for i in range(n):
        nb = np.asarray(tree.query_ball_point(c[i], fos), dtype=np.int64)
        nb = nb[np.hypot(*(c[nb] - c[i]).T) < fos]
        ok = True
        if m is not None:
            rows = np.round(np.arange(c[i, 0] - fos, c[i, 0] + fos + 0.5)).astype(np.int64) - 1
            cols = np.round(np.arange(c[i, 1] - fos, c[i, 1] + fos + 0.5)).astype(np.int64) - 1
            if rows[0] < 0 or cols[0] < 0 or rows[-1] >= m.shape[0] or cols[-1] >= m.shape[1]:
                ok = False  # the window leaves the image: the reference's catch
            else:
                win = m[np.ix_(rows, cols)]
                lab, _ = ndimage.label(win, structure=four)
                centre = lab[h, h]
                if centre == 0:
                    nb = nb[:0]
                else:
                    r = np.round(c[nb, 0] - (c[i, 0] - fos)).astype(np.int64)
                    s = np.round(c[nb, 1] - (c[i, 1] - fos)).astype(np.int64)
                    nb = nb[lab[r, s] == centre]
        if ok and nb.size > 3:
            A = np.stack([np.ones(nb.size), c[nb, 0] - c[i, 0], c[nb, 1] - c[i, 1]], axis=1)
            P, *_ = np.linalg.lstsq(A, uv[nb], rcond=None)  # columns: u, v
            U[2 * i:2 * i + 2] = P[0]
            F[4 * i:4 * i + 4] = [P[1, 0], P[1, 1], P[2, 0], P[2, 1]]
        else:
            U[2 * i:2 * i + 2] = uv[i]
    return U, F


def _nearest_fill(coords, values, finite):
    """scatteredInterpolant(coords(finite), values(finite), 'nearest',
    'nearest') evaluated at every node: the finite nodes keep their values,
    the others take their nearest finite node's (the lowest index on a
    tie)."""
    out = values.copy()
    src = np.flatnonzero(finite)
    dst = np.flatnonzero(~finite)
    if dst.size == 0 or src.size == 0:
        return out
    tree = cKDTree(coords[src])
    d, j = tree.query(coords[dst], k=1)
    for t, q in enumerate(dst):
        cand = tree.query_ball_point(coords[q], d[t] * (1 + 1e-12) + 1e-300)
        out[q] = values[src[min(cand)] if cand else src[j[t]]]
    return out


def local_icgn_refill(U, F, coordinatesFEM, mask, winstepsize):
    """Stages 1-3 of LocalICGN.m:144-292 on U (2N) and F (4N) whose bad nodes
    are NaN already; returns new U, F."""
    c = np.asarray(coordinatesFEM, dtype=np.float64)
    N = c.shape[0]
    U = np.asarray(U, dtype=np.float64).reshape(N, 2).copy()
    F = np.asarray(F, dtype=np.float64).reshape(N, 4).copy()
    m = np.asarray(mask) != 0
    step = float(np.mean(winstepsize))
    rad = 1.0 + 2.0 * step
    lab, ncomp = ndimage.label(m, structure=np.ones((3, 3), dtype=bool))
    px = np.round(c[:, 0]).astype(np.int64) - 1
    py = np.round(c[:, 1]).astype(np.int64) - 1
    inside = (px >= 0) & (py >= 0) & (px < m.shape[0]) & (py < m.shape[1])
    node_lab = np.zeros(N, dtype=np.int64)
    node_lab[inside] = lab[px[inside], py[inside]]
    good = np.isfinite(U[:, 0])
    for k in range(1, ncomp + 1):
        sel = np.flatnonzero(good & (node_lab == k))
        if sel.size == 0:
            continue  # the reference's try fails on an empty set and skips it
        Uc, Fc = comp_def_grad2(U[sel], c[sel], rad, m)
        U[sel] = Uc.reshape(-1, 2)
        F[sel] = Fc.reshape(-1, 4)
    # rmoutliers on each F component; a flagged node loses all four
    flag = np.zeros(N, dtype=bool)
    for q in range(4):
        flag |= isoutlier_movmedian(F[:, q], 1 + step)
    F[flag] = np.nan
    okU = np.isfinite(U[:, 0])
    okF = np.isfinite(F[:, 0])
    if not okU.all() or not okF.all():
        U = _nearest_fill(c, U, okU)
        F = _nearest_fill(c, F, okF)
    return U.reshape(-1), F.reshape(-1)


def subpb1_gridfit(U, coordinatesFEM):
    """Subpb1.m:186-199: gridfit of u and v through the nodes with a finite u
    on unique(x) x unique(y), read back at every node."""
    c = np.asarray(coordinatesFEM, dtype=np.float64)
    N = c.shape[0]
    U = np.asarray(U, dtype=np.float64).reshape(N, 2)
    xn = np.unique(c[:, 0])
    yn = np.unique(c[:, 1])
    ok = np.isfinite(U[:, 0])
    ix = np.searchsorted(xn, c[:, 0])
    iy = np.searchsorted(yn, c[:, 1])
    out = np.empty((N, 2))
    for k in range(2):
        z, _, _ = gridfit(c[ok, 0], c[ok, 1], U[ok, k], xn, yn)  # rows follow yn
        out[:, k] = z[iy, ix]
    return out.reshape(-1)
