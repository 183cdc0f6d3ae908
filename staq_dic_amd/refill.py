"""Refill of bad subsets after a local step, on the host (FP64).

LocalICGNQuadtree (func_quadtree/LocalICGNQuadtree.m:177-333) and
Subpb1Quadtree (func_quadtree/Subpb1Quadtree.m:189-256) replace the values of
bad subsets (NaN after the bad-subset classification) in two stages:

1. for every 8-connected component of the reference mask, a thin-plate radial
   basis function phi(r) = r^2 log(r + 1) with a linear polynomial term
   (rbfinterp/rbfcreate.m:148-200) is fitted to the good nodes whose rounded
   position lies in the component, and evaluated at every node of the
   component (good nodes get their own values back);
2. if NaNs remain (nodes outside every component, or a component with too few
   good nodes), every node is re-interpolated from the non-NaN nodes with
   scatteredInterpolant(..., 'natural', 'nearest').

Stage 1 solves the reference's dense (n+3) x (n+3) system [A P; P' 0]; the
matrix is the same for every field of a component, so it is factored once and
solved for all fields together.  That solve is O(n^3): past DENSE_MAX good
nodes in a component (the reference's own meshes hold about 1 500 nodes, the
large workloads here 10^4-10^5) each bad node instead takes the thin-plate
RBF through its LOCAL_K nearest good nodes of the component (local support),
and the good nodes keep their values, which the dense interpolant returns at
its centres.  Stage 2 is Sibson's natural-neighbour interpolation with
nearest-node extrapolation (``natural.natural_nearest``).
"""
from __future__ import annotations

import numpy as np
import scipy.linalg as sla
from scipy import ndimage
from scipy.spatial import cKDTree
from scipy.spatial.distance import cdist

from .natural import natural_nearest

__all__ = ["thinplate_fit", "thinplate_eval", "thinplate_fill_local", "refill_bad",
           "DENSE_MAX", "LOCAL_K"]

# the largest centre set that takes the reference's dense solve, and the
# neighbourhood of the local-support form beyond it
DENSE_MAX = 2000
LOCAL_K = 40


def _phi(r):
    return r * r * np.log(r + 1.0)


def thinplate_fit(centers, values):
    """Coefficients of rbfcreate(centers', values', 'RBFFunction', 'thinplate'):
    centers (n, 2), values (n, m) -> (n + 3, m)."""
    c = np.asarray(centers, dtype=np.float64)
    n = c.shape[0]
    r = cdist(c, c)
    A = np.zeros((n + 3, n + 3))
    A[:n, :n] = _phi(r)
    A[:n, n] = 1.0
    A[:n, n + 1:] = c
    A[n, :n] = 1.0
    A[n + 1:, :n] = c.T
    b = np.zeros((n + 3, values.shape[1]))
    b[:n] = values
    # NaN values give NaN coefficients, as MATLAB's backslash does (no error)
    return sla.solve(A, b, check_finite=False)


def thinplate_eval(centers, coeff, points):
    """rbfinterp(points', op): coefficient order [w_1..w_n, c0, cx, cy]."""
    c = np.asarray(centers, dtype=np.float64)
    p = np.asarray(points, dtype=np.float64)
    n = c.shape[0]
    r = cdist(p, c)
    return _phi(r) @ coeff[:n] + coeff[n] + p[:, :1] * coeff[n + 1] + p[:, 1:2] * coeff[n + 2]


def thinplate_fill_local(centers, values, targets, k=LOCAL_K):
    """At every target, the thin-plate RBF (with its linear term) through the
    target's k nearest centres, evaluated there: the local-support form of
    rbfcreate/rbfinterp (rbfcreate.m:148-200) for centre sets past DENSE_MAX.
    centers (n, 2), values (n, m), targets (t, 2) -> (t, m).  Linear fields
    are reproduced exactly, as by the dense form."""
    c = np.asarray(centers, dtype=np.float64)
    v = np.asarray(values, dtype=np.float64)
    p = np.asarray(targets, dtype=np.float64)
    if p.shape[0] == 0:
        return np.zeros((0, v.shape[1]))
    k = min(k, c.shape[0])
    _, nb = cKDTree(c).query(p, k=k)
    nb = nb.reshape(p.shape[0], k)
    out = np.zeros((p.shape[0], v.shape[1]))
    for s0 in range(0, p.shape[0], 4096):  # bounded memory: (4096, k+3, k+3) doubles
        sl = slice(s0, min(s0 + 4096, p.shape[0]))
        C = c[nb[sl]]  # (t, k, 2)
        t = C.shape[0]
        A = np.zeros((t, k + 3, k + 3))
        A[:, :k, :k] = _phi(np.linalg.norm(C[:, :, None, :] - C[:, None, :, :], axis=3))
        A[:, :k, k] = 1.0
        A[:, :k, k + 1:] = C
        A[:, k, :k] = 1.0
        A[:, k + 1:, :k] = np.transpose(C, (0, 2, 1))
        b = np.zeros((t, k + 3, v.shape[1]))
        b[:, :k] = v[nb[sl]]
        w = np.linalg.solve(A, b)
        r = np.linalg.norm(p[sl, None, :] - C, axis=2)  # (t, k)
        out[sl] = (np.einsum("tk,tkm->tm", _phi(r), w[:, :k]) + w[:, k] +
                   p[sl, :1] * w[:, k + 1] + p[sl, 1:2] * w[:, k + 2])
    return out


def refill_bad(U, F, coordinatesFEM, mask):
    """U (2N) and F (4N or None) with NaN at bad subsets -> refilled copies.
    mask is the reference mask (H, W) indexed [x-1, y-1]."""
    coords = np.asarray(coordinatesFEM, dtype=np.float64)
    N = coords.shape[0]
    U = np.array(U, dtype=np.float64).reshape(N, 2)
    Fm = None if F is None else np.array(F, dtype=np.float64).reshape(N, 4)
    m = np.asarray(mask) != 0
    H, W = m.shape
    lab, ncomp = ndimage.label(m, structure=np.ones((3, 3), dtype=bool))  # 8-connected
    rc = np.round(coords).astype(np.int64)
    inside = (rc[:, 0] >= 1) & (rc[:, 0] <= H) & (rc[:, 1] >= 1) & (rc[:, 1] <= W)
    comp = np.zeros(N, dtype=np.int64)
    comp[inside] = lab[rc[inside, 0] - 1, rc[inside, 1] - 1]
    # the centres of every fit are the nodes whose u is finite (notnanindex,
    # LocalICGNQuadtree.m:180-181), for F as for U (:228-249)
    good_u = np.isfinite(U[:, 0])
    good_f = good_u

    # stage 1: thin-plate RBF per mask component (:189-255)
    for k in range(1, ncomp + 1):
        members = np.flatnonzero(comp == k)
        if members.size == 0:
            continue
        for vals, good, cols in ((U, good_u, (0, 1)),) + (((Fm, good_f, (0, 1, 2, 3)),)
                                                          if Fm is not None else ()):
            src = members[good[members]]
            if src.size < 3:
                continue  # the reference's try/catch leaves the component as it is
            if src.size > DENSE_MAX:
                # local support: the bad members only (the good ones keep their values)
                dst = members[~good[members]]
                try:
                    vals[np.ix_(dst, list(cols))] = thinplate_fill_local(
                        np.round(coords[src]), vals[src][:, list(cols)], coords[dst])
                except np.linalg.LinAlgError:
                    continue
                continue
            try:
                coeff = thinplate_fit(np.round(coords[src]), vals[src][:, list(cols)])
            except (np.linalg.LinAlgError, ValueError):
                continue
            vals[np.ix_(members, list(cols))] = thinplate_eval(np.round(coords[src]), coeff,
                                                               coords[members])

    # stage 2: scatteredInterpolant over all nodes if NaNs remain (:313-330)
    nan_u = ~np.isfinite(U[:, 0])
    nan_f = np.zeros(N, dtype=bool) if Fm is None else ~np.isfinite(Fm[:, 0])
    if nan_u.any() or nan_f.any():
        U = natural_nearest(coords[~nan_u], U[~nan_u], coords)
        if Fm is not None:
            Fm = natural_nearest(coords[~nan_f], Fm[~nan_f], coords)
    return U.reshape(-1), (None if Fm is None else Fm.reshape(-1))
