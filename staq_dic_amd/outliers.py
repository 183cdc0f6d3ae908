"""Outlier filter of the integer search, on the host (FP64 / FP32 as noted).

``funRemoveOutliers(u, v, cc, 0, Thr0)`` (func/funRemoveOutliers.m:1-120, the
call at IntegerSearch.m:117-119): a normalised median test on the search grid
(funMedRemoveOutliers / medFilt2, :443-489), the points whose normalised
fluctuation magnitude exceeds Thr0 set to NaN and in-painted with the spring
method of ``inpaint_nans`` (func/inpaint_nans.m, method 4), the hole points put
back to NaN.  The q-factor branch (qDICOrNot = 1) is not restated.

Conventions kept from the reference:
  * the median test's input is first in-painted with ``inpaint_nans`` method 3
    (del^4 with reduced-order stencils near the edges);
  * ``medFilt2`` with a 2 x 2 window, symmetric padding after the array and the
    second neighbour skipped takes the median of three values (the element,
    the one below, the one below-right) and collects them in single
    precision, so the fluctuations and their normalisation are FP32;
  * ``inpaint_nans`` solves its sparse systems in the least-squares sense, as
    MATLAB's backslash does for a rectangular system.
Arrays are indexed like MATLAB's: A[r, c] with column-major linear indices.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

__all__ = ["inpaint_nans", "med_filt2", "funRemoveOutliers"]


def _lstsq_sparse(M, rhs):
    """Least-squares solution of the sparse system M x = rhs (normal
    equations; M has full column rank in every use here)."""
    M = sp.csr_matrix(M)
    A = (M.T @ M).tocsc()
    return spla.spsolve(A, M.T @ rhs)


def _nan_list(A):
    n, m = A.shape
    a = A.ravel(order="F")
    nan = np.flatnonzero(np.isnan(a))
    return a, nan, nan % n, nan // n


def _identify_neighbors(n, m, nr, nc, nan_idx, talks_to):
    rr = (nr[:, None] + talks_to[None, :, 0]).ravel()
    cc = (nc[:, None] + talks_to[None, :, 1]).ravel()
    ok = (rr >= 0) & (rr < n) & (cc >= 0) & (cc < m)
    idx = np.unique(rr[ok] + cc[ok] * n)
    return np.setdiff1d(idx, nan_idx)


def inpaint_nans(A, method):
    """inpaint_nans(A, method) for methods 3 (del^4) and 4 (springs)."""
    A = np.asarray(A, dtype=np.float64)
    n, m = A.shape
    a, nan, nr, nc = _nan_list(A)
    if nan.size == 0:
        return A.copy()
    nm = n * m
    known = np.setdiff1d(np.arange(nm), nan)
    if method == 3:
        talks = np.array([[-2, 0], [-1, -1], [-1, 0], [-1, 1], [0, -2], [0, -1], [0, 1], [0, 2],
                          [1, -1], [1, 0], [1, 1], [2, 0]])
        allv = np.concatenate([nan, _identify_neighbors(n, m, nr, nc, nan, talks)])
        r, c = allv % n, allv // n
        rows, cols, vals = [], [], []

        def add(sel, offs, coef):
            idx = allv[sel]
            for o, w in zip(offs, coef):
                rows.append(idx)
                cols.append(idx + o)
                vals.append(np.full(idx.size, float(w)))

        # r, c are 0-based here: the reference's 1-based bounds shift by one
        add((r >= 2) & (r <= n - 3) & (c >= 2) & (c <= m - 3),
            [-2 * n, -n - 1, -n, -n + 1, -2, -1, 0, 1, 2, n - 1, n, n + 1, 2 * n],
            [1, 2, -8, 2, 1, -8, 20, -8, 1, 2, -8, 2, 1])
        add((((r == 1) | (r == n - 2)) & (c >= 1) & (c <= m - 2)) |
            (((c == 1) | (c == m - 2)) & (r >= 1) & (r <= n - 2)),
            [-n, -1, 0, 1, n], [1, 1, -4, 1, 1])
        add(((r == 0) | (r == n - 1)) & (c >= 1) & (c <= m - 2), [-n, 0, n], [1, -2, 1])
        add(((c == 0) | (c == m - 1)) & (r >= 1) & (r <= n - 2), [-1, 0, 1], [1, -2, 1])
        fda = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                            shape=(nm, nm))
        rhs = -(fda[:, known] @ a[known])
        sub = fda[:, nan]
        k = np.flatnonzero(np.asarray(abs(sub).sum(axis=1)).ravel() > 0)
        x = _lstsq_sparse(sub[k], rhs[k])
    elif method == 4:
        pairs = []
        for dl, dr, dc in ((-1, -1, 0), (1, 1, 0), (-n, 0, -1), (n, 0, 1)):
            ok = (nr + dr >= 0) & (nr + dr < n) & (nc + dc >= 0) & (nc + dc < m)
            pairs.append(np.stack([nan[ok], nan[ok] + dl], axis=1))
        pairs = np.unique(np.sort(np.concatenate(pairs), axis=1), axis=0)
        nhv = pairs.shape[0]
        springs = sp.csr_matrix((np.tile([1.0, -1.0], nhv),
                                 (np.repeat(np.arange(nhv), 2), pairs.ravel())), shape=(nhv, nm))
        rhs = -(springs[:, known] @ a[known])
        x = _lstsq_sparse(springs[:, nan], rhs)
    else:
        raise NotImplementedError("only inpaint_nans methods 3 and 4 are restated")
    b = a.copy()
    b[nan] = x
    return b.reshape((n, m), order="F")


def med_filt2(V0):
    """medFilt2(V0, [2 2], 'symmetric', 2): median of (element, below,
    below-right), collected in single precision."""
    V0 = np.asarray(V0)
    V = np.pad(V0, ((0, 1), (0, 1)), mode="symmetric")
    n, m = V0.shape
    stack = np.stack([V[:n, :m], V[1:n + 1, :m], V[1:n + 1, 1:m + 1]], axis=2).astype(np.float32)
    return np.sort(stack, axis=2)[:, :, 1]


def _med_remove_outliers(u, epsilon=0.1):
    """funMedRemoveOutliers(u, epsilon) -> (medianU, normFluct), FP32."""
    u = inpaint_nans(u, 3)
    median_u = med_filt2(u)
    fluct = u.astype(np.float32) - median_u
    median_res = med_filt2(np.abs(fluct))
    norm = np.abs(fluct / (median_res + np.float32(epsilon)))
    return median_u, norm


def funRemoveOutliers(u, v, cc, qDICOrNot=0, Thr0=100):
    """[u,v,cc,BadptRow,BadptCol,RemoveOutliersList] = funRemoveOutliers(u,v,cc,0,Thr0)."""
    if qDICOrNot:
        raise NotImplementedError("the q-factor branch of funRemoveOutliers is not restated")
    if not Thr0:
        raise ValueError("an interactive threshold is not supported: pass Thr0 > 0")
    u = np.asarray(u, dtype=np.float64)
    v = np.asarray(v, dtype=np.float64)
    hole = np.isnan(u)
    _, nu = _med_remove_outliers(u)
    _, nv = _med_remove_outliers(v)
    mag = np.sqrt(nu * nu + nv * nv)
    out = np.flatnonzero((mag > Thr0).ravel(order="F"))
    u2 = u.copy().ravel(order="F")
    v2 = v.copy().ravel(order="F")
    u2[out] = np.nan
    v2[out] = np.nan
    u2 = inpaint_nans(u2.reshape(u.shape, order="F"), 4)
    v2 = inpaint_nans(v2.reshape(v.shape, order="F"), 4)
    u2[hole] = np.nan
    v2[hole] = np.nan
    return u2, v2, cc, np.zeros(0), np.zeros(0), out + 1
