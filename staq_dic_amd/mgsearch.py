"""Multigrid integer search (func/funIntegerSearchMg.m:3-285) and the surface
fit it ends with (func/gridfit.m, the 'triangle' interpolation with the
'springs' regularizer).

The correlations run on the GPU through the C ABI: the whole-region
normxcorr2 of level 1 (``staq_xcorr_full_peak``) and, per level, every box's
+-5 px normxcorr2 with findpeak in one call (``staq_xcorr_boxes``).  The box
bookkeeping between levels is host control logic, and gridfit is a sparse
least-squares solve on the host in FP64, as the reference keeps both in
MATLAB.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import _lib
from ._lib import d, lp

__all__ = ["gridfit", "funIntegerSearchMg"]

PAD = 5  # D = g(box +- 5) at levels >= 2 (funIntegerSearchMg.m:166)


def _mround(x):
    """MATLAB round (half away from zero) of a non-negative number."""
    return float(np.floor(x + 0.5))


# ---------------------------------------------------------------------------
# gridfit (func/gridfit.m:1-700): defaults smoothness 1, interp 'triangle',
# solver backslash, autoscale on, extend 'warning'; regularizer as given.
# ---------------------------------------------------------------------------
def gridfit(x, y, z, xnodes, ynodes, regularizer="springs", smoothness=1.0):
    """[zgrid, xgrid, ygrid] = gridfit(x, y, z, xnodes, ynodes, 'regularizer', ...).

    zgrid is ny x nx (rows follow ynodes).  Data outside the node range move
    the first / last node out to them (gridfit.m:394-432, 'extend' =
    'warning').  The least-squares system is solved through its normal
    equations with a sparse LU (the reference: MATLAB's sparse QR)."""
    if regularizer != "springs":
        raise NotImplementedError("only the 'springs' regularizer of funIntegerSearchMg is restated")
    x = np.asarray(x, dtype=float).ravel()
    y = np.asarray(y, dtype=float).ravel()
    z = np.asarray(z, dtype=float).ravel()
    keep = ~(np.isnan(x) | np.isnan(y) | np.isnan(z))
    x, y, z = x[keep], y[keep], z[keep]
    n = x.size
    if n < 3:
        raise ValueError("Insufficient data for surface estimation.")
    xnodes = np.array(xnodes, dtype=float).ravel()
    ynodes = np.array(ynodes, dtype=float).ravel()
    xnodes[0] = min(xnodes[0], x.min())
    xnodes[-1] = max(xnodes[-1], x.max())
    ynodes[0] = min(ynodes[0], y.min())
    ynodes[-1] = max(ynodes[-1], y.max())
    dx = np.diff(xnodes)
    dy = np.diff(ynodes)
    nx, ny = xnodes.size, ynodes.size
    ngrid = nx * ny
    xscale, yscale = dx.mean(), dy.mean()

    # cell of each point (histc, the last node folded into the last cell), 0-based
    ix = np.searchsorted(xnodes, x, side="right") - 1
    iy = np.searchsorted(ynodes, y, side="right") - 1
    ix[ix == nx - 1] = nx - 2
    iy[iy == ny - 1] = ny - 2
    ind = iy + ny * ix
    tx = np.clip((x - xnodes[ix]) / dx[ix], 0.0, 1.0)
    ty = np.clip((y - ynodes[iy]) / dy[iy], 0.0, 1.0)
    # triangle interpolation (:484-493)
    k = tx > ty
    L = np.where(k, ny, 1)
    t1 = np.minimum(tx, ty)
    t2 = np.maximum(tx, ty)
    rows = np.repeat(np.arange(n), 3)
    cols = np.stack([ind, ind + ny + 1, ind + L], axis=1).ravel()
    vals = np.stack([1 - t2, t1, t2 - t1], axis=1).ravel()
    A = sp.csr_matrix((vals, (rows, cols)), shape=(n, ngrid))

    # springs (:523-549): neighbours along y, along x, and both diagonals
    blocks = []
    I, J = np.meshgrid(np.arange(nx), np.arange(ny - 1))
    i_, j_ = I.ravel(order="F"), J.ravel(order="F")
    ii = j_ + ny * i_
    st = 1.0 / (dy[j_] / yscale)
    blocks.append((ii, ii + 1, st))
    I, J = np.meshgrid(np.arange(nx - 1), np.arange(ny))
    i_, j_ = I.ravel(order="F"), J.ravel(order="F")
    ii = j_ + ny * i_
    st = 1.0 / (dx[i_] / xscale)
    blocks.append((ii, ii + ny, st))
    I, J = np.meshgrid(np.arange(nx - 1), np.arange(ny - 1))
    i_, j_ = I.ravel(order="F"), J.ravel(order="F")
    ii = j_ + ny * i_
    st = 1.0 / np.sqrt((dx[i_] / xscale) ** 2 + (dy[j_] / yscale) ** 2)
    blocks.append((ii, ii + ny + 1, st))
    blocks.append((ii + 1, ii + ny, st))
    r0 = 0
    rr, cc, vv = [], [], []
    for a, b, s in blocks:
        m = a.size
        r = np.arange(r0, r0 + m)
        rr += [r, r]
        cc += [a, b]
        vv += [-s, s]
        r0 += m
    Areg = sp.csr_matrix((np.concatenate(vv), (np.concatenate(rr), np.concatenate(cc))),
                         shape=(r0, ngrid))
    # relative weight (:596-600): norm(., 1) is the largest column sum
    NA = abs(A).sum(axis=0).max()
    NR = abs(Areg).sum(axis=0).max()
    M = sp.vstack([A, Areg * (smoothness * NA / NR)]).tocsc()
    rhs = np.concatenate([z, np.zeros(r0)])
    sol = spla.spsolve((M.T @ M).tocsc(), M.T @ rhs)
    zgrid = sol.reshape((ny, nx), order="F")
    xgrid, ygrid = np.meshgrid(xnodes, ynodes)
    return zgrid, xgrid, ygrid


# ---------------------------------------------------------------------------
# funIntegerSearchMg
# ---------------------------------------------------------------------------
def _full_peak(c, slot, bf, bg):
    out = np.zeros(3)
    a = np.asarray(bf, dtype=np.int64)
    b = np.asarray(bg, dtype=np.int64)
    c.check(c.L.staq_xcorr_full_peak(c.h, slot, lp(a), lp(b), d(out)), "staq_xcorr_full_peak")
    return out


def _boxes(c, slot, bf, bg):
    nb = bf.shape[0]
    out = np.zeros(3 * nb)
    ok = np.zeros(nb, dtype=np.int32)
    bf = np.ascontiguousarray(bf, dtype=np.int64)
    bg = np.ascontiguousarray(bg, dtype=np.int64)
    c.check(c.L.staq_xcorr_boxes(c.h, slot, nb, lp(bf), lp(bg), PAD, d(out),
                                 ok.ctypes.data_as(C.POINTER(C.c_int))), "staq_xcorr_boxes")
    return out.reshape(nb, 3), ok


def _fit_boxes(H, W, gap, xf, yf, xg, yg):
    """Move g boxes that cross the border gap back inside and shrink the f
    boxes alike, then make every width even (funIntegerSearchMg.m:30-41,
    :221-254).  Arrays are nbox x 2, modified in place; NaN boxes stay NaN."""
    with np.errstate(invalid="ignore"):
        k = xg[:, 0] < gap
        xf[k, 0] += gap - xg[k, 0]
        xg[k, 0] = gap
        k = yg[:, 0] < gap
        yf[k, 0] += gap - yg[k, 0]
        yg[k, 0] = gap
        k = xg[:, 1] > H - gap
        xf[k, 1] += (H - gap) - xg[k, 1]
        xg[k, 1] = H - gap
        k = yg[:, 1] > W - gap
        yf[k, 1] += (W - gap) - yg[k, 1]
        yg[k, 1] = W - gap
        k = np.mod(xf[:, 1] - xf[:, 0], 2) == 1
        xf[k, 1] -= 1
        xg[k, 1] -= 1
        k = np.mod(yf[:, 1] - yf[:, 0], 2) == 1
        yf[k, 1] -= 1
        yg[k, 1] -= 1


def _split(lo_hi, on):
    """Halves of each interval where ``on``, else the interval twice."""
    mid = lo_hi.mean(axis=1)
    first = np.where(on[:, None], np.stack([lo_hi[:, 0], mid], axis=1), lo_hi)
    second = np.where(on[:, None], np.stack([mid, lo_hi[:, 1]], axis=1), lo_hi)
    return first, second


def mg_scatter(c, slot, H, W, gridx, gridy, winsize, last_boxes=None):
    """The level loop of funIntegerSearchMg (:3-246) on frame ``slot``:
    box centres, shifts and peak values of the last level, and the grid
    limits after the level-1 border gap.  ``last_boxes``, a dict, receives
    the last level's template boxes ('f') and g boxes with their +-5 ring
    ('g', -1 where the box left the image)."""
    gx = np.array([gridx[0], gridx[-1]], dtype=float)
    gy = np.array([gridy[0], gridy[-1]], dtype=float)
    gap = 1 + _mround(0.75 * winsize)
    gx[0] = max(gx[0], gap)
    gy[0] = max(gy[0], gap)
    gx[1] = min(gx[1], H - gap)
    gy[1] = min(gy[1], W - gap)
    box = [gx[0], gx[1], gy[0], gy[1]]
    xpk, ypk, _ = _full_peak(c, slot, box, box)
    # zero_disp = ceil((size(X) + 1) / 2), size(X) = 2 n - 1
    u = ypk - np.ceil((2 * (gx[1] - gx[0] + 1) - 1 + 1) / 2.0)
    v = xpk - np.ceil((2 * (gy[1] - gy[0] + 1) - 1 + 1) / 2.0)
    xf, yf = gx[None, :].copy(), gy[None, :].copy()
    xg, yg = xf + np.ceil(u), yf + np.ceil(v)
    _fit_boxes(H, W, gap, xf, yf, xg, yg)
    wx, wy = xf[0, 1] - xf[0, 0], yf[0, 1] - yf[0, 0]
    uc, vc = np.ceil(np.array([u])), np.ceil(np.array([v]))
    while True:
        ratio = wx / wy
        if ratio > 2:  # split along x only (:85-107)
            on = (xf[:, 1] - xf[:, 0]) > winsize
            a, b = _split(xf, on)
            ga, gb = _split(xg, on)
            xf2, xg2 = np.stack([a, b], 1).reshape(-1, 2), np.stack([ga, gb], 1).reshape(-1, 2)
            yf2, yg2 = np.repeat(yf, 2, axis=0), np.repeat(yg, 2, axis=0)
            un, vn = np.repeat(uc, 2), np.repeat(vc, 2)
            wx = wx / 2
        elif ratio < 0.5:  # split along y only (:108-130)
            on = (yf[:, 1] - yf[:, 0]) > winsize
            a, b = _split(yf, on)
            ga, gb = _split(yg, on)
            yf2, yg2 = np.stack([a, b], 1).reshape(-1, 2), np.stack([ga, gb], 1).reshape(-1, 2)
            xf2, xg2 = np.repeat(xf, 2, axis=0), np.repeat(xg, 2, axis=0)
            un, vn = np.repeat(uc, 2), np.repeat(vc, 2)
            wy = wy / 2
        else:  # split both ways, order (lo x, lo y), (hi x, lo y), (lo x, hi y), (hi x, hi y)
            on = ((xf[:, 1] - xf[:, 0]) > winsize) & ((yf[:, 1] - yf[:, 0]) > winsize)
            xa, xb = _split(xf, on)
            gxa, gxb = _split(xg, on)
            ya, yb = _split(yf, on)
            gya, gyb = _split(yg, on)
            xf2 = np.stack([xa, xb, xa, xb], 1).reshape(-1, 2)
            xg2 = np.stack([gxa, gxb, gxa, gxb], 1).reshape(-1, 2)
            yf2 = np.stack([ya, ya, yb, yb], 1).reshape(-1, 2)
            yg2 = np.stack([gya, gya, gyb, gyb], 1).reshape(-1, 2)
            un, vn = np.repeat(uc, 4), np.repeat(vc, 4)
            wx, wy = wx / 2, wy / 2
        # every box of the level on the GPU (:151-170)
        finite = np.isfinite(xg2).all(1) & np.isfinite(yg2).all(1)
        bf = np.zeros((len(xf2), 4), dtype=np.int64)
        bg = np.zeros((len(xf2), 4), dtype=np.int64)
        bf[:] = np.stack([xf2[:, 0], xf2[:, 1], yf2[:, 0], yf2[:, 1]], 1)
        bg[finite] = np.stack([xg2[finite, 0], xg2[finite, 1], yg2[finite, 0], yg2[finite, 1]], 1)
        bg[~finite] = -1  # outside the image: the reference's catch
        res, ok = _boxes(c, slot, bf, bg)
        ok = ok.astype(bool)
        # zero_disp of the 13 x 13 crop: ceil(13 / 2) = pad + 2
        un = np.where(ok, un + (res[:, 1] - (PAD + 2)), np.nan)
        vn = np.where(ok, vn + (res[:, 0] - (PAD + 2)), np.nan)
        phi = np.where(ok, res[:, 2], 0.0)
        if wx < winsize and wy < winsize:
            if last_boxes is not None:
                last_boxes["f"] = bf
                last_boxes["g"] = np.where(finite[:, None], bg + np.array([-PAD, PAD, -PAD, PAD]), -1)
            break
        xf, yf = xf2, yf2
        with np.errstate(invalid="ignore"):
            xg = xf2 + np.ceil(un)[:, None]
            yg = yf2 + np.ceil(vn)[:, None]
        xf = xf.copy()
        yf = yf.copy()
        _fit_boxes(H, W, gap, xf, yf, xg, yg)
        uc, vc = np.ceil(un), np.ceil(vn)
    tx = 0.5 * (xf2[:, 0] + xf2[:, 1])
    ty = 0.5 * (yf2[:, 0] + yf2[:, 1])
    # grid limits (:256-261)
    ix1, ix2, iy1, iy2 = np.argmin(tx), np.argmax(tx), np.argmin(ty), np.argmax(ty)
    with np.errstate(invalid="ignore"):
        g = lambda s: np.ceil(1 + 1.25 * winsize + s)
        if gx[0] < g(un[ix1]):
            gx[0] = g(un[ix1])
        if gy[0] < g(vn[iy1]):
            gy[0] = g(vn[iy1])
        if gx[1] > H - g(un[ix2]) + 1:
            gx[1] = H - g(un[ix2]) + 1
        if gy[1] > W - g(vn[iy2]) + 1:
            gy[1] = W - g(vn[iy2]) + 1
    return tx, ty, un, vn, phi, gx, gy


def funIntegerSearchMg(f, g, gridx, gridy, winsize, winstepsize, *varargin, ctx=None, slot=0,
                       qfactors=True):
    """[xGrid, yGrid, uGrid, vGrid, cc] = funIntegerSearchMg(f, g, gridx, gridy,
    winsize, winstepsize), the InitFFTSearchMethod = 0 search of
    IntegerSearch.m:122-125.  cc has 'max' (the fitted peak values), 'A' = []
    and 'qfactors' (one row per box of the last level, :170-196)."""
    from .api import _c
    c = _c(ctx)
    c.set_frame(f, g, None, slot=slot)
    H, W = np.shape(f)
    return mg_search_slot(c, slot, H, W, gridx, gridy, winsize, winstepsize, qfactors=qfactors)


def mg_search_slot(c, slot, H, W, gridx, gridy, winsize, winstepsize, qfactors=False):
    """funIntegerSearchMg on the frame pair resident in context slot ``slot``;
    with ``qfactors`` cc['qfactors'] holds the normalised [pce ppe] of the
    last level's boxes, NaN where a box failed (:185, :192-196)."""
    from .api import xcorr_qfactors
    last = {} if qfactors else None
    tx, ty, tu, tv, phi, gx, gy = mg_scatter(c, slot, H, W, gridx, gridy, int(winsize),
                                             last_boxes=last)
    step = np.broadcast_to(np.asarray(winstepsize, dtype=float).ravel(), (2,))
    xl = np.arange(gx[0], gx[1] + 1e-9, step[0])
    yl = np.arange(gy[0], gy[1] + 1e-9, step[1])
    sel = (tx > xl.min()) & (tx < xl.max()) & (ty > yl.min()) & (ty < yl.max())
    u, xg, yg = gridfit(tx[sel], ty[sel], tu[sel], xl, yl)
    v, _, _ = gridfit(tx[sel], ty[sel], tv[sel], xl, yl)
    p, _, _ = gridfit(tx[sel], ty[sel], phi[sel], xl, yl)
    qf = None
    if qfactors:
        qf = xcorr_qfactors(c, slot, last["f"], last["g"], fail=np.nan)
    return xg, yg, u, v, {"max": p, "A": [], "qfactors": qf}
