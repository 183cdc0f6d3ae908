"""Host-side mirror of the reference's MATLAB interface for the hot path.

Each function keeps the name, argument order and outputs of the MATLAB function
it replaces and runs its compute on the GPU through libstaqdic.so.  Arrays use
the reference's layout: images are ``(H, W)`` arrays indexed ``[x-1, y-1]``
(MATLAB's arrays after the transpose on load), coordinates are 1-based,
``coordinatesFEM`` is ``N x 2``, ``U`` is ``2N`` interleaved and ``F`` is ``4N``.

What stays on the host is the reference's bookkeeping around the per-subset
solver: bad-subset classification (LocalICGNQuadtree.m:151-177,
Subpb1Quadtree.m:168-187) and the refill of bad subsets (refill.py), the hole
mask of the integer search (IntegerSearch.m:69-77) and timing.  The plain
twins' refills (LocalICGN's plane fit, rmoutliers and nearest fill, Subpb1's
gridfit) are in refill_plain.py.
"""
from __future__ import annotations

import time

import numpy as np

from . import _lib
from .outliers import funRemoveOutliers
from .refill import refill_bad
from .refill_plain import local_icgn_refill, subpb1_gridfit
from ._lib import Batch, Context, IN_FOLD, IN_U0, IN_UOLD, IN_VDUAL, d, ip, lp

__all__ = [
    "default_context", "ba_interp2", "funImgGradient", "funIntegerSearch", "IntegerSearch",
    "LocalICGNQuadtree", "LocalICGN", "Subpb1Quadtree", "Subpb1", "local_icgn_raw",
    "subpb1_raw", "classify_local", "classify_subpb1",
]

_ctx = None


def default_context() -> Context:
    global _ctx
    if _ctx is None:
        _ctx = Context(0)
    return _ctx


def _c(ctx):
    return ctx if ctx is not None else default_context()


# ---------------------------------------------------------------------------
# ba_interp2.cpp:318-335
# ---------------------------------------------------------------------------
def ba_interp2(F, X, Y, method="cubic", ctx=None):
    """Z = ba_interp2(F, X, Y[, 'cubic']).  F is M x N (x P); X is the dim-2
    (column) and Y the dim-1 (row) coordinate, 1-based.  float32 inputs give a
    float32 result, float64 a float64 one; mixed classes are an error, as in
    the MEX file."""
    if method not in ("cubic", None):
        raise ValueError("only the cubic method is implemented (ba_interp2.cpp keeps cubic only)")
    F = np.asarray(F)
    X = np.asarray(X)
    Y = np.asarray(Y)
    if X.shape != Y.shape:
        raise ValueError("X, Y should have the same size")
    dt = {F.dtype, X.dtype, Y.dtype}
    if dt not in ({np.dtype(np.float64)}, {np.dtype(np.float32)}):
        raise TypeError("ba_interp2 takes only double or single arguments for IMG,X,Y, and all "
                        "must have the same class")
    c = _c(ctx)
    M, N = F.shape[0], F.shape[1]
    P = int(np.prod(F.shape[2:])) if F.ndim > 2 else 1
    Ff = np.asfortranarray(F)
    Xf = np.asfortranarray(X)
    Yf = np.asfortranarray(Y)
    out_shape = X.shape + F.shape[2:]
    Z = np.zeros(out_shape, dtype=F.dtype, order="F")
    if F.dtype == np.float64:
        rc = c.L.staq_ba_interp2(c.h, d(Ff), M, N, P, d(Xf), d(Yf), X.size, d(Z))
    else:
        rc = c.L.staq_ba_interp2_f32(c.h, _lib.f32(Ff), M, N, P, _lib.f32(Xf), _lib.f32(Yf),
                                     X.size, _lib.f32(Z))
    c.check(rc, "ba_interp2")
    return Z


# ---------------------------------------------------------------------------
# func/funImgGradient.m
# ---------------------------------------------------------------------------
def funImgGradient(ImgRef, ImgDef=None, ImgRefMask=None, ctx=None):
    """Df = funImgGradient(ImgRef, ImgDef, ImgRefMask): 7-tap central
    differences on the GPU; returns the reference's Df struct as a dict."""
    c = _c(ctx)
    f = np.asfortranarray(ImgRef, dtype=np.float64)
    H, W = f.shape
    mask = np.ones((H, W)) if ImgRefMask is None else np.asfortranarray(ImgRefMask, dtype=np.float64)
    dx = np.zeros((H - 6, W - 6), order="F")
    dy = np.zeros((H - 6, W - 6), order="F")
    c.check(c.L.staq_img_gradient(c.h, d(f), d(mask), H, W, d(dx), d(dy)), "funImgGradient")
    return {"DfDx": dx, "DfDy": dy, "DfCropWidth": 3, "DfAxis": [4, H - 3, 4, W - 3],
            "imgSize": (H, W), "ImgRefMask": mask}


# ---------------------------------------------------------------------------
# func/funIntegerSearch.m, func/IntegerSearch.m
# ---------------------------------------------------------------------------
def funIntegerSearch(f, g, tempSizeOfSearchRegion, gridx, gridy, winsize, winstepsize,
                     tempNoOfInitPt=0, *varargin, ctx=None, slot=0, qfactors=True):
    """[x,y,u,v,cc] = funIntegerSearch(f, g, r, gridx, gridy, winsize, winstepsize, 0).
    Whole-field mode (tempNoOfInitPt = 0).  cc is a dict with 'max' (ny x nx)
    and 'qfactors' (one row per search point, x fastest)."""
    if tempNoOfInitPt != 0:
        raise NotImplementedError("seed-point mode needs interactive input (funIntegerSearch.m:37-74)")
    c = _c(ctx)
    c.set_frame(f, g, None, slot=slot)
    return integer_search_slot(c, slot, tempSizeOfSearchRegion, gridx, gridy, winsize,
                               winstepsize, qfactors=qfactors)


def xcorr_qfactors(c, slot, box_f, box_g, fail=np.inf, normalise=True):
    """compute_qFactor (funIntegerSearch.m:211-264) for each box on the frame
    pair in slot ``slot``: C = f(box_f), D = g(box_g), boxes (n, 4) 1-based
    inclusive [x0 x1 y0 y1].  A box the reference's try rejects takes
    ``fail`` (Inf in funIntegerSearch, :156; NaN in funIntegerSearchMg,
    :185).  With ``normalise``, each column becomes (q - min q) / max(q -
    min q), NaN left out of min and max, as :174-177 and
    funIntegerSearchMg.m:192-196 do."""
    bf = np.ascontiguousarray(np.asarray(box_f, dtype=np.int64).reshape(-1, 4))
    bg = np.ascontiguousarray(np.asarray(box_g, dtype=np.int64).reshape(-1, 4))
    n = bf.shape[0]
    q = np.zeros(2 * n)
    ok = np.zeros(n, dtype=np.int32)
    c.check(c.L.staq_xcorr_qfactors(c.h, slot, n, lp(bf), lp(bg), d(q), ip(ok)),
            "compute_qFactor")
    q = q.reshape(n, 2)
    q[ok == 0] = fail
    if normalise and n:
        with np.errstate(invalid="ignore", divide="ignore"):
            for k in range(2):
                qf = q[:, k] - np.nanmin(q[:, k])
                q[:, k] = qf / np.nanmax(qf)
    return q


def integer_search_slot(c, slot, tempSizeOfSearchRegion, gridx, gridy, winsize, winstepsize,
                        qfactors=False):
    """funIntegerSearch on the frame pair already resident in context slot
    ``slot`` (f = its reference image, g = its deformed image).  With
    ``qfactors`` cc['qfactors'] holds the normalised [pce ppe] of every
    search point in the reference's order (x fastest)."""
    r = np.broadcast_to(np.asarray(tempSizeOfSearchRegion, dtype=np.int64), (2,))
    w = int(np.asarray(winsize).ravel()[0])
    step = int(np.asarray(winstepsize).ravel()[0])
    H, W = c.frames[slot]
    mx = (H // step + 2) * (W // step + 2)
    gx = np.asarray([int(gridx[0]), int(gridx[-1])], dtype=np.int64)
    gy = np.asarray([int(gridy[0]), int(gridy[-1])], dtype=np.int64)
    outs = [np.zeros(mx) for _ in range(5)]
    import ctypes as C
    nx, ny = C.c_long(), C.c_long()
    c.check(c.L.staq_integer_search(c.h, slot, int(r[0]), int(r[1]), lp(gx), lp(gy), w, step, mx,
                                    C.byref(nx), C.byref(ny), *[d(o) for o in outs]),
            "funIntegerSearch")
    n = nx.value * ny.value
    x, y, u, v, cmax = [o[:n].reshape((ny.value, nx.value), order="F") for o in outs]
    cc = {"max": cmax, "A": []}
    if qfactors:
        # point tempi = ci + nx (cj - 1) sits at x(cj, ci) = ii + w/2 (:194-195)
        ii = np.rint(x.ravel() - 0.5 * w).astype(np.int64)
        jj = np.rint(y.ravel() - 0.5 * w).astype(np.int64)
        bf = np.stack([ii, ii + w, jj, jj + w], axis=1)
        bg = np.stack([ii - r[0], ii + w + r[0], jj - r[1], jj + w + r[1]], axis=1)
        cc["qfactors"] = xcorr_qfactors(c, slot, bf, bg, fail=np.inf)
    return x, y, u, v, cc


def IntegerSearch(ImgRef, ImgDef, file_name=None, DICpara=None, ctx=None):
    """[DICpara,x0,y0,u,v,cc] = IntegerSearch(ImgRef, ImgDef, file_name, DICpara).

    DICpara['InitFFTSearchMethod'] 1 (the default, IntegerSearch.m:29-35):
    whole-field search with radius DICpara['SizeOfFFTSearchRegion'] (the
    reference hard-codes 6, :49), then the median-test outlier filter
    (:116-119).  0: the multigrid search funIntegerSearchMg (:122-125, as
    main_ALDIC_Quadtree_inc.m:91 sets it), no outlier filter.  2 (seed
    points) needs interactive input and is not provided.  Either way grid
    points on a hole of DICpara['ImgRefMask'] are NaN (:69-77, :128-135)."""
    P = dict(DICpara or {})
    H, W = np.shape(ImgRef)
    roi = P.get("gridxyROIRange", {"gridx": [1, H], "gridy": [1, W]})
    method = P.get("InitFFTSearchMethod", 1)
    if method == 0:
        from .mgsearch import funIntegerSearchMg
        r = 0
        x0, y0, u, v, cc = funIntegerSearchMg(ImgRef, ImgDef, roi["gridx"], roi["gridy"],
                                              P["winsize"], P["winstepsize"], ctx=ctx)
    elif method == 1:
        r = P.get("SizeOfFFTSearchRegion", 6)
        x0, y0, u, v, cc = funIntegerSearch(ImgRef, ImgDef, r, roi["gridx"], roi["gridy"],
                                            P["winsize"], P["winstepsize"], 0, ctx=ctx)
    else:
        raise NotImplementedError("seed-point search needs interactive input (IntegerSearch.m:58-64)")
    mask = P.get("ImgRefMask")
    if mask is not None and x0.size:
        xi = np.round(x0).astype(np.int64).ravel()
        yi = np.round(y0).astype(np.int64).ravel()
        m = np.asarray(mask)[xi - 1, yi - 1]
        hole = (m == 0).reshape(x0.shape)
        u = np.where(hole, np.nan, u)
        v = np.where(hole, np.nan, v)
    # median-test outlier filter, threshold 100 (IntegerSearch.m:116-119)
    if method == 1 and np.isfinite(u).sum() >= 4 and min(np.shape(u)) >= 2:
        u, v, cc, _, _, _ = funRemoveOutliers(u, v, cc, 0, 100)
    P["InitFFTSearchMethod"] = method
    P["SizeOfFFTSearchRegion"] = r
    return P, x0, y0, u, v, cc


# ---------------------------------------------------------------------------
# Local IC-GN (LocalICGNQuadtree / LocalICGN) and the AL step (Subpb1*)
# ---------------------------------------------------------------------------
def _winsizes(DICpara, N):
    return np.broadcast_to(np.asarray(DICpara["winsize"], dtype=np.int64), (N,)).copy()


def local_icgn_raw(U0, coords, winsize, f, g, mask=None, tol=1e-2, plain=False, ctx=None,
                   slot=None, **kw):
    """Raw per-subset GPU outputs: U (2N), F (4N), H (N x 21), status (N)."""
    c = _c(ctx)
    if f is not None:
        c.set_frame(f, g, mask, slot=0)
    coords = np.asfortranarray(coords, dtype=np.float64)
    b = Batch(c, coords, winsize, slot=slot)
    try:
        b.set_input(IN_U0, U0)
        b.icgn6(_lib.params(tol=tol, plain=plain, **kw))
        return b.outputs()
    finally:
        b.close()


def subpb1_raw(UOld, FOld, vdual, coords, wx, wy, mu, f, g, mask=None, tol=1e-2, plain=False,
               ctx=None, slot=None, **kw):
    """Raw per-subset GPU outputs of the 2-DOF AL step: U (2N), status (N)."""
    c = _c(ctx)
    if f is not None:
        c.set_frame(f, g, mask, slot=0)
    coords = np.asfortranarray(coords, dtype=np.float64)
    b = Batch(c, coords, wx, wy, slot=slot)
    try:
        b.set_input(IN_UOLD, UOld)
        b.set_input(IN_FOLD, FOld)
        b.set_input(IN_VDUAL, vdual)
        b.icgn2(_lib.params(tol=tol, mu=mu, plain=plain, **kw))
        U, _, _, st = b.outputs()
        return U, st
    finally:
        b.close()


def classify_local(status, plain=False):
    """Bad-subset set of LocalICGNQuadtree.m:151-169 (plain: LocalICGN.m:119-131).
    Returns (bad mask, LocalICGNBadPtNum)."""
    st = np.asarray(status)
    bad = (st < 0) | (st > 99)
    # 102 (mostly masked) is bad but not counted (LocalICGNQuadtree.m:153-157, LocalICGN.m:119-122)
    n_bad = int(bad.sum()) - int((st == 102).sum())
    good = ~bad
    if good.sum() >= 1:
        g = st[good].astype(float)
        mu = g.mean()
        sd = g.std(ddof=1) if g.size > 1 else 0.0
        thr = mu + 0.15 * sd if plain else max(mu + sd, 6.0)
        bad |= st > thr
    return bad, n_bad


def classify_subpb1(status, plain=False):
    """Bad-subset set of Subpb1Quadtree.m:173-186 (plain: Subpb1.m)."""
    st = np.asarray(status)
    bad = (st < 0) | (st > 99)
    if plain:
        return bad, int(bad.sum())
    n_bad = int(bad.sum()) - int((st == 102).sum())
    good = ~bad
    if good.sum() >= 1:
        g = st[good].astype(float)
        thr = max(g.mean() + 0.25 * (g.std(ddof=1) if g.size > 1 else 0.0), 10.0)
        bad |= st > thr
    return bad, n_bad


def _hole_nodes(coords, winsize, f, mask):
    """markCoordHoleStrain (LocalICGNQuadtree.m:71-82): nodes whose window holds
    a zero of f*mask, or cannot be read.  1-based indices."""
    H, W = f.shape
    # zeros of f*mask counted over each window from one summed-area table
    z = np.zeros((H + 1, W + 1), dtype=np.int64)
    z[1:, 1:] = np.cumsum(np.cumsum(np.asarray(f * mask) == 0, axis=0), axis=1)
    xy = np.round(np.asarray(coords, dtype=np.float64)).astype(np.int64)
    x, y = xy[:, 0], xy[:, 1]
    h = np.asarray(winsize, dtype=np.int64) // 2
    outside = (x - h < 1) | (y - h < 1) | (x + h > H) | (y + h > W)
    a0 = np.clip(x - h - 1, 0, H)
    a1 = np.clip(x + h, 0, H)
    b0 = np.clip(y - h - 1, 0, W)
    b1 = np.clip(y + h, 0, W)
    zeros = z[a1, b1] - z[a0, b1] - z[a1, b0] + z[a0, b0]
    return (np.flatnonzero(outside | (zeros > 0)) + 1).astype(np.int64)


def LocalICGNQuadtree(U0, coordinatesFEM, Df, ImgRef, ImgDef, DICpara, ICGNmethod="GaussNewton",
                      tol=1e-2, ctx=None, refill=True, raw=None):
    """[U,F,HtempPar,LocalTime,ConvItPerEle,LocalICGNBadPtNum,markCoordHoleStrain] =
    LocalICGNQuadtree(U0,coordinatesFEM,Df,ImgRef,ImgDef,DICpara,ICGNmethod,tol).
    Bad subsets are refilled as the reference does (refill=False leaves them NaN).
    raw: the GPU step's (U, F, H, status) when the caller ran it already (a
    resident batch), else it runs here."""
    if ICGNmethod != "GaussNewton":
        raise NotImplementedError("only the Gauss-Newton scheme runs on this path")
    coords = np.asarray(coordinatesFEM, dtype=np.float64)
    N = coords.shape[0]
    mask = Df.get("ImgRefMask") if isinstance(Df, dict) else None
    win = _winsizes(DICpara, N)
    t0 = time.perf_counter()
    if raw is None:
        U, F, H, st = local_icgn_raw(U0, coords, win, ImgRef, ImgDef, mask, tol=tol, ctx=ctx)
    else:
        U, F, H, st = raw
    LocalTime = time.perf_counter() - t0
    bad, nbad = classify_local(st)
    U = U.copy()
    F = F.copy()
    U[2 * np.flatnonzero(bad)] = np.nan
    U[2 * np.flatnonzero(bad) + 1] = np.nan
    for k in range(4):
        F[4 * np.flatnonzero(bad) + k] = np.nan
    m = np.ones(np.shape(ImgRef)) if mask is None else np.asarray(mask)
    if refill:
        U, F = refill_bad(U, F, coords, m)  # LocalICGNQuadtree.m:177-333
    hole = _hole_nodes(coords, win, np.asarray(ImgRef), m)
    return U, F, H, LocalTime, st.astype(np.float64), nbad, hole


def LocalICGN(U0, coordinatesFEM, Df, ImgRef, ImgDef, DICpara, ICGNmethod="GaussNewton", tol=1e-2,
              ctx=None, refill=True):
    """[U,F,HtempPar,LocalTime,ConvItPerEle,LocalICGNBadPtNum] = LocalICGN(...) —
    the uniform-grid twin (func/LocalICGN.m): bad subsets NaN, then the
    plane fit per mask component, the F outlier removal and the nearest fill
    of LocalICGN.m:144-292 (refill_plain.local_icgn_refill)."""
    if ICGNmethod != "GaussNewton":
        raise NotImplementedError("only the Gauss-Newton scheme runs on this path")
    coords = np.asarray(coordinatesFEM, dtype=np.float64)
    N = coords.shape[0]
    mask = Df.get("ImgRefMask") if isinstance(Df, dict) else None
    win = _winsizes(DICpara, N)
    t0 = time.perf_counter()
    U, F, H, st = local_icgn_raw(U0, coords, win, ImgRef, ImgDef, mask, tol=tol, plain=True,
                                 ctx=ctx)
    LocalTime = time.perf_counter() - t0
    bad, nbad = classify_local(st, plain=True)
    U = U.copy()
    F = F.copy()
    idx = np.flatnonzero(bad)
    U[2 * idx] = np.nan
    U[2 * idx + 1] = np.nan
    for k in range(4):
        F[4 * idx + k] = np.nan
    if refill:
        m = np.ones(np.shape(ImgRef)) if mask is None else np.asarray(mask)
        U, F = local_icgn_refill(U, F, coords, m, DICpara["winstepsize"])
    return U, F, H, LocalTime, st.astype(np.float64), nbad


def Subpb1Quadtree(UOld, FOld, udual, vdual, coordinatesFEM, Df, ImgRef, ImgDef, mu, beta, HPar,
                   ALSolveStep, DICpara, ICGNmethod="GaussNewton", tol=1e-2, ctx=None, refill=True,
                   raw=None):
    """[U,HPar,ALSub1Time,ConvItPerEle,LocalICGNBadPtNum] = Subpb1Quadtree(...).
    Window sizes come from DICpara['winsize_List'] (N x 2), else winsize.
    raw: the GPU step's (U, status) when the caller ran it already."""
    if ICGNmethod != "GaussNewton":
        raise NotImplementedError("only the Gauss-Newton scheme runs on this path")
    coords = np.asarray(coordinatesFEM, dtype=np.float64)
    N = coords.shape[0]
    wl = DICpara.get("winsize_List")
    if wl is None:
        wx = wy = _winsizes(DICpara, N)
    else:
        wl = np.asarray(wl, dtype=np.int64).reshape(N, 2)
        wx, wy = wl[:, 0].copy(), wl[:, 1].copy()
    mask = Df.get("ImgRefMask") if isinstance(Df, dict) else None
    t0 = time.perf_counter()
    if raw is None:
        U, st = subpb1_raw(UOld, FOld, vdual, coords, wx, wy, mu, ImgRef, ImgDef, mask, tol=tol,
                           ctx=ctx)
    else:
        U, st = raw
    ALSub1Time = time.perf_counter() - t0
    # a caught error leaves ConvItPerEle at its initial 0 here (Subpb1Quadtree.m:146-149)
    st = np.where(st == -1, 0, st)
    bad, nbad = classify_subpb1(st)
    U = U.copy()
    idx = np.flatnonzero(bad)
    U[2 * idx] = np.nan
    U[2 * idx + 1] = np.nan
    if refill:
        m = np.ones(np.shape(ImgRef)) if mask is None else np.asarray(mask)
        U, _ = refill_bad(U, None, coords, m)  # Subpb1Quadtree.m:189-256
    return U, HPar, ALSub1Time, st.astype(np.float64), nbad


def Subpb1(UOld, FOld, udual, vdual, coordinatesFEM, Df, ImgRef, ImgDef, mu, beta, HPar,
           ALSolveStep, DICpara, ICGNmethod="GaussNewton", tol=1e-2, ctx=None, refill=True):
    """[U,HPar,ALSub1Time,ConvItPerEle,LocalICGNBadPtNum] = Subpb1(...) — the
    uniform-grid twin (func/Subpb1.m, errors are -1): bad subsets NaN, then
    the gridfit of Subpb1.m:185-199 (refill_plain.subpb1_gridfit)."""
    coords = np.asarray(coordinatesFEM, dtype=np.float64)
    N = coords.shape[0]
    w = _winsizes(DICpara, N)
    mask = Df.get("ImgRefMask") if isinstance(Df, dict) else None
    t0 = time.perf_counter()
    U, st = subpb1_raw(UOld, FOld, vdual, coords, w, w, mu, ImgRef, ImgDef, mask, tol=tol,
                       plain=True, ctx=ctx)
    ALSub1Time = time.perf_counter() - t0
    bad, nbad = classify_subpb1(st, plain=True)
    U = U.copy()
    idx = np.flatnonzero(bad)
    U[2 * idx] = np.nan
    U[2 * idx + 1] = np.nan
    if refill:
        U = subpb1_gridfit(U, coords)
    return U, HPar, ALSub1Time, st.astype(np.float64), nbad
