"""ctypes binding of the FP64 CPU oracle (oracle.c) — TEST INFRASTRUCTURE ONLY.

Loaded by tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg,
never by the product package.  Arrays follow the reference's layout: images are
``(H, W)`` numpy arrays indexed ``[x-1, y-1]`` and passed column-major.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

_dp = C.POINTER(C.c_double)
_lp = C.POINTER(C.c_long)
_ip = C.POINTER(C.c_int)


class OrcFrame(C.Structure):
    _fields_ = [("f", _dp), ("g", _dp), ("mask", _dp), ("dfdx", _dp), ("dfdy", _dp),
                ("H", C.c_long), ("W", C.c_long)]


def build(force: bool = False) -> str:
    if force or not os.path.exists(_LIB) or os.path.getmtime(_LIB) < os.path.getmtime(
            os.path.join(_HERE, "oracle.c")):
        subprocess.check_call(["make", "-s", "-C", _HERE, "-B" if force else "all"])
    return _LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        L.orc_ba_interp2.argtypes = [_dp, C.c_long, C.c_long, _dp, _dp, C.c_long, _dp]
        L.orc_img_gradient.argtypes = [_dp, _dp, C.c_long, C.c_long, _dp, _dp]
        L.orc_bwselect.argtypes = [C.POINTER(C.c_ubyte), C.c_long, C.c_long, C.c_long, C.c_long,
                                   C.c_int, C.POINTER(C.c_ubyte)]
        L.orc_findpeak.argtypes = [_dp, C.c_long, C.c_long, _dp, _dp, _dp]
        L.orc_normxcorr2_kept.argtypes = [_dp, C.c_long, C.c_long, _dp, C.c_long, C.c_long, _dp]
        L.orc_normxcorr2_full.argtypes = [_dp, C.c_long, C.c_long, _dp, C.c_long, C.c_long, _dp]
        L.orc_integer_search.argtypes = [_dp, _dp, C.c_long, C.c_long, C.c_long, C.c_long, _lp, _lp,
                                         C.c_long, C.c_long, C.c_long, _lp, _lp, _dp, _dp, _dp,
                                         _dp, _dp]
        L.orc_integer_search.restype = C.c_int
        L.orc_icgn6.argtypes = [C.POINTER(OrcFrame), _dp, C.c_double, C.c_double, C.c_long,
                                C.c_double, C.c_int, _dp, _dp]
        L.orc_icgn6.restype = C.c_int
        L.orc_icgn2_al.argtypes = [C.POINTER(OrcFrame), C.c_double, C.c_double, C.c_long,
                                   C.c_long, C.c_double, _dp, _dp, _dp, C.c_double, _dp]
        L.orc_icgn2_al.restype = C.c_int
        L.orc_local_icgn_batch.argtypes = [C.POINTER(OrcFrame), _dp, _dp, _lp, C.c_long,
                                           C.c_double, C.c_int, C.c_int, _dp, _dp, _dp, _ip]
        L.orc_subpb1_batch.argtypes = [C.POINTER(OrcFrame), _dp, _dp, _dp, _dp, _lp, _lp,
                                       C.c_long, C.c_double, C.c_double, C.c_int, C.c_int, _dp,
                                       _ip]
        L.orc_znssd.argtypes = [C.POINTER(OrcFrame), _dp, _dp, _lp, C.c_long, _dp, _dp, _ip]
        L.orc_znssd.restype = C.c_double
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _F(a, dtype=np.float64):
    return np.asfortranarray(np.asarray(a, dtype=dtype))


class Frame:
    """Holds column-major copies of f, g, mask and the oracle's own gradient."""

    def __init__(self, f, g, mask=None):
        self.f = _F(f)
        self.g = _F(g)
        H, W = self.f.shape
        self.mask = _F(np.ones((H, W)) if mask is None else mask)
        self.dfdx, self.dfdy = img_gradient(self.f, self.mask)
        self.st = OrcFrame(_d(self.f), _d(self.g), _d(self.mask), _d(self.dfdx), _d(self.dfdy),
                           H, W)


def ba_interp2(F, X, Y):
    F = _F(F)
    X = _F(X)
    Y = _F(Y)
    out = np.empty_like(X)
    lib().orc_ba_interp2(_d(F), F.shape[0], F.shape[1] if F.ndim > 1 else 1, _d(X), _d(Y),
                         X.size, _d(out))
    return out


def img_gradient(img, mask=None):
    img = _F(img)
    H, W = img.shape
    mask = _F(np.ones((H, W)) if mask is None else mask)
    dfdx = np.zeros((H - 6, W - 6), order="F")
    dfdy = np.zeros((H - 6, W - 6), order="F")
    lib().orc_img_gradient(_d(img), _d(mask), H, W, _d(dfdx), _d(dfdy))
    return dfdx, dfdy


def bwselect(bw, c, r, conn):
    bw = np.asfortranarray(np.asarray(bw, dtype=np.uint8))
    out = np.zeros_like(bw)
    P = C.POINTER(C.c_ubyte)
    lib().orc_bwselect(bw.ctypes.data_as(P), bw.shape[0], bw.shape[1], c, r, conn,
                       out.ctypes.data_as(P))
    return out.astype(bool)


def findpeak(f):
    f = _F(f)
    xp, yp, mf = C.c_double(), C.c_double(), C.c_double()
    lib().orc_findpeak(_d(f), f.shape[0], f.shape[1], C.byref(xp), C.byref(yp), C.byref(mf))
    return xp.value, yp.value, mf.value


def normxcorr2_kept(Cm, D, r1, r2):
    Cm = _F(Cm)
    D = _F(D)
    out = np.zeros((2 * r1 + 3, 2 * r2 + 3), order="F")
    lib().orc_normxcorr2_kept(_d(Cm), Cm.shape[0], Cm.shape[1], _d(D), r1, r2, _d(out))
    return out


def normxcorr2_full(Cm, D):
    """normxcorr2(C, D): the whole (m1+d1-1) x (m2+d2-1) map, direct FP64."""
    Cm = _F(Cm)
    D = _F(D)
    out = np.zeros((Cm.shape[0] + D.shape[0] - 1, Cm.shape[1] + D.shape[1] - 1), order="F")
    lib().orc_normxcorr2_full(_d(Cm), Cm.shape[0], Cm.shape[1], _d(D), D.shape[0], D.shape[1],
                              _d(out))
    return out


QF_BINS = 30


def qfactor_stats(cc):
    """compute_qFactor (func/funIntegerSearch.m:211-264) of one normxcorr2 map:
    cc_min = cc - min(cc) (cc_max_find, :241-250), pce = max(cc_min)^2 /
    mean(cc_min.^2) (:222), ppe = 1 / the entropy (natural log) of cc_min's
    30-bin histogram (q_entropy, :252-264).  The bins are 30 equal ones over
    [0, max(cc_min)], bin k holding k w <= x < (k+1) w and the last one closed;
    MATLAB's histcounts(x, 30) places its edges at rounded positions by a rule
    of its own that the reference's files do not fix, so ppe's binning is
    parity unpinned against MATLAB (pce is not affected)."""
    x = np.asarray(cc, dtype=np.float64).ravel()
    x = x - x.min()
    peak = x.max()
    n = x.size
    pce = peak * peak / (np.sum(x * x) / n)
    width = peak / QF_BINS
    if width > 0:
        edges = np.arange(QF_BINS + 1, dtype=np.float64) * width
        k = np.clip(np.searchsorted(edges, x, side="right") - 1, 0, QF_BINS - 1)
    else:
        k = np.zeros(n, dtype=np.int64)
    p = np.bincount(k, minlength=QF_BINS) / n
    p = p[p > 0]
    ent = float(np.sum(p * np.log(1.0 / p)))
    ppe = 1.0 / ent if ent > 0 else np.inf
    return pce, ppe


def xcorr_qfactors(f, g, box_f, box_g):
    """[pce ppe] of normxcorr2(f(box_f), g(box_g)) per box (4 ints each,
    1-based inclusive) and ok (0 where normxcorr2 fails: a constant template
    or a box outside the image, the reference's catch)."""
    box_f = np.asarray(box_f, dtype=np.int64).reshape(-1, 4)
    box_g = np.asarray(box_g, dtype=np.int64).reshape(-1, 4)
    q = np.full((box_f.shape[0], 2), np.nan)
    ok = np.zeros(box_f.shape[0], dtype=np.int32)
    H, W = f.shape
    for b, (F, G) in enumerate(zip(box_f, box_g)):
        if min(F[0], F[2], G[0], G[2]) < 1 or max(F[1], G[1]) > H or max(F[3], G[3]) > W:
            continue
        Cm = f[F[0] - 1:F[1], F[2] - 1:F[3]]
        D = g[G[0] - 1:G[1], G[2] - 1:G[3]]
        if Cm.max() == Cm.min() or D.shape[0] < Cm.shape[0] or D.shape[1] < Cm.shape[1]:
            continue
        q[b] = qfactor_stats(normxcorr2_full(Cm, D))
        ok[b] = 1
    return q, ok


def integer_search_mg(f, g, gridx, gridy, winsize, winstepsize, trace=None):
    """funIntegerSearchMg (func/funIntegerSearchMg.m:3-285) with direct FP64
    normxcorr2 maps: returns the scattered box centres and shifts before the
    gridfit (tempx, tempy, tempu, tempv, tempPhi) and the final grid limits
    (gridx, gridy after :256-261).  Small images only (the level-1 map is
    computed by direct sums).  ``trace``: a list that receives one dict per
    level (level 1: the shift of the whole region; levels >= 2: the f and g
    boxes searched and the shifts found, before the border and even-width
    fixes of :221-254)."""
    f = np.asarray(f, dtype=np.float64)
    g = np.asarray(g, dtype=np.float64)
    H, W = f.shape
    gx = [int(gridx[0]), int(gridx[-1])]
    gy = [int(gridy[0]), int(gridy[-1])]
    bgap = 1 + int(np.floor(0.75 * winsize + 0.5))  # :6 (MATLAB round: half away from zero)
    gx[0] = max(gx[0], bgap)
    gy[0] = max(gy[0], bgap)
    gx[1] = min(gx[1], H - bgap)
    gy[1] = min(gy[1], W - bgap)

    def fpk(X):
        xp, yp, mf = findpeak(X)
        return xp, yp, mf

    # level 1 (:17-27)
    Cm = f[gx[0] - 1:gx[1], gy[0] - 1:gy[1]]
    Dm = g[gx[0] - 1:gx[1], gy[0] - 1:gy[1]]
    X = normxcorr2_full(Cm, Dm)
    v1, u1, _ = fpk(X)
    u = u1 - np.ceil((X.shape[0] + 1) / 2.0)
    v = v1 - np.ceil((X.shape[1] + 1) / 2.0)
    if trace is not None:
        trace.append({"xf": [list(map(float, gx))], "yf": [list(map(float, gy))],
                      "u": [u], "v": [v]})

    def clamp_boxes(xf, yf, xg, yg):
        # :30-37 / :221-245: move g boxes inside the border gap, f boxes alike
        for b in range(len(xf)):
            if xg[b][0] < bgap:
                t = xg[b][0]; xg[b][0] = bgap; xf[b][0] = xf[b][0] + bgap - t
            if yg[b][0] < bgap:
                t = yg[b][0]; yg[b][0] = bgap; yf[b][0] = yf[b][0] + bgap - t
            if xg[b][1] > H - bgap:
                t = xg[b][1]; xg[b][1] = H - bgap; xf[b][1] = xf[b][1] + (H - bgap) - t
            if yg[b][1] > W - bgap:
                t = yg[b][1]; yg[b][1] = W - bgap; yf[b][1] = yf[b][1] + (W - bgap) - t
        for b in range(len(xf)):  # :40-41 / :247-254: even widths
            if (xf[b][1] - xf[b][0]) % 2 == 1:
                xf[b][1] -= 1; xg[b][1] -= 1
            if (yf[b][1] - yf[b][0]) % 2 == 1:
                yf[b][1] -= 1; yg[b][1] -= 1

    xf = [list(map(float, gx))]
    yf = [list(map(float, gy))]
    xg = [[gx[0] + np.ceil(u), gx[1] + np.ceil(u)]]
    yg = [[gy[0] + np.ceil(v), gy[1] + np.ceil(v)]]
    clamp_boxes(xf, yf, xg, yg)
    wx = xf[0][1] - xf[0][0]
    wy = yf[0][1] - yf[0][0]
    uc = [np.ceil(u)]
    vc = [np.ceil(v)]
    while True:
        ratio = wx / wy
        nxf, nyf, nxg, nyg, un, vn = [], [], [], [], [], []
        for b in range(len(xf)):
            mx = 0.5 * (xf[b][0] + xf[b][1]); my = 0.5 * (yf[b][0] + yf[b][1])
            mgx = 0.5 * (xg[b][0] + xg[b][1]); mgy = 0.5 * (yg[b][0] + yg[b][1])
            if ratio > 2:
                if xf[b][1] - xf[b][0] > winsize:
                    parts = [([xf[b][0], mx], yf[b], [xg[b][0], mgx], yg[b]),
                             ([mx, xf[b][1]], yf[b], [mgx, xg[b][1]], yg[b])]
                else:
                    parts = [(xf[b], yf[b], xg[b], yg[b])] * 2
            elif ratio < 0.5:
                if yf[b][1] - yf[b][0] > winsize:
                    parts = [(xf[b], [yf[b][0], my], xg[b], [yg[b][0], mgy]),
                             (xf[b], [my, yf[b][1]], xg[b], [mgy, yg[b][1]])]
                else:
                    parts = [(xf[b], yf[b], xg[b], yg[b])] * 2
            else:
                if xf[b][1] - xf[b][0] > winsize and yf[b][1] - yf[b][0] > winsize:
                    lo_x, hi_x = [xf[b][0], mx], [mx, xf[b][1]]
                    lo_gx, hi_gx = [xg[b][0], mgx], [mgx, xg[b][1]]
                    lo_y, hi_y = [yf[b][0], my], [my, yf[b][1]]
                    lo_gy, hi_gy = [yg[b][0], mgy], [mgy, yg[b][1]]
                    parts = [(lo_x, lo_y, lo_gx, lo_gy), (hi_x, lo_y, hi_gx, lo_gy),
                             (lo_x, hi_y, lo_gx, hi_gy), (hi_x, hi_y, hi_gx, hi_gy)]
                else:
                    parts = [(xf[b], yf[b], xg[b], yg[b])] * 4
            for (a1, a2, a3, a4) in parts:
                nxf.append(list(a1)); nyf.append(list(a2)); nxg.append(list(a3)); nyg.append(list(a4))
                un.append(uc[b]); vn.append(vc[b])
        if ratio > 2:
            nwx, nwy = wx / 2, wy
        elif ratio < 0.5:
            nwx, nwy = wx, wy / 2
        else:
            nwx, nwy = wx / 2, wy / 2
        phi = np.zeros(len(nxf))
        for b in range(len(nxf)):  # :151-170
            try:
                vals = nxf[b] + nyf[b] + nxg[b] + nyg[b]
                if not all(np.isfinite(vals)):
                    raise ValueError
                a0, a1_ = int(nxf[b][0]), int(nxf[b][1])
                b0, b1 = int(nyf[b][0]), int(nyf[b][1])
                c0, c1 = int(nxg[b][0]) - 5, int(nxg[b][1]) + 5
                e0, e1 = int(nyg[b][0]) - 5, int(nyg[b][1]) + 5
                if min(a0, b0, c0, e0) < 1 or a1_ > H or c1 > H or b1 > W or e1 > W:
                    raise ValueError
                Cb = f[a0 - 1:a1_, b0 - 1:b1]
                Db = g[c0 - 1:c1, e0 - 1:e1]
                if Cb.shape[0] > Db.shape[0] or Cb.shape[1] > Db.shape[1] or Cb.max() == Cb.min():
                    raise ValueError
                Xb = normxcorr2_full(Cb, Db)
                w1, w2 = Cb.shape[0] - 1, Cb.shape[1] - 1
                crop = Xb[w1 - 1:Xb.shape[0] - w1 + 1, w2 - 1:Xb.shape[1] - w2 + 2]
                zshape = (Xb.shape[0] - 2 * w1 + 2, Xb.shape[1] - 2 * w2 + 2)
                xp, yp, mf = fpk(crop)
                un[b] = un[b] + (yp - np.ceil(zshape[0] / 2.0))
                vn[b] = vn[b] + (xp - np.ceil(zshape[1] / 2.0))
                phi[b] = mf
            except ValueError:
                un[b] = np.nan; vn[b] = np.nan; phi[b] = 0.0
        if trace is not None:
            trace.append({"xf": [list(r) for r in nxf], "yf": [list(r) for r in nyf],
                          "xg": [list(r) for r in nxg], "yg": [list(r) for r in nyg],
                          "u": list(un), "v": list(vn)})
        if nwx < winsize and nwy < winsize:
            break
        xf, yf = nxf, nyf
        xg = [[nxf[b][0] + np.ceil(un[b]), nxf[b][1] + np.ceil(un[b])] for b in range(len(nxf))]
        yg = [[nyf[b][0] + np.ceil(vn[b]), nyf[b][1] + np.ceil(vn[b])] for b in range(len(nyf))]
        xf = [list(r) for r in xf]; yf = [list(r) for r in yf]
        clamp_boxes(xf, yf, xg, yg)
        wx, wy = nwx, nwy
        uc = [np.ceil(x) for x in un]
        vc = [np.ceil(x) for x in vn]
    tx = np.array([0.5 * (a[0] + a[1]) for a in nxf])
    ty = np.array([0.5 * (a[0] + a[1]) for a in nyf])
    tu = np.array(un, dtype=float)
    tv = np.array(vn, dtype=float)
    # :262-270 final grid limits
    ix1, ix2 = int(np.argmin(tx)), int(np.argmax(tx))
    iy1, iy2 = int(np.argmin(ty)), int(np.argmax(ty))
    gap = lambda d: np.ceil(1 + 1.25 * winsize + d)
    if gx[0] < gap(tu[ix1]): gx[0] = gap(tu[ix1])
    if gy[0] < gap(tv[iy1]): gy[0] = gap(tv[iy1])
    if gx[1] > H - gap(tu[ix2]) + 1: gx[1] = H - gap(tu[ix2]) + 1
    if gy[1] > W - gap(tv[iy2]) + 1: gy[1] = W - gap(tv[iy2]) + 1
    return tx, ty, tu, tv, phi, gx, gy


def integer_search(f, g, r, gridx, gridy, winsize, step):
    """funIntegerSearch whole-field mode; returns x0, y0, u, v, ccmax (ny x nx)."""
    f = _F(f)
    g = _F(g)
    H, W = f.shape
    r1, r2 = (r, r) if np.isscalar(r) else r
    gx = (C.c_long * 2)(*gridx)
    gy = (C.c_long * 2)(*gridy)
    mx = (H // step + 2) * (W // step + 2)
    bufs = [np.zeros(mx) for _ in range(5)]
    nx, ny = C.c_long(), C.c_long()
    rc = lib().orc_integer_search(_d(f), _d(g), H, W, r1, r2, gx, gy, winsize, step, mx,
                                  C.byref(nx), C.byref(ny), *[_d(b) for b in bufs])
    if rc != 0:
        raise RuntimeError("integer search grid too large")
    n = nx.value * ny.value
    return [b[:n].reshape((ny.value, nx.value), order="F") for b in bufs]


def local_icgn(fr: Frame, U0, coords, winsize, tol=1e-3, plain=False, nthreads=0):
    """Raw per-subset outputs of LocalICGNQuadtree / LocalICGN before the bad-set
    logic: U (2N), F (4N), H (N x 21), status (N)."""
    coords = _F(coords)
    N = coords.shape[0]
    U0 = np.ascontiguousarray(U0, dtype=np.float64)
    win = np.ascontiguousarray(np.broadcast_to(winsize, (N,)), dtype=np.int64)
    U = np.zeros(2 * N)
    F = np.zeros(4 * N)
    Hp = np.zeros((N, 21), order="F")
    st = np.zeros(N, dtype=np.int32)
    lib().orc_local_icgn_batch(C.byref(fr.st), _d(U0), _d(coords),
                               win.ctypes.data_as(_lp), N, tol, int(plain), nthreads, _d(U),
                               _d(F), _d(Hp), st.ctypes.data_as(_ip))
    return U, F, Hp, st


def subpb1(fr: Frame, UOld, FOld, vdual, coords, wx, wy, mu, tol=1e-3, plain=False,
           nthreads=0):
    coords = _F(coords)
    N = coords.shape[0]
    UOld = np.ascontiguousarray(UOld, dtype=np.float64)
    FOld = np.ascontiguousarray(FOld, dtype=np.float64)
    vdual = np.ascontiguousarray(vdual, dtype=np.float64)
    wx = np.ascontiguousarray(np.broadcast_to(wx, (N,)), dtype=np.int64)
    wy = np.ascontiguousarray(np.broadcast_to(wy, (N,)), dtype=np.int64)
    U = np.zeros(2 * N)
    st = np.zeros(N, dtype=np.int32)
    lib().orc_subpb1_batch(C.byref(fr.st), _d(UOld), _d(FOld), _d(vdual), _d(coords),
                           wx.ctypes.data_as(_lp), wy.ctypes.data_as(_lp), N, mu, tol,
                           int(plain), nthreads, _d(U), st.ctypes.data_as(_ip))
    return U, st


def znssd(fr: Frame, P, coords, winsize, want_grad=True):
    coords = _F(coords)
    N = coords.shape[0]
    P = np.ascontiguousarray(P, dtype=np.float64).reshape(-1)
    win = np.ascontiguousarray(np.broadcast_to(winsize, (N,)), dtype=np.int64)
    cost = np.zeros(N)
    grad = np.zeros(6 * N)
    st = np.zeros(N, dtype=np.int32)
    tot = lib().orc_znssd(C.byref(fr.st), _d(P), _d(coords), win.ctypes.data_as(_lp), N,
                          _d(cost), _d(grad) if want_grad else None, st.ctypes.data_as(_ip))
    return tot, cost, grad.reshape(N, 6), st
