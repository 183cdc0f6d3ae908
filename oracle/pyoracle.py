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
