"""Pins for the FP64 oracle (CPU only).

The reference ships no tests or golden vectors for this path, so every oracle
function is checked against something other than its own formula: closed forms
(Keys kernel written as the piecewise cubic W(s), polynomial exactness of the
7-tap stencil), independent library routines (numpy lstsq for findpeak's fit,
np.corrcoef for ZNCC, scipy.ndimage.label for bwselect), known answers
(integer shifts, affine warps), forced status codes, and central finite
differences for the ZNSSD gradient.  A plausible mistake (swapped taps, a
dropped term, transposed coordinates) fails at least one of them.
"""
import numpy as np
import pytest
from scipy import ndimage


# ---------------------------------------------------------------- bicubic
def keys_W(s):
    """Keys (1981) cubic convolution kernel, a = -0.5, piecewise form."""
    s = np.abs(s)
    return np.where(s <= 1, 1.5 * s ** 3 - 2.5 * s ** 2 + 1,
                    np.where(s < 2, -0.5 * s ** 3 + 2.5 * s ** 2 - 4 * s + 2, 0.0))


def test_bicubic_matches_piecewise_kernel(oracle):
    rng = np.random.default_rng(0)
    F = rng.standard_normal((12, 15))
    X = rng.uniform(3, 13, 200)   # column coordinate (dim 2), interior: no clamping
    Y = rng.uniform(3, 10, 200)   # row coordinate (dim 1)
    got = oracle.ba_interp2(F, X, Y)
    rows = np.arange(1, 13)[:, None]
    cols = np.arange(1, 16)[None, :]
    ref = np.array([np.sum(F * keys_W(Y[i] - rows) * keys_W(X[i] - cols)) for i in range(200)])
    np.testing.assert_allclose(got, ref, atol=1e-12)


def test_bicubic_exactness_and_clamp(oracle):
    yy, xx = np.meshgrid(np.arange(1, 21), np.arange(1, 26), indexing="ij")  # rows, cols
    # quadratics are reproduced exactly by the a = -0.5 kernel
    Fq = 0.3 * yy ** 2 - 0.2 * xx * yy + 0.1 * xx ** 2 + yy - 2 * xx + 5
    rng = np.random.default_rng(1)
    X = rng.uniform(3, 22, 100)
    Y = rng.uniform(3, 17, 100)
    q = lambda x, y: 0.3 * y ** 2 - 0.2 * x * y + 0.1 * x ** 2 + y - 2 * x + 5
    np.testing.assert_allclose(oracle.ba_interp2(Fq, X, Y), q(X, Y), atol=1e-10)
    # interpolation at integer nodes returns the node value
    Fr = rng.standard_normal((20, 25))
    Xi = rng.integers(1, 26, 50).astype(float)
    Yi = rng.integers(1, 21, 50).astype(float)
    np.testing.assert_allclose(oracle.ba_interp2(Fr, Xi, Yi), Fr[Yi.astype(int) - 1, Xi.astype(int) - 1],
                               atol=1e-14)
    # beyond the border the taps repeat the edge value (ba_interp2.cpp:62-66)
    Fc = np.tile(np.arange(1, 26, dtype=float), (20, 1))  # varies along columns only
    np.testing.assert_allclose(oracle.ba_interp2(Fc, np.array([-5.0, 40.0]), np.array([3.0, 3.0])),
                               [1.0, 25.0], atol=1e-12)


# --------------------------------------------------------------- gradient
def test_gradient_polynomial_exact(oracle):
    H, W = 30, 34
    x, y = np.meshgrid(np.arange(1, H + 1, dtype=float), np.arange(1, W + 1, dtype=float),
                       indexing="ij")
    img = 1e-3 * x ** 5 - 0.02 * x ** 2 * y + 0.5 * y ** 3
    mask = np.ones((H, W))
    mask[10, 12] = 0
    dx, dy = oracle.img_gradient(img, mask)
    ex = (5e-3 * x ** 4 - 0.04 * x * y)[3:-3, 3:-3]
    ey = (-0.02 * x ** 2 + 1.5 * y ** 2)[3:-3, 3:-3]
    m = mask[3:-3, 3:-3]
    np.testing.assert_allclose(dx, ex * m, rtol=1e-10, atol=1e-8)
    np.testing.assert_allclose(dy, ey * m, rtol=1e-10, atol=1e-8)
    assert dx[10 - 3, 12 - 3] == 0 and dy[10 - 3, 12 - 3] == 0


# --------------------------------------------------------------- bwselect
@pytest.mark.parametrize("conn", [4, 8])
def test_bwselect_against_label(oracle, conn):
    rng = np.random.default_rng(conn)
    struct = ndimage.generate_binary_structure(2, 1 if conn == 4 else 2)
    for _ in range(40):
        bw = rng.uniform(size=(9, 11)) > 0.45
        r, c = rng.integers(1, 10), rng.integers(1, 12)
        got = oracle.bwselect(bw, c, r, conn)
        lab, _ = ndimage.label(bw, structure=struct)
        ref = (lab == lab[r - 1, c - 1]) & bw[r - 1, c - 1]
        np.testing.assert_array_equal(got, ref)


# --------------------------------------------------------------- findpeak
def test_findpeak_fit_against_lstsq(oracle):
    rng = np.random.default_rng(4)
    for _ in range(50):
        A = rng.normal(size=6)
        A[4] = -abs(A[4]) - 1.0
        A[5] = -abs(A[5]) - 1.0
        A[3] *= 0.2
        xs, ys = np.arange(7) - 3.0, np.arange(5) - 2.0
        ygrid, xgrid = np.meshgrid(ys, xs, indexing="ij")  # rows = y, cols = x
        f = (A[0] + A[1] * xgrid + A[2] * ygrid + A[3] * xgrid * ygrid + A[4] * xgrid ** 2
             + A[5] * ygrid ** 2 + 0.01 * rng.normal(size=xgrid.shape))
        xp, yp, mf = oracle.findpeak(f)
        iy, ix = np.unravel_index(np.argmax(f.ravel(order="F")), f.shape, order="F")
        if ix in (0, 6) or iy in (0, 4):
            assert (xp, yp) == (ix + 1, iy + 1)
            continue
        u = f[iy - 1:iy + 2, ix - 1:ix + 2].ravel(order="F")
        x = np.array([-1, -1, -1, 0, 0, 0, 1, 1, 1.0])
        y = np.array([-1, 0, 1, -1, 0, 1, -1, 0, 1.0])
        X = np.stack([np.ones(9), x, y, x * y, x ** 2, y ** 2], 1)
        c = np.linalg.lstsq(X, u, rcond=None)[0]
        d = c[3] ** 2 - 4 * c[4] * c[5]
        xo = (-c[2] * c[3] + 2 * c[5] * c[1]) / d
        yo = -1 / d * (c[3] * c[1] - 2 * c[4] * c[2])
        if abs(xo) > 1 or abs(yo) > 1:
            assert (xp, yp) == (ix + 1, iy + 1)
            continue
        rx = np.sign(1000 * xo) * np.floor(abs(1000 * xo) + 0.5) / 1000
        ry = np.sign(1000 * yo) * np.floor(abs(1000 * yo) + 0.5) / 1000
        assert abs(xp - (ix + 1 + rx)) < 1e-9 and abs(yp - (iy + 1 + ry)) < 1e-9
        ref_max = np.array([1, rx, ry, rx * ry, rx * rx, ry * ry]) @ c
        assert abs(mf - ref_max) < 1e-9


def test_findpeak_ties_and_edges(oracle):
    f = np.zeros((5, 6))
    f[2, 3] = f[1, 4] = 1.0  # equal maxima: the first in column-major order wins
    xp, yp, mf = oracle.findpeak(f)
    assert mf >= 1.0 and round(xp) == 4 and round(yp) == 3
    g = np.zeros((5, 6))
    g[0, 2] = 2.0            # on the edge: integer peak, no fit
    assert oracle.findpeak(g) == (3.0, 1.0, 2.0)


# ------------------------------------------------------ normxcorr2 / search
def test_normxcorr2_against_corrcoef(oracle):
    rng = np.random.default_rng(9)
    m, r = 7, 3
    C = rng.standard_normal((m, m))
    D = rng.standard_normal((m + 2 * r, m + 2 * r))
    got = oracle.normxcorr2_kept(C, D, r, r)
    Dp = np.pad(D, 1)  # normxcorr2 treats D as zero outside
    for a in range(2 * r + 3):
        for b in range(2 * r + 3):
            win = Dp[a:a + m, b:b + m]
            ref = np.corrcoef(C.ravel(), win.ravel())[0, 1]
            assert abs(got[a, b] - ref) < 1e-12


def test_integer_search_known_shift(oracle):
    rng = np.random.default_rng(12)
    H, W = 120, 110
    base = ndimage.gaussian_filter(rng.standard_normal((H + 20, W + 20)), 1.5)
    f = base[10:10 + H, 10:10 + W]
    su, sv = 3, -2  # g(x) = f(x - s): the pattern moved by +s
    g = base[10 - su:10 - su + H, 10 - sv:10 - sv + W]
    x, y, u, v, cc = oracle.integer_search(f, g, 5, [1, H], [1, W], 16, 16)
    assert x.size > 0
    # the integer peak sits at the shift (ZNCC = 1 there); the 3x3 quadratic
    # fit of findpeak then moves it by less than a pixel
    np.testing.assert_array_equal(np.round(u), np.full_like(u, su))
    np.testing.assert_array_equal(np.round(v), np.full_like(v, sv))
    assert np.abs(u - su).max() < 0.2 and np.abs(v - sv).max() < 0.2
    assert np.all(cc > 0.99)
    Cm = f[40:57, 40:57]
    Dm = g[40 - 5:57 + 5, 40 - 5:57 + 5]
    K = oracle.normxcorr2_kept(Cm, Dm, 5, 5)
    ia, ib = np.unravel_index(np.argmax(K), K.shape)
    assert (ia - 1 - 5, ib - 1 - 5) == (su, sv) and abs(K[ia, ib] - 1) < 1e-12


# ------------------------------------------------------------------ IC-GN
def _smooth_pair(seed=21, H=160, W=150, P=(0.01, -0.004, 0.006, -0.008, 1.3, -0.7)):
    """Analytic speckle g(x) = f(W^-1 x) for an affine warp about (80, 75)."""
    rng = np.random.default_rng(seed)
    cx, cy = rng.uniform(-10, H + 10, 900), rng.uniform(-10, W + 10, 900)
    s = rng.uniform(1.8, 3.0, 900)

    def pat(X, Y):
        out = np.zeros_like(X)
        for k in range(900):
            out += np.exp(-((X - cx[k]) ** 2 + (Y - cy[k]) ** 2) / (2 * s[k] ** 2))
        return out
    X, Y = np.meshgrid(np.arange(1, H + 1, dtype=float), np.arange(1, W + 1, dtype=float),
                       indexing="ij")
    f = pat(X, Y)
    A = np.array([[1 + P[0], P[2]], [P[1], 1 + P[3]]])
    x0, y0 = 80.0, 75.0
    Ai = np.linalg.inv(A)
    dx, dy = X - x0 - P[4], Y - y0 - P[5]
    Xr = Ai[0, 0] * dx + Ai[0, 1] * dy + x0
    Yr = Ai[1, 0] * dx + Ai[1, 1] * dy + y0
    g = pat(Xr, Yr)
    n = lambda a: (a - a.mean()) / a.std(ddof=1)
    return n(f), n(g), np.array(P)


def test_icgn6_known_affine(oracle):
    f, g, P = _smooth_pair()
    fr = oracle.Frame(f, g)
    coords = np.array([[80.0, 75.0]])
    for plain in (False, True):
        U, F, Hp, st = oracle.local_icgn(fr, np.array([1.0, -1.0]), coords, 40, tol=1e-6,
                                         plain=plain)
        assert 1 <= st[0] <= 100
        np.testing.assert_allclose(U, P[4:6], atol=5e-3)
        np.testing.assert_allclose(F, P[:4], atol=5e-4)
        # H is the Gauss-Newton Hessian: symmetric positive definite
        Hm = np.zeros((6, 6))
        Hm[np.tril_indices(6)] = Hp[0][[0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16,
                                        17, 18, 19, 20]][np.argsort(np.lexsort(np.tril_indices(6)))]
        Hm = Hm + Hm.T - np.diag(np.diag(Hm))
        assert np.all(np.linalg.eigvalsh(Hm) > 0)


def test_icgn2_limits(oracle):
    f, g, P = _smooth_pair(P=(0.0, 0.0, 0.0, 0.0, 1.4, -0.6))
    fr = oracle.Frame(f, g)
    coords = np.array([[80.0, 75.0]])
    UOld = np.array([1.0, -1.0])
    FOld = np.zeros(4)
    vdual = np.array([0.25, 0.1])
    # mu = 0: plain translation IC-GN recovers the shift
    U, st = oracle.subpb1(fr, UOld, FOld, vdual, coords, 40, 40, 0.0, tol=1e-8)
    np.testing.assert_allclose(U, P[4:6], atol=5e-3)
    # mu large: funICGN_Subpb1.m:379,391-394 adds +mu (P - UOld - vdual) to b
    # and the inverse-compositional update then subtracts the step, so the
    # penalty drives u away from UOld + vdual (a sign quirk of the reference,
    # kept for parity): the iterate leaves in the direction of -vdual and the
    # subset does not converge
    U, st = oracle.subpb1(fr, UOld, FOld, vdual, coords, 40, 40, 1e9, tol=1e-12)
    assert st[0] == 101
    assert np.all(np.sign(U - UOld) == -np.sign(vdual))


@pytest.mark.parametrize("mu", [1e-3, 1e-2, 3e-2])
def test_icgn2_penalty_weight_at_the_fixed_point(oracle, mu):
    """funICGN_Subpb1.m:111,379 stops where tempb = 2 b / bottomf^2 +
    mu (P - UOld - vdual) vanishes.  b is taken here from the ZNSSD cost alone,
    by central differences of orc_znssd (itself pinned by finite differences
    above) through the first-order identity b = -(bottomf bottomg / 2) dC/dP
    (the steepest-descent vector of funICGNQuadtree.m:362-371 against the
    derivative of C = sum (fhat - ghat)^2), with bottomf and bottomg from
    the window's pixels.  Dropping the 2/bottomf^2 weight (from H2 and tempb
    alike) or the factor 2 moves the fixed point by far more than the 10 %
    allowed; a penalty of the other sign fails the direction test."""
    f, g, _ = _smooth_pair(P=(0.0, 0.0, 0.0, 0.0, 1.4, -0.6))
    fr = oracle.Frame(f, g)
    x0, y0, w = 80.0, 75.0, 40
    coords = np.array([[x0, y0]])
    UOld, FOld, vdual = np.array([1.0, -1.0]), np.zeros(4), np.array([0.25, 0.1])
    target = UOld + vdual
    U0, st0 = oracle.subpb1(fr, UOld, FOld, vdual, coords, w, w, 0.0, tol=1e-12)
    U, st = oracle.subpb1(fr, UOld, FOld, vdual, coords, w, w, mu, tol=1e-12)
    assert 1 <= st0[0] <= 100 and 1 <= st[0] <= 100
    # the window's pixels: f on the 41 x 41 window, g bicubic at the warped
    # positions (the oracle's ba_interp2 is pinned by the Keys kernel above)
    XX, YY = np.meshgrid(np.arange(x0 - 20, x0 + 21), np.arange(y0 - 20, y0 + 21), indexing="ij")
    fw = f[int(x0) - 21:int(x0) + 20, int(y0) - 21:int(y0) + 20]
    gw = oracle.ba_interp2(g, (YY + U[1]).ravel(), (XX + U[0]).ravel())
    bf = np.sqrt(np.sum((fw - fw.mean()) ** 2))
    bg = np.sqrt(np.sum((gw - gw.mean()) ** 2))
    h = 1e-4
    b = np.zeros(2)
    for k in range(2):
        Pp = np.zeros(6)
        Pp[4:6] = U
        Pm = Pp.copy()
        Pp[4 + k] += h
        Pm[4 + k] -= h
        cp = oracle.znssd(fr, Pp[None], coords, w, want_grad=False)[1][0]
        cm = oracle.znssd(fr, Pm[None], coords, w, want_grad=False)[1][0]
        b[k] = -(bf * bg / 2.0) * (cp - cm) / (2 * h)
    pen = mu * (U - target)
    assert np.all(np.abs(2 * b / bf ** 2 + pen) <= 0.1 * np.abs(pen))
    # the penalty moves the solution away from UOld + vdual (the sign of :379
    # under the inverse-compositional update)
    assert np.all(np.sign(U - U0) == np.sign(U0 - target))


def _two_motion_pair(transposed):
    """A window split in two by a masked band, each side with its own integer
    shift (g sampled exactly from the same analytic speckle pattern, so
    bicubic samples at integer positions reproduce it bit for bit).  Window
    (wx, wy) = (40, 20) at (80, 80), or (20, 40) with the band across y."""
    H = W = 160
    x0 = y0 = 80
    rng = np.random.default_rng(7)
    cx, cy = rng.uniform(-10, H + 10, 1000), rng.uniform(-10, W + 10, 1000)
    s = rng.uniform(1.8, 3.0, 1000)

    def pat(X, Y):
        out = np.zeros_like(X)
        for k in range(1000):
            out += np.exp(-((X - cx[k]) ** 2 + (Y - cy[k]) ** 2) / (2 * s[k] ** 2))
        return out

    X, Y = np.meshgrid(np.arange(1, H + 1, dtype=float), np.arange(1, W + 1, dtype=float),
                       indexing="ij")
    if not transposed:
        sA, sB, coord, lo, hi, k, wx, wy = [2.0, -1.0], [-1.0, 1.0], X, x0 - 10, x0 - 3, 0, 40, 20
    else:
        sA, sB, coord, lo, hi, k, wx, wy = [-1.0, 2.0], [1.0, -1.0], Y, y0 - 9, y0 - 2, 1, 20, 40
    sA, sB = np.array(sA), np.array(sB)
    mask = np.where((coord >= lo) & (coord <= hi), 0.0, 1.0)
    f = pat(X, Y) * mask
    g = np.where(coord < lo + sA[k], pat(X - sA[0], Y - sA[1]),
                 np.where(coord > hi + sB[k], pat(X - sB[0], Y - sB[1]), 0.0))
    return f, g, mask, sA, sB, (x0, y0, wx, wy)


@pytest.mark.parametrize("transposed", [False, True])
def test_icgn2_masked_region_seed(oracle, transposed):
    """The 2-DOF masked path (funICGN_Subpb1.m:58-93,150-190): the window is
    grown for its masked pixels (:74-88) and every iteration keeps the
    4-connected component of nonzero g through bwselect(BW, c, r) with
    c = floor((wx+1)/2) a COLUMN (dim-2) index and r = floor((wy+1)/2) a ROW
    index; f, its gradient, H and the f statistics are rebuilt on that region.
    The band splits the window into side A, holding the seed, and side B,
    holding the centre and the swapped seed (row c, column r); the sides move
    by sA and sB, 3.6 px apart.  Started exactly at sA, the residual on A is
    zero, so the first norm test stops with U = sA (a region taken from B, or
    f statistics taken over the whole window, leave a nonzero b); started
    0.3 px away it converges next to sA (the rows next to the band hold
    bicubic samples that straddle it, which moves the fixed point by a few
    hundredths of a pixel), far from sB."""
    f, g, mask, sA, sB, (x0, y0, wx, wy) = _two_motion_pair(transposed)
    fr = oracle.Frame(f, g, mask)
    coords = np.array([[x0, y0]], dtype=float)
    # construction check: with the window grown as :74-79 grows it (47 x 23,
    # or 25 x 45; 168 masked pixels), the component of nonzero g at sA through
    # the reference's seed is side A only
    gwx, gwy = (46, 22) if not transposed else (24, 44)
    xa, ya = x0 - gwx // 2, y0 - gwy // 2
    XX, YY = np.meshgrid(np.arange(xa, xa + gwx + 1, dtype=float),
                         np.arange(ya, ya + gwy + 1, dtype=float), indexing="ij")
    gs = oracle.ba_interp2(g, (YY + sA[1]).ravel(), (XX + sA[0]).ravel()).reshape(XX.shape)
    lab, _ = ndimage.label(gs != 0)
    r, c = (gwy + 1) // 2, (gwx + 1) // 2  # bwselect(BW, c, r): row r, column c (1-based)
    region = lab == lab[r - 1, c - 1]
    assert lab[c - 1, r - 1] != lab[r - 1, c - 1]  # the swapped seed lies on side B
    fw = f[xa - 1:xa + gwx, ya - 1:ya + gwy]
    assert np.array_equal(fw[region], gs[region]) and region.sum() > 250
    # the exact start stops at the first test with U = sA
    U, st = oracle.subpb1(fr, sA, np.zeros(4), np.zeros(2), coords, wx, wy, 0.0, tol=1e-6)
    assert st[0] == 1 and np.array_equal(U, sA)
    for start in (sA + [0.3, -0.2], sA + [-0.3, 0.25]):
        U, st = oracle.subpb1(fr, start, np.zeros(4), np.zeros(2), coords, wx, wy, 0.0, tol=1e-6)
        assert 1 <= st[0] <= 100
        assert np.abs(U - sA).max() < 0.2 and np.hypot(*(U - sB)) > 3.0


def test_icgn_status_codes(oracle):
    rng = np.random.default_rng(11)
    H = W = 96
    f = rng.standard_normal((H, W))
    mask = np.ones((H, W))
    mask[60:, 60:] = 0
    f = f * mask
    fr = oracle.Frame(f, f.copy(), mask)
    coords = np.array([[48, 48], [48, 48], [75, 75], [48, 48], [5, 48]], dtype=float)
    U0 = np.array([0, 0, 60, 0, 0, 0, np.nan, 0, 0, 0.0])
    U, F, Hp, st = oracle.local_icgn(fr, U0, coords, 16, tol=1e-3)
    assert list(st) == [1, 101, 102, -1, -1]
    assert np.all(np.isnan(U[6:8])) and np.all(Hp[2] == 0)
    # the plain twin: NaN initial guess gives status 1 with NaN output
    Up, Fp, Hpp, stp = oracle.local_icgn(fr, U0, coords, 16, tol=1e-3, plain=True)
    assert stp[3] == 1 and np.isnan(Up[6]) and Up[7] == 0.0


# ------------------------------------------------------------------ ZNSSD
def test_znssd_gradient_fd_and_steepest_descent(oracle):
    f, g, P = _smooth_pair()
    fr = oracle.Frame(f, g)
    coords = np.array([[80.0, 75.0], [70.0, 80.0]])
    rng = np.random.default_rng(3)
    Pb = np.tile(P, (2, 1)) + rng.normal(0, [2e-3] * 4 + [0.1] * 2, (2, 6))
    tot, c, gr, st = oracle.znssd(fr, Pb, coords, 30)
    assert np.all(st == 0)
    h = np.array([1e-5] * 4 + [1e-3] * 2)
    for i in range(2):
        for k in range(6):
            Pp, Pm = Pb.copy(), Pb.copy()
            Pp[i, k] += h[k]
            Pm[i, k] -= h[k]
            fd = (oracle.znssd(fr, Pp, coords, 30)[0] - oracle.znssd(fr, Pm, coords, 30)[0]) / (2 * h[k])
            assert abs(fd - gr[i, k]) <= 2e-3 * np.abs(gr[i]).max() + 1e-8
    # near the solution the FA gradient matches the IC steepest-descent vector
    # -2 b / (bottomf bottomg) up to the warp Jacobian (small strain)
    U, F, Hp, stt = oracle.local_icgn(fr, np.array([1.0, -1.0]), coords[:1], 30, tol=1e-12)
    Pc = np.concatenate([F, U])[None, :] + np.array([[0, 0, 0, 0, 0.05, -0.04]])
    _, _, gc, _ = oracle.znssd(fr, Pc, coords[:1], 30)
    assert gc[0, 4] > 0 and gc[0, 5] < 0  # descent points back to the optimum


def test_oracle_on_reference_frames(oracle, ref_frames):
    """The reference's own data: frames 000/001 of its circular-hole sequence,
    its 1434-node mesh (winsize 32) and its stored answer for the pair
    (ResultDisp{1}.U, after three AL steps).  One 6-DOF local step started from
    the rounded stored answer must land on it for the subsets that converge
    (the AL steps move converged subsets by hundredths of a pixel at most),
    and the count of not-converged subsets must match the stored bad-subset
    count of the first local step (ALSub1BadPtNum(1) = 22).  Masked windows
    whose size changes at the second iteration end with -1 here; the stored run
    predates the current funICGNQuadtree.m, which no longer rebuilds f on the
    grown window (:180-205 commented out), so they are excluded."""
    r = ref_frames
    N = r["coords"].shape[0]
    U1 = r["U1"]
    of = oracle.Frame(r["f"], r["g"], r["mask"])
    win = np.full(N, r["winsize"])
    U, F, H, st = oracle.local_icgn(of, np.round(U1), r["coords"], win, tol=1e-2)
    n101 = int((st == 101).sum())
    assert n101 == 22
    conv = (st >= 1) & (st <= 100)
    assert conv.sum() > 650
    d = np.hypot(U[0::2] - U1[0::2], U[1::2] - U1[1::2])[conv]
    assert np.median(d) < 3e-3
    assert np.percentile(d, 90) < 0.06
    assert d.max() < 0.25
    dF = np.abs(F.reshape(-1, 4) - r["F1"].reshape(-1, 4))[conv]
    assert np.all(np.median(dF, axis=0) < 2e-4)
    # statuses other than iteration counts are the reference's codes only
    assert set(np.unique(st[~conv]).tolist()) <= {-1, 101, 102, 103}
