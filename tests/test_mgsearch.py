"""Multigrid integer search (funIntegerSearchMg) and gridfit.

CPU: the oracle's whole-map normxcorr2 against its kept-block restatement
(itself pinned by numpy.corrcoef in test_oracle.py) and known integer shifts;
the oracle's level loop on a known translation; gridfit's invariants
(constants reproduced exactly, planes as the smoothing weight vanishes, the
node-extension rule).  GPU: the two correlation entry points against the
oracle map by map and box by box, and the whole search against the oracle's
level loop followed by the same gridfit.
"""
import numpy as np
import pytest


def _img(n1, n2, seed):
    from staq_dic_amd import synth
    im = synth.speckle_image(n1, n2, int(n1 * n2 * 0.012), seed)
    return synth._f32(synth.normalize(im))


def test_full_map_matches_kept_block(oracle):
    rng = np.random.default_rng(3)
    for m, r in [(9, 3), (12, 4)]:
        Cm = rng.normal(size=(m, m))
        D = rng.normal(size=(m + 2 * r, m + 2 * r))
        full = oracle.normxcorr2_full(Cm, D)
        assert full.shape == (2 * m + 2 * r - 1,) * 2
        kept = oracle.normxcorr2_kept(Cm, D, r, r)
        # kept placement s in [-1, 2r+1] is full-map index s + m - 1
        np.testing.assert_allclose(full[m - 2:m + 2 * r + 1, m - 2:m + 2 * r + 1], kept,
                                   rtol=1e-12, atol=1e-12)


def test_full_map_known_shift(oracle):
    """g = f moved by (du, dv); the template f(16:35, 14:35) lies whole inside
    D = g(11:40, 9:40) at its true placement, where the map peaks at 1; map
    entry (k, l) puts the template's top-left at D(k - m1 + 1, l - m2 + 1)."""
    f = _img(60, 50, 11)
    Cm = f[15:35, 13:35]
    for du, dv in [(3, -2), (-5, 4), (0, 0)]:
        g = np.roll(np.roll(f, du, axis=0), dv, axis=1)
        D = g[10:40, 8:40]
        X = oracle.normxcorr2_full(Cm, D)
        i, j = np.unravel_index(np.argmax(X), X.shape)
        assert (i - (Cm.shape[0] - 1), j - (Cm.shape[1] - 1)) == (5 + du, 5 + dv)
        assert abs(X[i, j] - 1.0) < 1e-12


def test_oracle_levels_recover_a_translation(oracle):
    f = _img(200, 180, 21)
    g = np.roll(np.roll(f, 3, axis=0), -2, axis=1)
    tx, ty, tu, tv, phi, gx, gy = oracle.integer_search_mg(f, g, [1, 200], [1, 180], 24, 12)
    ok = np.isfinite(tu)
    assert ok.mean() > 0.9
    # g(x) = f(x - t): the displacement of every box is the roll, up to
    # findpeak's sub-pixel fit of the correlation peak
    assert np.abs(tu[ok] - 3).max() < 0.5 and np.abs(tv[ok] + 2).max() < 0.5
    assert abs(np.median(tu[ok]) - 3) < 0.02 and abs(np.median(tv[ok]) + 2) < 0.02
    assert np.median(phi[ok]) > 0.99


def _piecewise_pair():
    """256 x 256 pair sampled from one analytic speckle pattern (rounded to
    FP32, as frames are held): moved by s1 in the quadrant x <= 128, y >= 126
    of g and by s0 elsewhere."""
    H = W = 256
    rng = np.random.default_rng(17)
    n = 800
    cx, cy = rng.uniform(-10, H + 10, n), rng.uniform(-10, W + 10, n)
    s = rng.uniform(1.8, 3.0, n)

    def pat(X, Y):
        out = np.zeros_like(X)
        for k in range(n):
            out += np.exp(-((X - cx[k]) ** 2 + (Y - cy[k]) ** 2) / (2 * s[k] ** 2))
        return out

    X, Y = np.meshgrid(np.arange(1, H + 1, dtype=float), np.arange(1, W + 1, dtype=float),
                       indexing="ij")
    s0 = np.array([-0.35, -0.4])
    s1 = np.array([2.4, -3.4])
    f = pat(X, Y)
    g = np.where((X <= 128) & (Y >= 126), pat(X - s1[0], Y - s1[1]), pat(X - s0[0], Y - s0[1]))
    r32 = lambda a: a.astype(np.float32).astype(np.float64)
    return r32(f), r32(g), s0, s1


def test_oracle_levels_on_a_piecewise_field(oracle):
    """The level bookkeeping of funIntegerSearchMg.m:17-254 on a field with
    two motions: g is sampled exactly from one analytic speckle pattern moved
    by s1 = (2.4, -3.4) in the quadrant x <= 128, y >= 126 of g and by
    s0 = (-0.35, -0.4) elsewhere (256 x 256, winsize 32).  The boxes follow
    from the rules by hand: border gap 1 + round(24) = 25; level 1 finds
    u in (0, 1), v in (-1, 0), so its g box moves by ceil = (1, 0) and is
    clamped at 256 - 25, which shrinks the f box to x in [25, 230], trimmed to
    the even width [25, 229] (:30-41); levels 2-4 split every box in four
    (ratio 1), keep the halves' shared midpoints, and trim odd widths after
    levels 2 and 3 (:221-254); the s1 quadrant's g boxes move by
    ceil(u) = 3, ceil(v) = -3 from level 2 on.  The last level holds
    4^3 = 64 boxes whose centres are the products of the eight values below,
    in the (lo x, lo y), (hi x, lo y), (lo x, hi y), (hi x, hi y) order of
    :131-145 at every level, and each box away from the quadrant boundary
    carries its own quadrant's shift (ceil'd shifts accumulated level by
    level, :164-168, :240).  The grid limits of :256-261 take the shifts of
    the first box at the smallest and largest centres."""
    f, g, s0, s1 = _piecewise_pair()
    H, W = f.shape
    tr = []
    tx, ty, tu, tv, phi, gx, gy = oracle.integer_search_mg(f, g, [1, H], [1, W], 32, 16, trace=tr)
    assert len(tr) == 4 and [len(t["xf"]) for t in tr] == [1, 4, 16, 64]
    # level 1: the whole region [25, 231]^2
    assert tr[0]["xf"] == [[25.0, 231.0]] and tr[0]["yf"] == [[25.0, 231.0]]
    assert 0 < tr[0]["u"][0] < 1 and -1 < tr[0]["v"][0] < 0
    # level 2: the four quadrants of the clamped, trimmed region
    l2 = tr[1]
    assert [list(map(float, b)) for b in l2["xf"]] == [[25, 127], [127, 229], [25, 127], [127, 229]]
    assert [list(map(float, b)) for b in l2["yf"]] == [[25, 128], [25, 128], [128, 231], [128, 231]]
    assert [list(map(float, b)) for b in l2["xg"]] == [[26, 128], [128, 230], [26, 128], [128, 230]]
    assert [list(map(float, b)) for b in l2["yg"]] == [[25, 128], [25, 128], [128, 231], [128, 231]]
    want2 = np.array([s0, s0, s1, s0])
    got2 = np.stack([l2["u"], l2["v"]], axis=1)
    assert np.abs(got2 - want2).max() < 0.2
    # last level: centres and order
    xs = [37.5, 62.5, 88.5, 113.5, 139.5, 164.5, 190.5, 215.5]
    ys = [37.5, 62.5, 88.5, 113.5, 140.5, 165.5, 191.5, 216.5]
    exp_x, exp_y = [], []
    for b2 in range(4):  # level-2 box: (lo x, lo y), (hi x, lo y), (lo x, hi y), (hi x, hi y)
        for b3 in range(4):
            for b4 in range(4):
                ix = 4 * (b2 % 2) + 2 * (b3 % 2) + (b4 % 2)
                iy = 4 * (b2 // 2) + 2 * (b3 // 2) + (b4 // 2)
                exp_x.append(xs[ix])
                exp_y.append(ys[iy])
    np.testing.assert_array_equal(tx, exp_x)
    np.testing.assert_array_equal(ty, exp_y)
    # shifts of the boxes clear of the quadrant boundary
    in1 = (tx < 113) & (ty > 141)
    clear = ~np.isin(tx, [113.5, 139.5]) & ~np.isin(ty, [113.5, 140.5])
    want = np.where(in1[:, None], s1, s0)
    err = np.abs(np.stack([tu, tv], 1) - want)[clear]
    assert clear.sum() == 36 and err.max() < 0.2
    for sel, sh in ((in1 & clear, s1), (~in1 & clear, s0)):
        assert np.abs(np.median(tu[sel]) - sh[0]) < 0.05 and np.abs(np.median(tv[sel]) - sh[1]) < 0.05
    assert np.all(phi[clear] > 0.98)
    # grid limits: ceil(1 + 1.25 * 32 + shift) of the first box at each extreme
    # centre; the first box at the largest y centre lies in the s1 quadrant
    assert gx == [41.0, 216.0] and gy == [41.0, 219.0]


def test_gridfit_constant_and_plane():
    from staq_dic_amd import gridfit
    rng = np.random.default_rng(5)
    x = rng.uniform(10, 90, 300)
    y = rng.uniform(5, 75, 300)
    xn = np.arange(10, 91, 8.0)
    yn = np.arange(5, 76, 7.0)
    z, xg, yg = gridfit(x, y, np.full(300, 2.5), xn, yn)
    assert z.shape == (yn.size, xn.size) and xg.shape == z.shape
    np.testing.assert_allclose(z, 2.5, rtol=0, atol=1e-10)
    plane = lambda a, b: 0.3 + 0.02 * a - 0.01 * b
    z, xg, yg = gridfit(x, y, plane(x, y), xn, yn, smoothness=1e-7)
    np.testing.assert_allclose(z, plane(xg, yg), atol=1e-5)
    # the springs pull the surface towards a constant (gridfit.m:135-137):
    # smoothness 1 flattens the plane, and never past the range of the data
    zd = plane(x, y)
    z1, _, _ = gridfit(x, y, zd, xn, yn)
    assert zd.min() - 1e-9 <= z1.min() and z1.max() <= zd.max() + 1e-9
    assert np.ptp(z1) < np.ptp(plane(xg, yg))
    # data beyond the last node move that node out to them (extend = 'warning')
    z2, xg2, _ = gridfit(np.append(x, 95.0), np.append(y, 40.0), np.full(301, 1.0), xn, yn)
    assert xg2[0, -1] == 95.0 and np.allclose(z2, 1.0)


@pytest.mark.gpu
def test_full_peak_against_oracle(ctx, oracle):
    from staq_dic_amd import mgsearch
    from staq_dic_amd import synth
    field = synth.AffineField(4.2, -2.7, 0.01, -0.005, 0.004, 0.008)
    fr = synth.make_frame(180, 160, 350, 31, field)
    f, g = fr.f, fr.g
    ctx.set_frame(f, g, None, slot=0)
    for bf, bg in [([20, 150, 14, 141], [20, 150, 14, 141]), ([31, 90, 40, 101], [25, 110, 30, 120])]:
        out = mgsearch._full_peak(ctx, 0, bf, bg)
        X = oracle.normxcorr2_full(f[bf[0] - 1:bf[1], bf[2] - 1:bf[3]], g[bg[0] - 1:bg[1], bg[2] - 1:bg[3]])
        xp, yp, mf = oracle.findpeak(X)
        assert (out[0], out[1]) == (xp, yp), (out, (xp, yp))
        assert abs(out[2] - mf) < 1e-9


@pytest.mark.gpu
def test_boxes_against_oracle(ctx, oracle):
    from staq_dic_amd import mgsearch
    f = _img(170, 150, 41)
    g = np.roll(np.roll(f, 2, axis=0), 1, axis=1)
    g[60:64, 70:90] = 0.0  # a flat patch inside some boxes
    ctx.set_frame(f, g, None, slot=0)
    rng = np.random.default_rng(8)
    bf, bg = [], []
    for _ in range(60):
        n1, n2 = rng.integers(6, 90), rng.integers(6, 90)
        x0, y0 = rng.integers(1, 170 - n1), rng.integers(1, 150 - n2)
        sx, sy = rng.integers(-3, 4, 2)
        bf.append([x0, x0 + n1 - 1, y0, y0 + n2 - 1])
        bg.append([x0 + sx, x0 + sx + n1 - 1, y0 + sy, y0 + sy + n2 - 1])
    bf.append([5, 5, 10, 40])  # one-row template: the reference's catch
    bg.append([5, 5, 10, 40])
    bf = np.array(bf)
    bg = np.array(bg)
    out, ok = mgsearch._boxes(ctx, 0, bf, bg)
    n_ok = 0
    for b in range(len(bf)):
        F, G = bf[b], bg[b]
        p = mgsearch.PAD
        inside = G[0] - p >= 1 and G[2] - p >= 1 and G[1] + p <= 170 and G[3] + p <= 150
        Cm = f[F[0] - 1:F[1], F[2] - 1:F[3]]
        if not inside or min(Cm.shape) < 2 or Cm.max() == Cm.min():
            assert ok[b] == 0 and np.isnan(out[b]).all()
            continue
        X = oracle.normxcorr2_full(Cm, g[G[0] - p - 1:G[1] + p, G[2] - p - 1:G[3] + p])
        w1, w2 = Cm.shape[0] - 1, Cm.shape[1] - 1
        crop = X[w1 - 1:X.shape[0] - w1 + 1, w2 - 1:X.shape[1] - w2 + 2]
        assert crop.shape == (2 * p + 3, 2 * p + 4)
        xp, yp, mf = oracle.findpeak(crop)
        assert ok[b] == 1
        assert (out[b, 0], out[b, 1]) == (xp, yp), (b, out[b], xp, yp)
        assert abs(out[b, 2] - mf) < 1e-9
        n_ok += 1
    assert n_ok > 30


@pytest.mark.gpu
def test_mg_search_against_oracle(ctx, oracle):
    from staq_dic_amd import gridfit, mgsearch, synth
    from staq_dic_amd.api import IntegerSearch
    field = synth.AffineField(2.6, -1.4, 0.012, -0.006, 0.005, 0.009)
    fr = synth.make_frame(200, 176, 420, 51, field)
    H, W = fr.f.shape
    ctx.set_frame(fr.f, fr.g, None, slot=0)
    tx, ty, tu, tv, phi, gx, gy = mgsearch.mg_scatter(ctx, 0, H, W, [1, H], [1, W], 24)
    ox, oy, ou, ov, ophi, ogx, ogy = oracle.integer_search_mg(fr.f, fr.g, [1, H], [1, W], 24, 12)
    np.testing.assert_array_equal(tx, ox)
    np.testing.assert_array_equal(ty, oy)
    np.testing.assert_array_equal(tu, ou)
    np.testing.assert_array_equal(tv, ov)
    np.testing.assert_allclose(phi, ophi, rtol=0, atol=1e-9)
    assert list(gx) == list(ogx) and list(gy) == list(ogy)
    # the whole search: the same gridfit of the same scatter
    P, x0, y0, u, v, cc = IntegerSearch(fr.f, fr.g, None, {
        "winsize": 24, "winstepsize": 12, "InitFFTSearchMethod": 0,
        "gridxyROIRange": {"gridx": [1, H], "gridy": [1, W]}}, ctx=ctx)
    xl = np.arange(gx[0], gx[1] + 1e-9, 12)
    yl = np.arange(gy[0], gy[1] + 1e-9, 12)
    sel = (ox > xl.min()) & (ox < xl.max()) & (oy > yl.min()) & (oy < yl.max())
    uo, xo, yo = gridfit(ox[sel], oy[sel], ou[sel], xl, yl)
    np.testing.assert_allclose(u, uo, rtol=0, atol=1e-9)
    np.testing.assert_array_equal(x0, xo)
    # integer-pixel initial guess: within 1 px of the analytic field
    tu_, tv_ = field(x0, y0)
    assert np.mean(np.abs(u - tu_) <= 1.0) > 0.9 and np.mean(np.abs(v - tv_) <= 1.0) > 0.9
    assert P["InitFFTSearchMethod"] == 0


@pytest.mark.gpu
def test_mg_scatter_on_a_piecewise_field(ctx, oracle):
    """The device multigrid levels against the oracle's on the two-motion
    field whose boxes and shifts the CPU test above fixes by hand."""
    from staq_dic_amd import mgsearch
    f, g, _, _ = _piecewise_pair()
    H, W = f.shape
    ctx.set_frame(f, g, None, slot=0)
    tx, ty, tu, tv, phi, gx, gy = mgsearch.mg_scatter(ctx, 0, H, W, [1, H], [1, W], 32)
    ox, oy, ou, ov, ophi, ogx, ogy = oracle.integer_search_mg(f, g, [1, H], [1, W], 32, 16)
    np.testing.assert_array_equal(tx, ox)
    np.testing.assert_array_equal(ty, oy)
    np.testing.assert_array_equal(tu, ou)
    np.testing.assert_array_equal(tv, ov)
    np.testing.assert_allclose(phi, ophi, rtol=0, atol=1e-9)
    assert list(gx) == list(ogx) and list(gy) == list(ogy)
