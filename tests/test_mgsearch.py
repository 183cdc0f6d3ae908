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
