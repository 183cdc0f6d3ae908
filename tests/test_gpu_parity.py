"""GPU parity: libstaqdic.so against the FP64 oracle on the same inputs.

Tolerances: integer-search offsets must be identical; bicubic FP64 gathers
agree to 1e-12; FP32 paths to FP32 rounding; every IC-GN subset converged on
both sides, whatever its iteration count, within 1e-3 px in displacement and
1e-5 in displacement gradient. The stop test that sits within FP32 rounding
of the tolerance is decided by an FP64 re-solve, so the reference's codes and
the iteration counts agree except for the rare subset the FP64 re-solve and
the oracle still split (at most 0.2 %).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


def test_ba_interp2_f64_f32(ctx, oracle):
    from staq_dic_amd import ba_interp2
    rng = np.random.default_rng(7)
    F = rng.standard_normal((40, 53))
    X = rng.uniform(-3, 57, (31, 17))  # columns, including out-of-range (clamped)
    Y = rng.uniform(-3, 44, (31, 17))
    X[0, :4] = [1, 2, 53, 30]          # exact nodes and the last column
    Y[0, :4] = [1, 40, 7, 20]
    ref = oracle.ba_interp2(F, X, Y)
    got = ba_interp2(F, X, Y, ctx=ctx)
    assert got.dtype == np.float64
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12)
    got32 = ba_interp2(F.astype(np.float32), X.astype(np.float32), Y.astype(np.float32), ctx=ctx)
    assert got32.dtype == np.float32
    ref32 = oracle.ba_interp2(F.astype(np.float32).astype(float), X.astype(np.float32).astype(float),
                              Y.astype(np.float32).astype(float))
    np.testing.assert_allclose(got32, ref32, rtol=0, atol=2e-5)
    with pytest.raises(TypeError):
        ba_interp2(F, X.astype(np.float32), Y, ctx=ctx)


def test_ba_interp2_channels_nan(ctx, oracle):
    from staq_dic_amd import ba_interp2
    rng = np.random.default_rng(8)
    F = rng.standard_normal((20, 30, 3))
    X = rng.uniform(1, 30, 50)
    Y = rng.uniform(1, 20, 50)
    X[3] = np.nan
    got = ba_interp2(F, X, Y, ctx=ctx)
    assert got.shape == (50, 3)
    for p in range(3):
        ref = oracle.ba_interp2(F[:, :, p], X, Y)
        np.testing.assert_allclose(got[:, p], ref, rtol=0, atol=1e-12)
    assert np.all(np.isnan(got[3]))


def test_gradient(ctx, oracle):
    from staq_dic_amd import funImgGradient
    rng = np.random.default_rng(3)
    img = rng.standard_normal((70, 90)).astype(np.float32).astype(np.float64)
    mask = (rng.uniform(size=img.shape) > 0.1).astype(float)
    Df = funImgGradient(img, img, mask, ctx=ctx)
    dx, dy = oracle.img_gradient(img, mask)
    np.testing.assert_allclose(Df["DfDx"], dx, rtol=0, atol=2e-6)
    np.testing.assert_allclose(Df["DfDy"], dy, rtol=0, atol=2e-6)


def test_integer_search_plain(ctx, oracle, plain_cfg):
    from staq_dic_amd import funIntegerSearch
    fr, _, _ = plain_cfg
    H, W = fr.f.shape
    _print("integer search 512^2, r=10, w=32, step 16")
    x, y, u, v, cc = funIntegerSearch(fr.f, fr.g, 10, [1, H], [1, W], 32, 16, ctx=ctx)
    xo, yo, uo, vo, cco = oracle.integer_search(fr.f, fr.g, 10, [1, H], [1, W], 32, 16)
    assert x.shape == xo.shape == (26, 26)
    np.testing.assert_array_equal(x, xo)
    np.testing.assert_array_equal(y, yo)
    np.testing.assert_array_equal(np.floor(u), np.floor(uo))  # integer part
    np.testing.assert_array_equal(np.floor(v), np.floor(vo))
    assert np.mean(u == uo) > 0.99 and np.mean(v == vo) > 0.99  # 1e-3 rounded subpixel
    np.testing.assert_allclose(u, uo, atol=1e-3 + 1e-12)
    np.testing.assert_allclose(cc["max"], cco, atol=1e-9)
    # the field is u = 3.4 + 0.02x + 0.01y, v = -1.7 + 0.005x - 0.01y; where it
    # stays inside the search radius the peak is the true shift
    tu = 3.4 + 0.02 * x + 0.01 * y
    tv = -1.7 + 0.005 * x - 0.01 * y
    inside = (np.abs(tu) < 9) & (np.abs(tv) < 9)
    assert inside.sum() > 100
    assert np.max(np.abs(u - tu)[inside]) < 0.6 and np.max(np.abs(v - tv)[inside]) < 0.6


def _cmp_icgn(st_g, st_o, U_g, U_o, F_g=None, F_o=None, max_status_mismatch=0.002, name=""):
    """Every subset converged on both sides, whatever its iteration count:
    |dU| < 1e-3 px and |dF| < 1e-5. The reference's codes (-1, 101, 102, the
    converged range) must agree on all but `max_status_mismatch` of the
    subsets; the count-mismatch fraction and the worst differences are
    printed."""
    conv_g = (st_g >= 1) & (st_g <= 100)
    conv_o = (st_o >= 1) & (st_o <= 100)
    codes = lambda s: np.where((s >= 1) & (s <= 100), 1, s)
    code_mis = np.mean(codes(st_g) != codes(st_o))
    both = conv_g & conv_o
    assert both.sum() > 0
    cnt_mis = np.sum(both & (st_g != st_o))
    du = np.abs(U_g.reshape(-1, 2) - U_o.reshape(-1, 2)).max(axis=1)
    line = (f"{name} N={len(st_g)} converged(both)={both.sum()} code mismatch={code_mis:.4%} "
            f"count mismatch={cnt_mis} ({cnt_mis / both.sum():.4%}) max|dU|={np.max(du[both]):.3e}")
    if F_g is not None:
        dF = np.abs(F_g.reshape(-1, 4) - F_o.reshape(-1, 4)).max(axis=1)
        line += f" max|dF|={np.max(dF[both]):.3e}"
    _print(line)
    assert code_mis <= max_status_mismatch, (
        np.unique(st_g, return_counts=True), np.unique(st_o, return_counts=True))
    assert np.all(np.isfinite(du[both]))
    assert np.max(du[both]) < 1e-3, np.max(du[both])
    if F_g is not None:
        assert np.max(dF[both]) < 1e-5, np.max(dF[both])
    err = (st_g == -1)
    assert np.all(np.isnan(U_g.reshape(-1, 2)[err]))
    return both.sum()


@pytest.mark.parametrize("plain", [False, True])
def test_icgn6_plain_cfg(ctx, oracle, plain_cfg, plain):
    from staq_dic_amd import synth
    from staq_dic_amd.api import local_icgn_raw
    fr, coords, win = plain_cfg
    truth = synth.truth_at(fr.field, coords)
    U0 = np.round(truth[:, 4:6]).reshape(-1)
    _print("icgn6 512^2 plain" if plain else "icgn6 512^2 quadtree rules")
    Ug, Fg, Hg, sg = local_icgn_raw(U0, coords, win, fr.f, fr.g, fr.mask, tol=1e-3, plain=plain,
                                    ctx=ctx)
    of = oracle.Frame(fr.f, fr.g, fr.mask)
    Uo, Fo, Ho, so = oracle.local_icgn(of, U0, coords, win, tol=1e-3, plain=plain)
    n = _cmp_icgn(sg, so, Ug, Uo, Fg, Fo)
    assert n > 800
    ok = (sg >= 1) & (sg <= 100) & (so >= 1) & (so <= 100)
    # the device gradient is FP32: compare H relative to each subset's largest entry
    scale = np.abs(Ho[ok]).max(axis=1, keepdims=True)
    assert np.max(np.abs(Hg[ok] - Ho[ok]) / scale) < 2e-5
    err = np.hypot(Ug[0::2] - truth[:, 4], Ug[1::2] - truth[:, 5])[ok]
    assert np.sqrt(np.mean(err ** 2)) < 0.01


def test_icgn2_al(ctx, oracle, plain_cfg):
    from staq_dic_amd import synth
    from staq_dic_amd.api import subpb1_raw
    fr, coords, win = plain_cfg
    N = coords.shape[0]
    truth = synth.truth_at(fr.field, coords)
    rng = np.random.default_rng(5)
    UOld = (truth[:, 4:6] + rng.uniform(-0.3, 0.3, (N, 2))).reshape(-1)
    FOld = truth[:, :4].reshape(-1)
    vdual = rng.uniform(-0.05, 0.05, 2 * N)
    for mu in (1e-3, 0.0):
        _print("icgn2 mu", mu)
        Ug, sg = subpb1_raw(UOld, FOld, vdual, coords, win, win, mu, fr.f, fr.g, fr.mask,
                            tol=1e-3, ctx=ctx)
        of = oracle.Frame(fr.f, fr.g, fr.mask)
        Uo, so = oracle.subpb1(of, UOld, FOld, vdual, coords, win, win, mu, tol=1e-3)
        _cmp_icgn(sg, so, Ug, Uo)


def test_icgn2_plain_and_plain_refills(ctx, oracle, plain_cfg):
    """The uniform-grid 2-DOF step (func/Subpb1.m calling funICGN_Subpb1.m with
    the node coordinates as given) against the oracle's plain mode; then the
    plain twins end to end: Subpb1 with its gridfit and LocalICGN with its
    plane fit, rmoutliers and nearest fill, against the analytic field."""
    from staq_dic_amd import synth
    from staq_dic_amd.api import LocalICGN, Subpb1, subpb1_raw
    fr, coords, win = plain_cfg
    N = coords.shape[0]
    truth = synth.truth_at(fr.field, coords)
    rng = np.random.default_rng(6)
    UOld = (truth[:, 4:6] + rng.uniform(-0.3, 0.3, (N, 2))).reshape(-1)
    FOld = truth[:, :4].reshape(-1)
    vdual = rng.uniform(-0.05, 0.05, 2 * N)
    of = oracle.Frame(fr.f, fr.g, fr.mask)
    for mu in (1e-3, 0.0):
        _print("icgn2 plain mu", mu)
        Ug, sg = subpb1_raw(UOld, FOld, vdual, coords, win, win, mu, fr.f, fr.g, fr.mask,
                            tol=1e-3, plain=True, ctx=ctx)
        Uo, so = oracle.subpb1(of, UOld, FOld, vdual, coords, win, win, mu, tol=1e-3, plain=True)
        _cmp_icgn(sg, so, Ug, Uo, name="icgn2 plain")
    para = {"winsize": int(win[0]), "winstepsize": 16}
    Df = {"ImgRefMask": fr.mask}
    U1, _, _, st, nb = Subpb1(UOld, FOld, 0 * FOld, vdual, coords, Df, fr.f, fr.g, 1e-3, 0.0,
                              None, 2, para, tol=1e-3, ctx=ctx)
    e = np.hypot(U1[0::2] - truth[:, 4], U1[1::2] - truth[:, 5])
    _print(f"Subpb1 + gridfit: {nb} bad, median {np.median(e):.4f} px, max {e.max():.4f} px")
    assert np.isfinite(U1).all() and np.median(e) < 0.05
    U0 = np.round(truth[:, 4:6]).reshape(-1)
    U6, F6, _, _, st6, nb6 = LocalICGN(U0, coords, Df, fr.f, fr.g, para, tol=1e-3, ctx=ctx)
    e6 = np.hypot(U6[0::2] - truth[:, 4], U6[1::2] - truth[:, 5])
    eF = np.abs(F6.reshape(-1, 4) - truth[:, :4]).max(axis=1)
    _print(f"LocalICGN + plane fit: {nb6} bad, median {np.median(e6):.4f} px, "
           f"median |dF| {np.median(eF):.2e}")
    assert np.isfinite(U6).all() and np.isfinite(F6).all()
    assert np.median(e6) < 0.02 and np.median(eF) < 2e-3


def test_hole_masked_path(ctx, oracle):
    """Quadtree mesh around a hole: masked windows, window growth, bwselect
    regions, the -1 size-mismatch quirk and 102 all go through the device."""
    from staq_dic_amd import synth
    from staq_dic_amd.api import local_icgn_raw, subpb1_raw
    fr, coords, win = synth.config_hole(H=384, W=384, n_speckles=1700, max_disp=3.0, radius=60,
                                        margin=24, band=(96, 40))
    truth = synth.truth_at(fr.field, coords)
    U0 = np.round(truth[:, 4:6]).reshape(-1)
    _print("hole 384^2: icgn6, N =", coords.shape[0])
    Ug, Fg, Hg, sg = local_icgn_raw(U0, coords, win, fr.f, fr.g, fr.mask, tol=1e-3, ctx=ctx)
    of = oracle.Frame(fr.f, fr.g, fr.mask)
    Uo, Fo, Ho, so = oracle.local_icgn(of, U0, coords, win, tol=1e-3)
    assert (so == -1).sum() > 5 and (so == 102).sum() > 0
    _cmp_icgn(sg, so, Ug, Uo, Fg, Fo)
    _print("hole 384^2: icgn2")
    FOld = np.where(np.isfinite(Fo), Fo, 0.0)
    UOld = np.where(np.isfinite(Uo), Uo, U0)
    vd = np.zeros_like(UOld)
    Ug2, sg2 = subpb1_raw(UOld, FOld, vd, coords, win, win, 1e-3, fr.f, fr.g, fr.mask, tol=1e-3,
                          ctx=ctx)
    Uo2, so2 = oracle.subpb1(of, UOld, FOld, vd, coords, win, win, 1e-3, tol=1e-3)
    _cmp_icgn(sg2, so2, Ug2, Uo2)


def test_reference_frames(ctx, oracle, ref_frames):
    """The reference's own circular-hole frames 000/001, its 1434-node mesh and
    winsize_List: integer search (as IntegerSearch.m is called by the driver,
    g and f swapped, radius 6), the 6-DOF local step (winsize 32 everywhere,
    masked windows, growth, -1 / 102 codes) and the 2-DOF AL step on the
    per-node windows, each against the oracle."""
    from staq_dic_amd import funIntegerSearch
    from staq_dic_amd.api import local_icgn_raw, subpb1_raw
    r = ref_frames
    f, g, m = r["f"], r["g"], r["mask"]
    H, W = f.shape
    _print("reference frames: integer search r=6, w=32, step 16 (g, f order)")
    x, y, u, v, cc = funIntegerSearch(g, f, 6, [1, H], [1, W], 32, 16, ctx=ctx, slot=2)
    xo, yo, uo, vo, cco = oracle.integer_search(g, f, 6, [1, H], [1, W], 32, 16)
    np.testing.assert_array_equal(x, xo)
    np.testing.assert_array_equal(y, yo)
    fin = np.isfinite(uo)
    np.testing.assert_array_equal(np.isfinite(u), fin)
    np.testing.assert_array_equal(np.floor(u[fin]), np.floor(uo[fin]))
    np.testing.assert_array_equal(np.floor(v[fin]), np.floor(vo[fin]))
    np.testing.assert_allclose(u[fin], uo[fin], atol=1e-3 + 1e-12)
    np.testing.assert_allclose(v[fin], vo[fin], atol=1e-3 + 1e-12)

    N = r["coords"].shape[0]
    U1 = r["U1"]
    U0 = np.round(U1)
    win = np.full(N, r["winsize"])
    _print("reference frames: icgn6, N =", N)
    Ug, Fg, Hg, sg = local_icgn_raw(U0, r["coords"], win, f, g, m, tol=1e-2, ctx=ctx)
    of = oracle.Frame(f, g, m)
    Uo, Fo, Ho, so = oracle.local_icgn(of, U0, r["coords"], win, tol=1e-2)
    assert (so == -1).sum() > 50 and (so == 102).sum() > 100
    assert _cmp_icgn(sg, so, Ug, Uo, Fg, Fo) > 650

    wl = r["winsize_List"]
    _print("reference frames: icgn2 on winsize_List")
    vd = np.zeros(2 * N)
    Ug2, sg2 = subpb1_raw(U1, r["F1"], vd, r["coords"], wl[:, 0], wl[:, 1], 1e-3, f, g, m,
                          tol=1e-2, ctx=ctx)
    Uo2, so2 = oracle.subpb1(of, U1, r["F1"], vd, r["coords"], wl[:, 0], wl[:, 1], 1e-3, tol=1e-2)
    _cmp_icgn(sg2, so2, Ug2, Uo2)


def test_forced_status_codes(ctx, oracle):
    """Each failure code of the reference, forced from the host."""
    from staq_dic_amd.api import local_icgn_raw
    rng = np.random.default_rng(11)
    H = W = 96
    f = rng.standard_normal((H, W)).astype(np.float32).astype(float)
    g = f.copy()
    mask = np.ones((H, W))
    mask[60:, 60:] = 0      # a masked corner
    f = f * mask
    f[10:30, 60:80] = 0.0   # featureless (exactly zero) patch
    coords = np.array([[48, 48],    # fine
                       [48, 48],    # U0 pushes the warp off the image -> 101
                       [75, 75],    # fully masked -> 102
                       [20, 70],    # zero patch -> 102 (>= 40 % zeros)
                       [48, 48],    # NaN initial guess -> -1
                       [5, 48]],    # window outside the gradient array -> -1
                      dtype=float)
    U0 = np.array([0, 0, 60, 0, 0, 0, 0, 0, np.nan, 0, 0, 0], dtype=float)
    win = np.full(6, 16)
    Ug, Fg, Hg, sg = local_icgn_raw(U0, coords, win, f, g, mask, tol=1e-3, ctx=ctx)
    of = oracle.Frame(f, g, mask)
    Uo, Fo, Ho, so = oracle.local_icgn(of, U0, coords, win, tol=1e-3)
    # subset 0 matches f exactly: b vanishes up to FP32 rounding on the device
    assert so[0] == 1 and sg[0] in (1, 2)
    np.testing.assert_array_equal(sg[1:], so[1:])
    assert list(so[1:]) == [101, 102, 102, -1, -1]
    assert np.all(Hg[2] == 0) and np.all(Hg[4] == 0)


def test_znssd_fwd_bwd(ctx, oracle, plain_cfg):
    from staq_dic_amd import synth
    fr, coords, win = plain_cfg
    sel = np.arange(0, coords.shape[0], 7)
    coords = coords[sel]
    win = win[sel]
    truth = synth.truth_at(fr.field, coords)
    rng = np.random.default_rng(2)
    P = truth + rng.normal(0, [0.003, 0.003, 0.003, 0.003, 0.2, 0.2], truth.shape)
    ctx.set_frame(fr.f, fr.g, fr.mask)
    import ctypes as C
    from staq_dic_amd._lib import d, ip, lp
    N = coords.shape[0]
    cf = np.asfortranarray(coords)
    Pc = np.ascontiguousarray(P)
    w = np.ascontiguousarray(win, dtype=np.int64)
    cost = np.zeros(N)
    st = np.zeros(N, dtype=np.int32)
    ctx.check(ctx.L.staq_znssd_fwd(ctx.h, N, d(Pc), d(cf), lp(w), None, d(cost), ip(st)))
    gP = np.zeros(6 * N)
    gU = np.zeros(2 * N)
    ctx.check(ctx.L.staq_znssd_bwd(ctx.h, N, d(Pc), d(cf), lp(w), None, None, d(gP), d(gU)))
    of = oracle.Frame(fr.f, fr.g, fr.mask)
    tot, co, go, so = oracle.znssd(of, P, coords, win)
    np.testing.assert_array_equal(st, so)
    ok = so == 0
    np.testing.assert_allclose(cost[ok], co[ok], rtol=1e-4, atol=1e-6)
    gP = gP.reshape(N, 6)
    scale = np.abs(go[ok]).max(axis=0) + 1e-12
    assert np.max(np.abs(gP[ok] - go[ok]) / scale) < 2e-3
    np.testing.assert_allclose(gU.reshape(N, 2), gP[:, 4:6])


def test_batch_forms_agree(ctx, plain_cfg):
    """Per-bin launches, the persistent work-queue grid, the one-subset-per-
    launch form and a replayed hipGraph give identical results."""
    from staq_dic_amd import Batch, params, synth
    from staq_dic_amd._lib import IN_U0
    fr, coords, win = plain_cfg
    sel = np.arange(0, coords.shape[0], 3)
    coords = coords[sel]
    win = win[sel]
    truth = synth.truth_at(fr.field, coords)
    U0 = np.round(truth[:, 4:6]).reshape(-1)
    ctx.set_frame(fr.f, fr.g, fr.mask)
    b = Batch(ctx, coords, win)
    b.set_input(IN_U0, U0)
    p = params(tol=1e-3)
    b.icgn6(p)
    ref = b.outputs()
    b.icgn6(params(tol=1e-3, persistent=True))
    got = b.outputs()
    for r, g in zip(ref, got):
        np.testing.assert_array_equal(r, g)
    # the per-subset form runs the workgroup kernel (block reductions), the
    # batch the wavefront kernel: same arithmetic up to FP32 summation order
    b.icgn6_per_subset(params(tol=1e-3, block_size=128))
    got = b.outputs()
    assert np.mean(ref[3] != got[3]) < 0.01
    ok = (ref[3] == got[3]) & (ref[3] >= 1) & (ref[3] <= 100)
    np.testing.assert_allclose(ref[0].reshape(-1, 2)[ok], got[0].reshape(-1, 2)[ok], atol=1e-4)
    np.testing.assert_allclose(ref[1].reshape(-1, 4)[ok], got[1].reshape(-1, 4)[ok], atol=1e-5)
    b.graph_capture(p, 6)
    for _ in range(3):
        b.graph_launch()
        got = b.outputs()
        for r, g in zip(ref, got):
            np.testing.assert_array_equal(r, g)
    # two and four wavefronts per subset (the schedule for long subsets), and
    # the autotuned schedule: the FP32 summation order changes, nothing else
    for form in (-2, -4):
        b.icgn6(params(tol=1e-3, block_size=form))
        got = b.outputs()
        assert np.mean(ref[3] != got[3]) < 0.01, form
        ok = (ref[3] == got[3]) & (ref[3] >= 1) & (ref[3] <= 100)
        np.testing.assert_allclose(ref[0].reshape(-1, 2)[ok], got[0].reshape(-1, 2)[ok], atol=1e-4)
        np.testing.assert_allclose(ref[1].reshape(-1, 4)[ok], got[1].reshape(-1, 4)[ok], atol=1e-5)
    b.autotune(p, 6)
    b.icgn6(p)
    got = b.outputs()
    assert np.mean(ref[3] != got[3]) < 0.01
    ok = (ref[3] == got[3]) & (ref[3] >= 1) & (ref[3] <= 100)
    np.testing.assert_allclose(ref[0].reshape(-1, 2)[ok], got[0].reshape(-1, 2)[ok], atol=1e-4)
    b.close()


@pytest.mark.parametrize("form", [0, -1, -2, -4, 1, 4, 64, 256])
def test_multi_wave_forms_on_masked_subsets(ctx, oracle, form):
    """Every kernel form forced on every bin through the masked path (region
    floods, window growth, -1 and 102) and the 2-DOF step, against the oracle:
    one, two and four wavefronts per subset (0 / -1, -2, -4), the workgroup
    kernel (64, 256); forms 1 and 4 force the lean kernel on every bin, so
    every subset that touches the mask goes through its hand-over to the
    fallback launch."""
    from staq_dic_amd import Batch, params, synth
    from staq_dic_amd._lib import IN_FOLD, IN_U0, IN_UOLD
    fr, coords, win = synth.config_hole(H=384, W=384, n_speckles=1700, max_disp=3.0, radius=60,
                                        margin=24, band=(96, 40))
    truth = synth.truth_at(fr.field, coords)
    U0 = np.round(truth[:, 4:6]).reshape(-1)
    _print("hole 384^2 with kernel form", form, "N =", coords.shape[0])
    ctx.set_frame(fr.f, fr.g, fr.mask)
    b = Batch(ctx, coords, win)
    b.set_input(IN_U0, U0)
    b.icgn6(params(tol=1e-3, block_size=form))
    Ug, Fg, Hg, sg = b.outputs()
    of = oracle.Frame(fr.f, fr.g, fr.mask)
    Uo, Fo, Ho, so = oracle.local_icgn(of, U0, coords, win, tol=1e-3)
    _cmp_icgn(sg, so, Ug, Uo, Fg, Fo)
    # a second step on the same inputs gives the same bits
    b.icgn6(params(tol=1e-3, block_size=form))
    for x, y in zip(b.outputs(), (Ug, Fg, Hg, sg)):
        np.testing.assert_array_equal(x, y)
    FOld = np.where(np.isfinite(Fo), Fo, 0.0)
    UOld = np.where(np.isfinite(Uo), Uo, U0)
    b.set_input(IN_UOLD, UOld)
    b.set_input(IN_FOLD, FOld)
    b.icgn2(params(tol=1e-3, mu=1e-3, block_size=form))
    Ug2, _, _, sg2 = b.outputs()
    Uo2, so2 = oracle.subpb1(of, UOld, FOld, np.zeros_like(UOld), coords, win, win, 1e-3, tol=1e-3)
    _cmp_icgn(sg2, so2, Ug2, Uo2)
    # the same step again on the same inputs gives the same bits (a kernel
    # whose results depend on what ran before, or on timing, fails here)
    b.icgn2(params(tol=1e-3, mu=1e-3, block_size=form))
    Ur, _, _, sr = b.outputs()
    np.testing.assert_array_equal(sr, sg2)
    np.testing.assert_array_equal(Ur, Ug2)
    b.close()


def test_znssd_autograd(ctx, oracle, plain_cfg):
    """The torch.autograd.Function: forward cost and the gradient reaching a
    nodal displacement tensor through P[:, 4:6] = U, against the oracle."""
    import torch
    from staq_dic_amd import synth
    from staq_dic_amd.autograd import znssd
    fr, coords, win = plain_cfg
    sel = np.arange(0, coords.shape[0], 5)
    coords, win = coords[sel], win[sel]
    truth = synth.truth_at(fr.field, coords)
    rng = np.random.default_rng(4)
    P0 = truth + rng.normal(0, [0.002, 0.002, 0.002, 0.002, 0.15, 0.15], truth.shape)
    ctx.set_frame(fr.f, fr.g, fr.mask)
    _print("znssd autograd, N =", coords.shape[0])
    dev = torch.device("cuda", 0)
    U = torch.tensor(P0[:, 4:6], dtype=torch.float64, device=dev, requires_grad=True)
    A = torch.tensor(P0[:, :4], dtype=torch.float64, device=dev, requires_grad=True)
    P = torch.cat([A, U], dim=1)
    cost, status = znssd(P, torch.tensor(coords, device=dev),
                         torch.tensor(win, dtype=torch.int32, device=dev), ctx)
    w = torch.linspace(0.5, 1.5, cost.shape[0], dtype=torch.float64, device=dev)
    (w * cost).sum().backward()
    torch.cuda.synchronize()
    of = oracle.Frame(fr.f, fr.g, fr.mask)
    tot, co, go, so = oracle.znssd(of, P0, coords, win)
    st = status.cpu().numpy()
    np.testing.assert_array_equal(st, so)
    ok = so == 0
    np.testing.assert_allclose(cost.detach().cpu().numpy()[ok], co[ok], rtol=1e-4, atol=1e-6)
    wn = w.cpu().numpy()[:, None]
    gP = np.concatenate([A.grad.cpu().numpy(), U.grad.cpu().numpy()], axis=1)
    ref = go * wn
    scale = np.abs(ref[ok]).max(axis=0) + 1e-12
    assert np.max(np.abs(gP[ok] - ref[ok]) / scale) < 2e-3


def test_multi_pair_batch(ctx, oracle):
    """Several frame pairs resident at once, one batch over all of them (each
    subset names its pair by slot), against the oracle pair by pair: the
    multi-instance form, 6-DOF and 2-DOF."""
    from staq_dic_amd import Batch, params, synth
    from staq_dic_amd._lib import IN_FOLD, IN_U0, IN_UOLD
    pairs = []
    for j, (H, W, field) in enumerate([
            (384, 448, synth.AffineField(2.2, -1.1, 0.015, -0.008, 0.004, 0.012)),
            (448, 384, synth.AffineField(-3.1, 0.7, -0.01, 0.006, 0.009, -0.004))]):
        fr = synth.make_frame(H, W, 2000, 5000 + j, field)
        coords = synth.plain_grid_nodes(H, W, 24, 24, margin=24)
        coords = coords[(coords[:, 0] <= H - 28) & (coords[:, 1] <= W - 28)]
        pairs.append((fr, coords))
    slots = [3, 1]
    for (fr, _), sl in zip(pairs, slots):
        ctx.set_frame(fr.f, fr.g, fr.mask, slot=sl)
    coords = np.concatenate([c for _, c in pairs])
    slot = np.concatenate([np.full(len(c), sl, dtype=np.int32) for (_, c), sl in zip(pairs, slots)])
    win = np.where(np.arange(len(coords)) % 2 == 0, 24, 16).astype(np.int64)
    truth = np.concatenate([synth.truth_at(fr.field, c) for fr, c in pairs])
    U0 = np.round(truth[:, 4:6]).reshape(-1)
    _print("two pairs in slots", slots, "N =", len(coords))
    b = Batch(ctx, coords, win, slot=slot)
    b.set_input(IN_U0, U0)
    b.icgn6(params(tol=1e-3))
    Ug, Fg, Hg, sg = b.outputs()
    b.set_input(IN_UOLD, Ug)
    b.set_input(IN_FOLD, np.where(np.isfinite(Fg), Fg, 0.0))
    b.icgn2(params(tol=1e-3, mu=1e-3))
    Ug2, _, _, sg2 = b.outputs()
    b.close()
    i0 = 0
    for fr, c in pairs:
        n = len(c)
        sel = slice(i0, i0 + n)
        of = oracle.Frame(fr.f, fr.g, fr.mask)
        Uo, Fo, Ho, so = oracle.local_icgn(of, U0.reshape(-1, 2)[sel].reshape(-1), c, win[sel],
                                           tol=1e-3)
        _cmp_icgn(sg[sel], so, Ug.reshape(-1, 2)[sel].reshape(-1), Uo,
                  Fg.reshape(-1, 4)[sel].reshape(-1), Fo)
        UOld = Ug.reshape(-1, 2)[sel].reshape(-1)
        FOld = np.where(np.isfinite(Fg), Fg, 0.0).reshape(-1, 4)[sel].reshape(-1)
        Uo2, so2 = oracle.subpb1(of, UOld, FOld, np.zeros_like(UOld), c, win[sel], win[sel], 1e-3,
                                 tol=1e-3)
        _cmp_icgn(sg2[sel], so2, Ug2.reshape(-1, 2)[sel].reshape(-1), Uo2)
        i0 += n


def test_integer_search_large_radius(ctx, oracle):
    """The large-deformation search geometry: window 64, radius 48, a field
    with ~30 px translation, against the oracle point by point."""
    from staq_dic_amd import funIntegerSearch, synth
    field = synth.AffineField(30.0, -25.0, 0.01, -0.005, 0.004, 0.008)
    fr = synth.make_frame(640, 640, 6000, 6001, field)
    H, W = fr.f.shape
    _print("integer search 640^2, r=48, w=64, step 32")
    x, y, u, v, cc = funIntegerSearch(fr.f, fr.g, 48, [1, H], [1, W], 64, 32, ctx=ctx, slot=2)
    xo, yo, uo, vo, cco = oracle.integer_search(fr.f, fr.g, 48, [1, H], [1, W], 64, 32)
    assert x.size >= 100
    np.testing.assert_array_equal(x, xo)
    np.testing.assert_array_equal(y, yo)
    fin = np.isfinite(uo)
    np.testing.assert_array_equal(np.isfinite(u), fin)
    np.testing.assert_array_equal(np.floor(u[fin]), np.floor(uo[fin]))
    np.testing.assert_array_equal(np.floor(v[fin]), np.floor(vo[fin]))
    np.testing.assert_allclose(u[fin], uo[fin], atol=1e-3 + 1e-12)
    np.testing.assert_allclose(v[fin], vo[fin], atol=1e-3 + 1e-12)
    tu = 30.0 + 0.01 * x - 0.005 * y
    tv = -25.0 + 0.004 * x + 0.008 * y
    assert np.max(np.abs(u - tu)[fin]) < 0.75 and np.max(np.abs(v - tv)[fin]) < 0.75


def test_reference_level_local_steps(ctx, ref_frames):
    """LocalICGNQuadtree and Subpb1Quadtree as the reference's driver calls them
    (bad-subset classification and refill included) on the reference's own
    frames: no NaN left, and the result lands on the stored answer."""
    from staq_dic_amd import LocalICGNQuadtree, Subpb1Quadtree
    r = ref_frames
    N = r["coords"].shape[0]
    Df = {"ImgRefMask": r["mask"]}
    DICpara = {"winsize": r["winsize"], "winsize_List": r["winsize_List"]}
    _print("LocalICGNQuadtree on the reference frames")
    U, F, Hp, t, conv, nbad, hole = LocalICGNQuadtree(np.round(r["U1"]), r["coords"], Df, r["f"],
                                                      r["g"], DICpara, "GaussNewton", 1e-2, ctx=ctx)
    assert np.isfinite(U).all() and np.isfinite(F).all()
    assert nbad == int(((conv < 0) | (conv > 99)).sum() - (conv == 102).sum())
    from staq_dic_amd import classify_local, classify_subpb1
    bad, _ = classify_local(conv)
    d = np.hypot(U[0::2] - r["U1"][0::2], U[1::2] - r["U1"][1::2])
    # kept subsets land on the stored answer; refilled ones (about half of the
    # nodes here: the masked 102 and -1 subsets) follow it loosely
    assert np.median(d[~bad]) < 5e-3 and np.percentile(d[~bad], 90) < 0.06
    assert np.median(d[bad]) < 0.1 and np.percentile(d, 90) < 0.25
    _print("Subpb1Quadtree on the reference frames")
    U2, _, _, conv2, nbad2 = Subpb1Quadtree(U, F, 0 * F, 0 * U, r["coords"], Df, r["f"], r["g"],
                                            1e-3, 0.0, None, 2, DICpara, "GaussNewton", 1e-2,
                                            ctx=ctx)
    assert np.isfinite(U2).all()
    bad2, _ = classify_subpb1(conv2)
    d2 = np.hypot(U2[0::2] - r["U1"][0::2], U2[1::2] - r["U1"][1::2])
    assert np.median(d2[~bad2]) < 0.02 and np.percentile(d2, 90) < 0.25


def test_reference_level_integer_search(ctx, oracle, ref_frames):
    """IntegerSearch as the reference's driver calls it (g and f swapped,
    radius 6, hole points NaN, median-test outlier filter): the device search
    followed by the host filter equals the oracle search followed by the same
    filter."""
    from staq_dic_amd import IntegerSearch
    from staq_dic_amd.outliers import funRemoveOutliers
    r = ref_frames
    H, W = r["f"].shape
    P = {"winsize": 32, "winstepsize": 16, "SizeOfFFTSearchRegion": 6,
         "gridxyROIRange": {"gridx": [1, H], "gridy": [1, W]}, "ImgRefMask": r["mask"]}
    _print("IntegerSearch on the reference frames")
    P2, x0, y0, u, v, cc = IntegerSearch(r["g"], r["f"], None, P, ctx=ctx)
    xo, yo, uo, vo, cco = oracle.integer_search(r["g"], r["f"], 6, [1, H], [1, W], 32, 16)
    hole = r["mask"][xo.astype(int) - 1, yo.astype(int) - 1] == 0
    uo = np.where(hole, np.nan, uo)
    vo = np.where(hole, np.nan, vo)
    uo, vo, _, _, _, _ = funRemoveOutliers(uo, vo, {}, 0, 100)
    np.testing.assert_array_equal(np.isnan(u), np.isnan(uo))
    fin = np.isfinite(uo)
    np.testing.assert_allclose(u[fin], uo[fin], atol=1e-3 + 1e-9)
    np.testing.assert_allclose(v[fin], vo[fin], atol=1e-3 + 1e-9)
    assert np.isnan(u[hole]).all() and fin.sum() > 100


def test_lean_kernel_matches_general(ctx, oracle):
    """The lean kernel (windows without masked pixels, the default form of
    such bins) against the general wavefront kernel on the same bins with the
    same number of wavefronts, and both against the oracle: the same
    arithmetic up to the FP32 summation order of the per-iteration sums, so
    both hold the oracle's bar on every subset converged on both sides.
    6-DOF and 2-DOF."""
    from staq_dic_amd import Batch, params, synth
    from staq_dic_amd._lib import IN_FOLD, IN_U0, IN_UOLD
    fr, coords, win = synth.config_hole(H=384, W=384, n_speckles=1700, max_disp=3.0, radius=60,
                                        margin=24, band=(96, 40))
    truth = synth.truth_at(fr.field, coords)
    U0 = np.round(truth[:, 4:6]).reshape(-1)
    ctx.set_frame(fr.f, fr.g, fr.mask)
    b = Batch(ctx, coords, win)
    clear = [x["bs6"] in (1, 2, 4, 8, 16) for x in b.bins()]
    assert any(clear) and not all(clear), b.bins()
    b.set_input(IN_U0, U0)
    p6 = params(tol=1e-3)
    p2 = params(tol=1e-3, mu=1e-3)

    def run(mode, forms):
        b.set_forms(forms, mode)
        (b.icgn6 if mode == 6 else b.icgn2)(p6 if mode == 6 else p2)
        return b.outputs()

    of = oracle.Frame(fr.f, fr.g, fr.mask)
    Uo6, Fo6, _, so6 = oracle.local_icgn(of, U0, coords, win, tol=1e-3)

    def compare(out_l, out_g, oracle_out, with_f):
        Ul, Fl, _, sl = out_l
        Ug, Fg, _, sg = out_g
        Uo, Fo, so = oracle_out
        _cmp_icgn(sl, sg, Ul, Ug, Fl if with_f else None, Fg if with_f else None,
                  name="lean vs general")
        _cmp_icgn(sl, so, Ul, Uo, Fl if with_f else None, Fo if with_f else None,
                  name="lean vs oracle")
        _cmp_icgn(sg, so, Ug, Uo, Fg if with_f else None, Fo if with_f else None,
                  name="general vs oracle")

    for nw in (1, 2, 4):
        general = 0 if nw == 1 else -nw
        _print("6-DOF: lean kernel with", nw, "wavefronts against the general kernel")
        out_l = run(6, [nw if c else general for c in clear])
        out_g = run(6, [general] * len(clear))
        compare(out_l, out_g, (Uo6, Fo6, so6), True)
    UOld = np.where(np.isfinite(out_g[0]), out_g[0], U0)
    FOld = np.where(np.isfinite(out_g[1]), out_g[1], 0.0)
    b.set_input(IN_UOLD, UOld)
    b.set_input(IN_FOLD, FOld)
    Uo2, so2 = oracle.subpb1(of, UOld, FOld, np.zeros_like(UOld), coords, win, win, 1e-3, tol=1e-3)
    for nw in (1, 2):
        general = 0 if nw == 1 else -nw
        _print("2-DOF: lean kernel with", nw, "wavefronts against the general kernel")
        out_l = run(2, [nw if c else general for c in clear])
        out_g = run(2, [general] * len(clear))
        compare(out_l, out_g, (Uo2, None, so2), False)
    b.close()


def test_handed_back_subsets_move_to_the_general_kernel(ctx, plain_cfg):
    """Subsets the lean kernel hands to the fallback launch (here: windows
    whose warped support meets a patch of exact zeros planted in g, which the
    quadtree rules treat as masked) move to the general kernel's bins at the
    next reorder; the results stay the same up to the fallback's summation
    order, and nothing is handed back afterwards."""
    from staq_dic_amd import Batch, params, synth
    from staq_dic_amd._lib import IN_U0
    fr, coords, win = plain_cfg
    truth = synth.truth_at(fr.field, coords)
    U0 = np.round(truth[:, 4:6]).reshape(-1)
    g = fr.g.copy()
    g[240:246, 240:246] = 0.0
    g[100:104, 380:384] = 0.0
    ctx.set_frame(fr.f, g, fr.mask)
    b = Batch(ctx, coords, win)
    b.set_input(IN_U0, U0)
    p = params(tol=1e-3)

    def general_count():
        return sum(x["count"] for x in b.bins() if x["bs6"] not in (1, 2, 4, 8, 16))

    n0 = general_count()
    b.icgn6(p)
    first = b.outputs()
    b.reorder()
    n1 = general_count()
    _print("subsets in general-kernel bins:", n0, "->", n1)
    assert n1 > n0
    b.icgn6(p)
    second = b.outputs()
    _cmp_icgn(first[3], second[3], first[0], second[0], first[1], second[1])
    assert np.mean(first[3] == second[3]) > 0.99
    b.reorder()
    assert general_count() == n1
    b.close()


@pytest.mark.parametrize("form", [8, 16, 256])
def test_wide_forms_on_64px_windows(ctx, oracle, plain_cfg, form):
    """64-px windows (4225 pixels): the lean kernel with eight (the default
    form of such bins) and sixteen wavefronts per subset and the workgroup
    kernel, each against the oracle, 6-DOF then 2-DOF on its result."""
    from staq_dic_amd import Batch, params, synth
    from staq_dic_amd._lib import IN_FOLD, IN_U0, IN_UOLD
    fr, _, _ = plain_cfg
    H, W = fr.f.shape
    xs = np.arange(40, 441, 12)  # warped windows stay inside the 512^2 frame
    X, Y = np.meshgrid(xs, xs, indexing="ij")
    coords = np.stack([X.ravel(order="F"), Y.ravel(order="F")], axis=1).astype(float)
    win = np.full(len(coords), 64, dtype=np.int64)
    truth = synth.truth_at(fr.field, coords)
    U0 = np.round(truth[:, 4:6]).reshape(-1)
    ctx.set_frame(fr.f, fr.g, fr.mask)
    _print("64-px windows, form", form, "N =", len(coords))
    b = Batch(ctx, coords, win)
    assert len(b.bins()) == 1 and b.bins()[0]["bs6"] == 8, b.bins()
    b.set_forms([form], 6)
    b.set_forms([form], 2)
    b.set_input(IN_U0, U0)
    b.icgn6(params(tol=1e-3))
    Ug, Fg, _, sg = b.outputs()
    b.reorder()
    # the lean forms kept every subset: nothing was handed to the fallback
    assert len(b.bins()) == 1, b.bins()
    of = oracle.Frame(fr.f, fr.g, fr.mask)
    Uo, Fo, _, so = oracle.local_icgn(of, U0, coords, win, tol=1e-3)
    assert _cmp_icgn(sg, so, Ug, Uo, Fg, Fo) > 0.95 * len(coords)
    UOld = np.where(np.isfinite(Ug), Ug, U0)
    FOld = np.where(np.isfinite(Fg), Fg, 0.0)
    b.set_input(IN_UOLD, UOld)
    b.set_input(IN_FOLD, FOld)
    b.icgn2(params(tol=1e-3, mu=1e-3))
    Ug2, _, _, sg2 = b.outputs()
    b.close()
    Uo2, so2 = oracle.subpb1(of, UOld, FOld, np.zeros_like(UOld), coords, win, win, 1e-3,
                             tol=1e-3)
    _cmp_icgn(sg2, so2, Ug2, Uo2)


def test_large_deformation_incremental_pair(ctx, oracle):
    """The large-deformation configuration at 768^2: a 16-bit pattern shown
    in frames 39 and 40 of the stretch-and-bump sequence, the incremental pair
    on a quadtree refined towards the bump (leaves 4-64 px, windows 16-64 px),
    all in one batch: 6-DOF step against the oracle subset by subset, the
    incremental displacement against the analytic field, and the integer
    search (window 64, radius 48) against the oracle point by point."""
    from staq_dic_amd import Batch, params, synth
    from staq_dic_amd.api import integer_search_slot
    from staq_dic_amd._lib import IN_U0
    fr, coords, win = synth.config_large(H=768, W=768, n_speckles=6750, radii=(90, 170, 280, 400),
                                         margin=48)
    assert set(np.unique(win)) == {16, 32, 64}, np.unique(win)
    truth = synth.truth_at(fr.field, coords)
    U0 = np.round(truth[:, 4:6]).reshape(-1)
    ctx.set_frame(fr.f, fr.g, fr.mask, slot=0)
    _print("large-deformation pair 768^2, N =", len(coords))
    b = Batch(ctx, coords, win)
    b.set_input(IN_U0, U0)
    b.icgn6(params(tol=1e-3))
    Ug, Fg, _, sg = b.outputs()
    b.close()
    of = oracle.Frame(fr.f, fr.g, fr.mask)
    Uo, Fo, _, so = oracle.local_icgn(of, U0, coords, win, tol=1e-3)
    assert _cmp_icgn(sg, so, Ug, Uo, Fg, Fo) > 0.9 * len(coords)
    ok = (sg >= 1) & (sg <= 100) & (win >= 32)
    err = np.hypot(Ug[0::2] - truth[:, 4], Ug[1::2] - truth[:, 5])[ok]
    assert np.sqrt(np.mean(err ** 2)) < 0.01
    H, W = fr.f.shape
    x, y, u, v, cc = integer_search_slot(ctx, 0, 48, [1, H], [1, W], 64, 64)
    xo, yo, uo, vo, co = oracle.integer_search(fr.f, fr.g, 48, [1, H], [1, W], 64, 64)
    np.testing.assert_array_equal(x, xo)
    np.testing.assert_array_equal(u, uo)
    np.testing.assert_array_equal(v, vo)


def test_znssd_sinusoid_field(ctx, oracle):
    """The adjoint configuration's field (u = 4 sin(2 pi x/512), v = 3 cos(2 pi
    y/768)) at 768^2 on its 16-px node grid: forward cost and vector-Jacobian
    product against the FP64 oracle, and the gradient against central finite
    differences of the oracle's cost (h = 1e-3 px) on u and v."""
    from staq_dic_amd import synth
    from staq_dic_amd.autograd import znssd
    import torch
    fr, coords, win = synth.config_adjoint(H=768, W=768, n_speckles=6750)
    truth = synth.truth_at(fr.field, coords)
    rng = np.random.default_rng(7)
    P0 = truth + rng.normal(0, [0.002, 0.002, 0.002, 0.002, 0.1, 0.1], truth.shape)
    ctx.set_frame(fr.f, fr.g, fr.mask, slot=0)
    _print("znssd sinusoid 768^2, N =", len(coords))
    dev = torch.device("cuda", 0)
    P = torch.tensor(P0, dtype=torch.float64, device=dev, requires_grad=True)
    cost, status = znssd(P, torch.tensor(coords, device=dev),
                         torch.tensor(win, dtype=torch.int32, device=dev), ctx)
    cost.sum().backward()
    torch.cuda.synchronize()
    of = oracle.Frame(fr.f, fr.g, fr.mask)
    tot, co, go, so = oracle.znssd(of, P0, coords, win)
    st = status.cpu().numpy()
    np.testing.assert_array_equal(st, so)
    ok = so == 0
    assert ok.mean() > 0.95
    np.testing.assert_allclose(cost.detach().cpu().numpy()[ok], co[ok], rtol=1e-4, atol=1e-6)
    g = P.grad.cpu().numpy()
    scale = np.abs(go[ok]).max(axis=0) + 1e-12
    assert np.max(np.abs(g[ok] - go[ok]) / scale) < 2e-3
    # central differences of the FP64 cost on the translation parameters
    h = 1e-3
    sel = np.flatnonzero(ok)[:: max(1, ok.sum() // 40)]
    for q in (4, 5):
        Pp, Pm = P0[sel].copy(), P0[sel].copy()
        Pp[:, q] += h
        Pm[:, q] -= h
        _, cp, _, _ = oracle.znssd(of, Pp, coords[sel], win[sel], want_grad=False)
        _, cm, _, _ = oracle.znssd(of, Pm, coords[sel], win[sel], want_grad=False)
        fd = (cp - cm) / (2 * h)
        np.testing.assert_allclose(g[sel, q], fd, rtol=0.02, atol=2e-3 * np.abs(fd).max())


def test_two_contexts_from_two_threads():
    """Distinct contexts used from distinct host threads at the same time (the
    threading rule of include/staq_dic.h: a context is not re-entrant,
    distinct contexts are independent): each thread's local steps equal the
    same steps run alone, bit for bit."""
    import threading

    from staq_dic_amd import Context, synth
    from staq_dic_amd.api import local_icgn_raw
    cases = []
    for seed in (1101, 1102):
        fr, coords, win = synth.config_hole(seed=seed, H=384, W=384, n_speckles=1700, max_disp=3.0,
                                            radius=60, margin=24, band=(96, 40))
        U0 = np.round(synth.truth_at(fr.field, coords)[:, 4:6]).reshape(-1)
        cases.append((fr, coords, win, U0))
    _print("two contexts: each case alone")
    alone = []
    for fr, coords, win, U0 in cases:
        c = Context(0)
        try:
            alone.append(local_icgn_raw(U0, coords, win, fr.f, fr.g, fr.mask, tol=1e-3, ctx=c))
        finally:
            c.close()
    _print("two contexts: both at once from two threads, three rounds each")
    results = [[], []]
    errors = []

    def work(i):
        fr, coords, win, U0 = cases[i]
        c = Context(0)
        try:
            for _ in range(3):
                results[i].append(local_icgn_raw(U0, coords, win, fr.f, fr.g, fr.mask, tol=1e-3,
                                                 ctx=c))
        except Exception as e:  # reported below
            errors.append(e)
        finally:
            c.close()

    th = [threading.Thread(target=work, args=(i,)) for i in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=120)
    assert not errors, errors
    for i in range(2):
        assert len(results[i]) == 3
        for out in results[i]:
            for a, b in zip(out, alone[i]):
                np.testing.assert_array_equal(a, b)


def test_bad_arguments_with_a_context(ctx):
    """Argument errors through a live context raise with the library's
    message and leave the context usable: no subsets, a window below 2 px,
    an unknown kernel form, a graph launch before any capture."""
    from staq_dic_amd import Batch, StaqError, params, synth
    from staq_dic_amd.api import local_icgn_raw
    fr, coords, win = synth.config_plain(H=256, W=256, n_speckles=800)
    ctx.set_frame(fr.f, fr.g, fr.mask)
    with pytest.raises(StaqError, match="bad batch arguments"):
        Batch(ctx, coords[:0], win[:0])
    with pytest.raises(StaqError):
        Batch(ctx, coords[:4], np.ones(4, dtype=np.int64))
    b = Batch(ctx, coords, win)
    try:
        with pytest.raises(StaqError, match="no captured graph"):
            b.graph_launch()
        with pytest.raises(StaqError):
            b.set_forms([3] * len(b.bins()), 6)
        with pytest.raises(StaqError, match="block_size"):
            b.icgn6(params(tol=1e-3, block_size=3))
    finally:
        b.close()
    # the context still runs a step
    U0 = np.round(synth.truth_at(fr.field, coords)[:, 4:6]).reshape(-1)
    U, F, H, st = local_icgn_raw(U0, coords, win, fr.f, fr.g, fr.mask, tol=1e-3, ctx=ctx)
    assert ((st >= 1) & (st <= 100)).mean() > 0.8


def test_full_size_workloads_against_oracle(ctx, oracle):
    """The measured workloads at their full sizes, every subset against the
    oracle: the 1024^2 hole pair with the reference's 32-px windows (6016
    subsets), with the 0.01 px accuracy gate over the subsets the reference
    keeps; the 4096^2 16-bit large-deformation pair (about 98 k subsets, one
    batch) with the reference's 32-px windows and with the leaf-size rule."""
    from staq_dic_amd import Batch, params, synth
    from staq_dic_amd.api import classify_local
    from staq_dic_amd._lib import IN_U0
    fr, coords, _ = synth.config_hole()
    win = np.full(len(coords), 32, dtype=np.int64)
    truth = synth.truth_at(fr.field, coords)
    U0 = np.round(truth[:, 4:6]).reshape(-1)
    ctx.set_frame(fr.f, fr.g, fr.mask, slot=0)
    _print("hole pair 1024^2, reference windows, N =", len(coords))
    b = Batch(ctx, coords, win)
    b.set_input(IN_U0, U0)
    b.icgn6(params(tol=1e-3))
    Ug, Fg, _, sg = b.outputs()
    b.close()
    Uo, Fo, _, so = oracle.local_icgn(oracle.Frame(fr.f, fr.g, fr.mask), U0, coords, win, tol=1e-3)
    assert _cmp_icgn(sg, so, Ug, Uo, Fg, Fo, name="hole 1024^2 w32") > 0.95 * len(coords)
    bad, _ = classify_local(sg)
    kept = (sg >= 1) & (sg <= 100) & ~bad
    err = np.hypot(Ug[0::2] - truth[:, 4], Ug[1::2] - truth[:, 5])[kept]
    assert np.sqrt(np.mean(err ** 2)) <= 0.01

    fr, coords, leaf = synth.config_large()
    truth = synth.truth_at(fr.field, coords)
    U0 = np.round(truth[:, 4:6]).reshape(-1)
    ctx.set_frame(fr.f, fr.g, fr.mask, slot=0)
    assert len(coords) > 90000
    of = oracle.Frame(fr.f, fr.g, fr.mask)
    for rule, win in (("w32", np.full(len(coords), 32, dtype=np.int64)), ("leaf", leaf)):
        _print("large-deformation pair 4096^2,", rule, "windows, N =", len(coords))
        b = Batch(ctx, coords, win)
        b.set_input(IN_U0, U0)
        b.icgn6(params(tol=1e-3))
        Ug, Fg, _, sg = b.outputs()
        b.close()
        Uo, Fo, _, so = oracle.local_icgn(of, U0, coords, win, tol=1e-3)
        assert _cmp_icgn(sg, so, Ug, Uo, Fg, Fo, name="4096^2 " + rule) > 0.9 * len(coords)
        ok = (sg >= 1) & (sg <= 100) & (win >= 32)
        err = np.hypot(Ug[0::2] - truth[:, 4], Ug[1::2] - truth[:, 5])[ok]
        _print("   RMS error against the analytic field, windows >= 32 px:",
               np.sqrt(np.mean(err ** 2)))
        assert np.sqrt(np.mean(err ** 2)) <= 0.01


def test_full_size_multi_pair_and_adjoint(ctx, oracle):
    """The multi-instance batch at full size (16 pairs of 512^2-2048^2 in
    slots 0-15, 88 k subsets, one batch) against the oracle on every subset,
    pair by pair; the adjoint pair at full size (2048^2, 15 625 subsets)
    against the FP64 oracle's cost and gradient on every subset."""
    import torch

    from staq_dic_amd import Batch, params, synth
    from staq_dic_amd._lib import IN_U0
    from staq_dic_amd.autograd import znssd
    pairs = synth.config_multi()
    for j, (fr, _, _) in enumerate(pairs):
        ctx.set_frame(fr.f, fr.g, fr.mask, slot=j)
    coords = np.concatenate([c for _, c, _ in pairs])
    win = np.concatenate([w for _, _, w in pairs])
    slot = np.concatenate([np.full(len(c), j, dtype=np.int32) for j, (_, c, _) in enumerate(pairs)])
    U0 = np.concatenate([np.round(synth.truth_at(fr.field, c)[:, 4:6]) for fr, c, _ in pairs])
    _print("multi-pair batch, N =", len(coords))
    b = Batch(ctx, coords, win, slot=slot)
    b.set_input(IN_U0, U0.reshape(-1))
    b.icgn6(params(tol=1e-3))
    Ug, Fg, _, sg = b.outputs()
    b.close()
    Uo = np.empty_like(Ug)
    Fo = np.empty_like(Fg)
    so = np.empty_like(sg)
    for j, (fr, c, w) in enumerate(pairs):
        idx = np.flatnonzero(slot == j)
        u, f_, _, s = oracle.local_icgn(oracle.Frame(fr.f, fr.g, fr.mask), U0[idx].reshape(-1),
                                        coords[idx], win[idx], tol=1e-3)
        Uo.reshape(-1, 2)[idx] = u.reshape(-1, 2)
        Fo.reshape(-1, 4)[idx] = f_.reshape(-1, 4)
        so[idx] = s
    assert _cmp_icgn(sg, so, Ug, Uo, Fg, Fo, name="16 pairs") > 0.9 * len(coords)

    fr, coords, win = synth.config_adjoint()
    truth = synth.truth_at(fr.field, coords)
    P0 = truth + np.random.default_rng(7).normal(0, [0.002, 0.002, 0.002, 0.002, 0.1, 0.1],
                                                 truth.shape)
    ctx.set_frame(fr.f, fr.g, fr.mask, slot=0)
    _print("adjoint pair 2048^2, N =", len(coords))
    dev = torch.device("cuda", 0)
    P = torch.tensor(P0, dtype=torch.float64, device=dev, requires_grad=True)
    cost, status = znssd(P, torch.tensor(coords, device=dev),
                         torch.tensor(win, dtype=torch.int32, device=dev), ctx)
    cost.sum().backward()
    torch.cuda.synchronize()
    _, co, go, so = oracle.znssd(oracle.Frame(fr.f, fr.g, fr.mask), P0, coords, win)
    st = status.cpu().numpy()
    np.testing.assert_array_equal(st, so)
    ok = so == 0
    assert ok.mean() > 0.95
    np.testing.assert_allclose(cost.detach().cpu().numpy()[ok], co[ok], rtol=1e-4, atol=1e-6)
    g = P.grad.cpu().numpy()
    scale = np.abs(go[ok]).max(axis=0) + 1e-12
    rel = np.max(np.abs(g[ok] - go[ok]) / scale)
    _print("   adjoint: largest gradient error relative to each parameter's scale", rel)
    assert rel < 2e-3


def test_graph_replays_on_a_new_frame_pair(ctx, oracle):
    """A local step captured once as a hipGraph and replayed after the frame
    pair in its slot is replaced (same size: the slot keeps its buffers) gives
    the same results as an eager step on the new pair, and those match the
    oracle: the per-frame replay of an image sequence."""
    from staq_dic_amd import Batch, params, synth
    from staq_dic_amd._lib import IN_U0
    fa, coords, win = synth.config_plain(seed=1101)
    fb, _, _ = synth.config_plain(seed=1102)
    p6 = params(tol=1e-3)
    ctx.set_frame(fa.f, fa.g, fa.mask, slot=0)
    b = Batch(ctx, coords, win)
    ta = synth.truth_at(fa.field, coords)
    b.set_input(IN_U0, np.round(ta[:, 4:6]).reshape(-1))
    b.icgn6(p6)
    b.reorder()
    b.graph_capture(p6, 6)
    ctx.set_frame(fb.f, fb.g, fb.mask, slot=0)
    tb = synth.truth_at(fb.field, coords)
    U0 = np.round(tb[:, 4:6]).reshape(-1)
    b.set_input(IN_U0, U0)
    b.graph_launch()
    Ug, Fg, Hg, sg = b.outputs()
    b.icgn6(p6)
    Ue, Fe, He, se = b.outputs()
    np.testing.assert_array_equal(sg, se)
    np.testing.assert_array_equal(Ug, Ue)
    np.testing.assert_array_equal(Fg, Fe)
    b.close()
    of = oracle.Frame(fb.f, fb.g, fb.mask)
    Uo, Fo, Ho, so = oracle.local_icgn(of, U0, coords, win, tol=1e-3)
    _cmp_icgn(sg, so, Ug, Uo, Fg, Fo, name="graph on a new pair")
