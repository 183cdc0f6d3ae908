"""Bad-subset refill (thin-plate RBF per mask component, then scattered
interpolation of leftovers) against closed forms and invariants."""
import numpy as np

from staq_dic_amd.refill import refill_bad, thinplate_eval, thinplate_fit


def test_thinplate_interpolates_and_reproduces_linear():
    rng = np.random.default_rng(0)
    c = rng.uniform(0, 100, (60, 2)).round()
    c = np.unique(c, axis=0)
    y = np.stack([np.sin(c[:, 0] / 17.0) + c[:, 1] / 50.0, 0.3 + 0.01 * c[:, 0] - 0.02 * c[:, 1]], 1)
    coeff = thinplate_fit(c, y)
    np.testing.assert_allclose(thinplate_eval(c, coeff, c), y, atol=1e-8)
    # the linear column is reproduced everywhere with zero radial weights
    p = rng.uniform(0, 100, (40, 2))
    np.testing.assert_allclose(thinplate_eval(c, coeff, p)[:, 1],
                               0.3 + 0.01 * p[:, 0] - 0.02 * p[:, 1], atol=1e-8)
    np.testing.assert_allclose(coeff[:len(c), 1], 0.0, atol=1e-9)
    # the radial part: phi(r) = r^2 log(r + 1) (rbfcreate.m:195)
    c2 = np.array([[0.0, 0.0], [3.0, 4.0], [6.0, 0.0], [0.0, 8.0]])
    y2 = np.array([[1.0], [0.0], [0.0], [0.0]])
    w = thinplate_fit(c2, y2)
    r = 5.0
    val = w[0, 0] * 25 * np.log(r + 1) + w[1, 0] * 0 + w[2, 0] * 25 * np.log(6) + \
        w[3, 0] * (9 + 16) * np.log(6) + w[4, 0] + 3 * w[5, 0] + 4 * w[6, 0]
    np.testing.assert_allclose(val, 0.0, atol=1e-10)


def test_refill_recovers_linear_fields_per_component(ref_frames):
    """Two mask components carrying different linear fields: each component is
    refilled from its own good nodes only (8-connected components of the mask,
    LocalICGNQuadtree.m:189-195)."""
    c = ref_frames["coords"]
    N = c.shape[0]
    mask = np.ones((400, 400))
    mask[199:201, :] = 0  # split the image into two components along x = 200
    left = np.round(c[:, 0]) < 200
    U = np.where(left[:, None], np.stack([0.5 + 0.01 * c[:, 0], -0.2 + 0.02 * c[:, 1]], 1),
                 np.stack([-1.0 - 0.03 * c[:, 1], 2.0 + 0.005 * c[:, 0]], 1)).ravel()
    F = np.where(left[:, None], np.tile([0.01, 0.0, 0.0, 0.02], (N, 1)),
                 np.tile([0.0, 0.005, -0.03, 0.0], (N, 1))).ravel()
    rng = np.random.default_rng(3)
    bad = rng.random(N) < 0.15
    on_line = np.isin(np.round(c[:, 0]), [199, 200])
    Ub = U.copy().reshape(N, 2)
    Fb = F.copy().reshape(N, 4)
    Ub[bad] = np.nan
    Fb[bad] = np.nan
    Uf, Ff = refill_bad(Ub.ravel(), Fb.ravel(), c, mask)
    ok = ~on_line
    np.testing.assert_allclose(Uf.reshape(N, 2)[ok], U.reshape(N, 2)[ok], atol=1e-6)
    np.testing.assert_allclose(Ff.reshape(N, 4)[ok], F.reshape(N, 4)[ok], atol=1e-8)
    assert np.isfinite(Uf).all() and np.isfinite(Ff).all()


def test_refill_leftovers_outside_every_component(ref_frames):
    """Nodes whose rounded position is outside the mask keep NaN after the RBF
    stage and are then interpolated from the non-NaN nodes (:313-330); for a
    linear field the interpolation inside the node hull is exact."""
    c = ref_frames["coords"]
    N = c.shape[0]
    mask = ref_frames["mask"]
    U = np.stack([0.1 + 0.002 * c[:, 0] - 0.001 * c[:, 1], -0.3 + 0.004 * c[:, 1]], 1).ravel()
    rc = np.round(c).astype(int)
    out = mask[rc[:, 0] - 1, rc[:, 1] - 1] == 0
    Ub = U.copy().reshape(N, 2)
    Ub[out] = np.nan
    Uf, _ = refill_bad(Ub.ravel(), None, c, mask)
    assert np.isfinite(Uf).all()
    np.testing.assert_allclose(Uf.reshape(N, 2)[~out], U.reshape(N, 2)[~out], atol=1e-6)
    from scipy.spatial import Delaunay
    hull = Delaunay(c[~out]).find_simplex(c[out]) >= 0
    if hull.any():
        np.testing.assert_allclose(Uf.reshape(N, 2)[out][hull], U.reshape(N, 2)[out][hull],
                                   atol=1e-9)


def test_f_takes_the_centres_of_u():
    """Every fit of the refill and of the AL loop's refit takes the nodes whose
    u is finite as centres, F's included (LocalICGNQuadtree.m:180-249,
    main_ALDIC_Quadtree_inc.m:166-222): an F value at a node whose u is NaN is
    replaced by the fit through the others, and an F of NaN at a node whose u
    is finite makes that fit NaN (MATLAB's backslash returns NaN without an
    error), so only the scattered-interpolation stage fills F there."""
    from staq_dic_amd.aldic import rbf_refit
    rng = np.random.default_rng(4)
    c = np.unique(rng.uniform(10, 190, (80, 2)).round(), axis=0)
    N = c.shape[0]
    lin = lambda b0, bx, by: b0 + bx * c[:, 0] + by * c[:, 1]
    U = np.stack([lin(0.3, 0.01, -0.02), lin(-1.0, 0.005, 0.01)], 1)
    F = np.stack([lin(0.01, 1e-4, 0), lin(0.02, 0, 1e-4), lin(-0.01, 0, 0), lin(0.0, 2e-4, 0)], 1)
    k = 7
    U[k] = np.nan
    F_planted = F.copy()
    F_planted[k] = [5.0, 5.0, 5.0, 5.0]  # finite F where u is NaN
    Uo, Fo = rbf_refit(U.ravel(), F_planted.ravel(), c)
    # the F fit excludes node k (u is NaN there): the linear field comes back
    np.testing.assert_allclose(Fo.reshape(N, 4)[k], F[k], atol=1e-8)
    np.testing.assert_allclose(Uo.reshape(N, 2)[k], [0.3 + 0.01 * c[k, 0] - 0.02 * c[k, 1],
                                                     -1.0 + 0.005 * c[k, 0] + 0.01 * c[k, 1]],
                               atol=1e-8)
    # F of NaN at a node whose u is finite: the F fit is NaN everywhere
    F_nan = F.copy()
    F_nan[3] = np.nan
    _, Fo2 = rbf_refit(U.ravel(), F_nan.ravel(), c)
    assert np.all(np.isnan(Fo2))
    # the refill (one mask component): F at the node whose u is NaN is
    # refitted through the other nodes, like u
    mask = np.ones((200, 200))
    U2, F2 = refill_bad(U.ravel(), F_planted.ravel(), c, mask)
    np.testing.assert_allclose(F2.reshape(N, 4)[k], F[k], atol=1e-8)
    np.testing.assert_allclose(U2.reshape(N, 2)[k], Uo.reshape(N, 2)[k], atol=1e-8)


def test_local_support_fill(ref_frames, monkeypatch):
    """The local-support thin-plate fill that replaces the dense solve past
    DENSE_MAX centres: linear fields exactly; on the reference's 1434-node
    mesh with its stored answer and 100 nodes set to NaN, the refill with the
    local form (forced by lowering DENSE_MAX) stays within 5e-3 px (u, v) and
    5e-4 (F) of the dense form at the refilled nodes (2e-3 px and 1.5e-4 at
    most, 2e-5 px median) and leaves the others unchanged."""
    import staq_dic_amd.refill as rf
    from staq_dic_amd.refill import thinplate_fill_local
    rng = np.random.default_rng(6)
    c = np.unique(rng.uniform(0, 300, (500, 2)).round(), axis=0)
    lin = np.stack([0.2 + 0.01 * c[:, 0] - 0.03 * c[:, 1], -1 + 0.002 * c[:, 1]], 1)
    t = rng.uniform(20, 280, (60, 2))
    got = thinplate_fill_local(c, lin, t)
    np.testing.assert_allclose(got, np.stack([0.2 + 0.01 * t[:, 0] - 0.03 * t[:, 1],
                                              -1 + 0.002 * t[:, 1]], 1), atol=1e-8)
    coords = ref_frames["coords"]
    N = coords.shape[0]
    U = ref_frames["U1"].copy()
    F = ref_frames["F1"].copy()
    ok = np.isfinite(U[0::2]) & np.isfinite(F[0::4])
    idx = rng.choice(np.flatnonzero(ok), 100, replace=False)
    U[2 * idx] = np.nan
    U[2 * idx + 1] = np.nan
    for k in range(4):
        F[4 * idx + k] = np.nan
    mask = ref_frames["mask"]
    Ud, Fd = rf.refill_bad(U, F, coords, mask)
    monkeypatch.setattr(rf, "DENSE_MAX", 10)
    Ul, Fl = rf.refill_bad(U, F, coords, mask)
    du = np.abs(Ul.reshape(N, 2) - Ud.reshape(N, 2))
    dF = np.abs(Fl.reshape(N, 4) - Fd.reshape(N, 4))
    assert np.nanmax(du[idx]) < 5e-3 and np.nanmax(dF[idx]) < 5e-4
    rest = np.setdiff1d(np.flatnonzero(ok), idx)
    np.testing.assert_allclose(Ul.reshape(N, 2)[rest], ref_frames["U1"].reshape(N, 2)[rest],
                               atol=1e-6)
