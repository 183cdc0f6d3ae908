"""The plain twins' refills (staq_dic_amd.refill_plain): LocalICGN.m:117-287
with funCompDefGrad2.m, and Subpb1.m:171-199.

isoutlier's moving rule is pinned by cases a plausible mistake fails: the
MAD scale 1.4826 (a point 3.5 unscaled MADs out is not an outlier, 5 is), the
shrunk end windows (a first element against one neighbour only), NaN left out.
The plane fit is pinned by exactness on piecewise-linear fields across a mask
gap (only the node's own 4-connected side enters its fit).  The in-loop F fill
of funCompDefGrad2.m:160-176 is not restated (parity unpinned; the module
docstring gives the reason)."""
import numpy as np
import pytest

from staq_dic_amd.refill_plain import (comp_def_grad2, isoutlier_movmedian, local_icgn_refill,
                                       subpb1_gridfit)


def test_isoutlier_scale_and_threshold():
    # whole-vector windows: median 0.5, MAD 1.0, scaled MAD 1.4826
    base = [-1.0, -1.0, 0.0, 1.0, 1.0]
    assert not isoutlier_movmedian(base + [4.0], 13).any()  # 3.5 MADs, 2.36 scaled
    tf = isoutlier_movmedian(base + [5.5], 13)              # 5 MADs, 3.37 scaled
    np.testing.assert_array_equal(tf, [False] * 5 + [True])


def test_isoutlier_shrinks_at_the_ends_and_skips_nan():
    a = np.array([10.0, 0, 0, 0, 0, 0, 0])
    # the first window is [10, 0] (median 5, MAD 5): no outlier there
    assert not isoutlier_movmedian(a, 3)[0]
    b = np.array([0.0, 0.1, -0.1, 0.05, 9.0, 0.0, np.nan, 0.02, -0.05])
    tf = isoutlier_movmedian(b, 5)
    assert tf[4] and not tf[6] and tf.sum() == 1
    # a NaN inside a window is left out of its statistics, not propagated
    c = np.array([0.0, 0.1, np.nan, -0.1, 0.05, 0.0, 0.02])
    assert not isoutlier_movmedian(c, 5).any()


def _grid(n=24, step=8.0, x0=40.0):
    g = x0 + step * np.arange(n)
    X, Y = np.meshgrid(g, g, indexing="ij")
    return np.stack([X.ravel(), Y.ravel()], axis=1)


def test_plane_fit_is_exact_on_each_side_of_a_mask_gap():
    H = W = 300
    c = _grid()
    mask = np.ones((H, W))
    gap = 130  # rows x = 131..132 (1-based) are masked out
    mask[gap:gap + 2, :] = 0
    side = c[:, 0] > gap + 1
    Ua = np.stack([0.3 + 0.01 * c[:, 0] - 0.02 * c[:, 1], -1.0 + 0.005 * c[:, 0] + 0.015 * c[:, 1]], 1)
    Ub = np.stack([2.0 - 0.03 * c[:, 0] + 0.01 * c[:, 1], 0.5 + 0.02 * c[:, 0] - 0.01 * c[:, 1]], 1)
    uv = np.where(side[:, None], Ub, Ua)
    U, F = comp_def_grad2(uv, c, 17.0, mask)
    np.testing.assert_allclose(U.reshape(-1, 2), uv, atol=1e-10)
    Fa = np.array([0.01, 0.005, -0.02, 0.015])
    Fb = np.array([-0.03, 0.02, 0.01, -0.01])
    np.testing.assert_allclose(F.reshape(-1, 4), np.where(side[:, None], Fb, Fa), atol=1e-12)
    # without the mask the fits near the gap mix both sides
    U2, _ = comp_def_grad2(uv, c, 17.0, None)
    assert np.abs(U2.reshape(-1, 2) - uv).max() > 1e-3


def test_plane_fit_few_neighbours_and_window_off_the_image():
    c = np.array([[50.0, 50.0], [58.0, 50.0], [50.0, 58.0], [150.0, 150.0], [5.0, 100.0]])
    uv = np.arange(10.0).reshape(5, 2)
    U, F = comp_def_grad2(uv, c, 17.0, np.ones((200, 200)))
    # three neighbours or fewer (every node here), or a window that leaves the
    # image (the last node): u kept, F NaN
    np.testing.assert_array_equal(U.reshape(-1, 2), uv)
    assert np.isnan(F).all()


def test_local_icgn_refill_fills_bad_nodes_from_their_nearest():
    H = W = 300
    c = _grid()
    N = c.shape[0]
    mask = np.ones((H, W))
    U = np.stack([0.1 + 0.01 * c[:, 0], 0.2 - 0.01 * c[:, 1]], 1)
    F = np.tile([0.01, 0.0, 0.0, -0.01], (N, 1))
    bad = np.array([30, 31, 200, 401])
    U[bad] = np.nan
    F[bad] = np.nan
    Uo, Fo = local_icgn_refill(U.ravel(), F.ravel(), c, mask, 8.0)
    Uo, Fo = Uo.reshape(N, 2), Fo.reshape(N, 4)
    good = np.setdiff1d(np.arange(N), bad)
    np.testing.assert_allclose(Uo[good], U[good], atol=1e-10)
    np.testing.assert_allclose(Fo, np.tile([0.01, 0.0, 0.0, -0.01], (N, 1)), atol=1e-12)
    for b in bad:
        d = np.hypot(*(c[good] - c[b]).T)
        near = good[d == d.min()].min()
        np.testing.assert_array_equal(Uo[b], Uo[near])


def test_subpb1_gridfit_constant_and_index_order():
    xs = 20.0 + 10.0 * np.arange(9)
    ys = 30.0 + 12.0 * np.arange(6)  # a non-square grid: a transposed read fails
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    c = np.stack([X.ravel(), Y.ravel()], axis=1)
    N = c.shape[0]
    U = np.tile([1.25, -0.5], N).astype(float)
    U[[2 * 7, 2 * 7 + 1, 2 * 20, 2 * 20 + 1]] = np.nan
    np.testing.assert_allclose(subpb1_gridfit(U, c), np.tile([1.25, -0.5], N), atol=1e-10)
    from staq_dic_amd.mgsearch import gridfit
    rng = np.random.default_rng(4)
    V = rng.normal(size=2 * N)
    out = subpb1_gridfit(V, c).reshape(N, 2)
    z, _, _ = gridfit(c[:, 0], c[:, 1], V[0::2], xs, ys)
    for k in (0, 5, 17, N - 1):
        ix = int(np.flatnonzero(xs == c[k, 0])[0])
        iy = int(np.flatnonzero(ys == c[k, 1])[0])
        assert out[k, 0] == z[iy, ix]
