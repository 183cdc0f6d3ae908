"""Initial guess on the uniform mesh: regularizeNd, Init and the drivers'
search -> RBF -> regularizeNd -> MeshSetUp -> Init chain.

regularizeNd is pinned by what the mathematics fixes: a plane is reproduced
at any smoothness (the interpolation reproduces it and its second
differences vanish); a very large smoothness leaves the least-squares
bilinear fit (the null space of the second differences); the result does not
change when all coordinates are scaled (the (range)^2 factor against the
1/h^2 of the second derivative) or when every point is given twice (the
sqrt(n / rows) factor); a vanishing smoothness returns data given at the
nodes.  Init: the node order equals MeshSetUp's, and the index-1 passes
remove a spike from a linear field, which the springs then restore exactly
(the border aside, which every pass clears and the springs refill).
On the GPU the chain runs on the reference's own circular-hole frames, and the
local step started from its result lands on the reference's stored answer.
"""
import numpy as np
import pytest

from staq_dic_amd.initguess import Init, regularizeNd
from staq_dic_amd.mesh import MeshSetUp


def _scatter(seed, n=400):
    rng = np.random.default_rng(seed)
    xg = np.array([0.0, 1.5, 3.0, 5.0, 6.0, 8.5, 10.0])
    yg = np.array([-2.0, 0.0, 1.0, 4.0, 7.0])
    x = np.stack([rng.uniform(0, 10, n), rng.uniform(-2, 7, n)], axis=1)
    return x, xg, yg


@pytest.mark.parametrize("method", ["linear", "cubic"])
def test_regularize_plane_at_any_smoothness(method):
    x, xg, yg = _scatter(1)
    plane = lambda a, b: 0.7 - 0.3 * a + 0.12 * b
    for s in (1e-3, 0.5, 30.0):
        z = regularizeNd(x, plane(x[:, 0], x[:, 1]), [xg, yg], s, method)
        X, Y = np.meshgrid(xg, yg, indexing="ij")
        np.testing.assert_allclose(z, plane(X, Y), atol=1e-7)


def test_regularize_large_smoothness_is_the_bilinear_fit():
    x, xg, yg = _scatter(2)
    rng = np.random.default_rng(3)
    y = np.sin(x[:, 0]) + 0.2 * x[:, 1] ** 2 + 0.05 * rng.normal(size=x.shape[0])
    # the distance to the null space falls as 1 / smoothness^2 (5e-3 at 1,
    # 5e-5 at 10); far beyond 100 the normal equations lose the digits
    z = regularizeNd(x, y, [xg, yg], 100.0)
    B = np.stack([np.ones(x.shape[0]), x[:, 0], x[:, 1], x[:, 0] * x[:, 1]], axis=1)
    c = np.linalg.lstsq(B, y, rcond=None)[0]
    X, Y = np.meshgrid(xg, yg, indexing="ij")
    np.testing.assert_allclose(z, c[0] + c[1] * X + c[2] * Y + c[3] * X * Y, atol=5e-6)


def test_regularize_scale_invariances():
    x, xg, yg = _scatter(4, n=150)
    y = np.cos(0.6 * x[:, 0]) * np.exp(0.1 * x[:, 1])
    z = regularizeNd(x, y, [xg, yg], 0.05)
    # scaling every coordinate: the second derivative's 1/h^2 against range^2
    z2 = regularizeNd(7.0 * x, y, [7.0 * xg, 7.0 * yg], 0.05)
    np.testing.assert_allclose(z2, z, rtol=1e-9, atol=1e-11)
    # each point twice: the fidelity doubles, and so does sqrt(n / rows)^2
    z3 = regularizeNd(np.concatenate([x, x]), np.concatenate([y, y]), [xg, yg], 0.05)
    np.testing.assert_allclose(z3, z, rtol=1e-9, atol=1e-11)
    # a different smoothness per dimension differs from either scalar
    za = regularizeNd(x, y, [xg, yg], [0.05, 2.0])
    assert np.abs(za - z).max() > 1e-3


def test_regularize_nodes_and_nearest():
    xg = np.array([0.0, 2.0, 3.0, 6.0])
    yg = np.array([1.0, 2.0, 4.0, 5.0])
    X, Y = np.meshgrid(xg, yg, indexing="ij")
    vals = np.arange(X.size, dtype=float).reshape(X.shape) ** 1.3
    pts = np.stack([X.ravel(order="F"), Y.ravel(order="F")], axis=1)
    for m in ("linear", "nearest", "cubic"):
        z = regularizeNd(pts, vals.ravel(order="F"), [xg, yg], 1e-12, m)
        np.testing.assert_allclose(z, vals, atol=1e-8)
    # nearest: a midpoint goes to the upper node (MATLAB's round)
    p = np.array([[1.0, 1.0], [2.5, 3.0]])
    z = regularizeNd(np.concatenate([pts, p]), np.concatenate([vals.ravel(order="F"), [100.0, 100.0]]),
                     [xg, yg], 1e-12, "nearest")
    assert z[1, 0] > vals[1, 0] + 1 and abs(z[0, 0] - vals[0, 0]) < 1e-6
    with pytest.raises(ValueError):
        regularizeNd(np.array([[7.0, 1.0]]), np.array([1.0]), [xg, yg])


def test_init_order_and_spike():
    xn = np.arange(20.0, 141.0, 16.0)
    yn = np.arange(20.0, 109.0, 16.0)
    X, Y = np.meshgrid(xn, yn, indexing="ij")
    coords, _ = MeshSetUp(xn, yn)
    U0 = Init(X, Y, None, X.T, Y.T, 0)
    np.testing.assert_array_equal(U0[0::2], coords[:, 0])
    np.testing.assert_array_equal(U0[1::2], coords[:, 1])
    u = 0.5 + 0.01 * X - 0.02 * Y
    v = -1.0 + 0.03 * X + 0.005 * Y
    uh = u.copy()
    uh[3, 2] = np.nan
    np.testing.assert_allclose(Init(uh, v, None, X.T, Y.T, 0)[0::2], u.ravel(), atol=1e-9)
    # index 1: the spike and its 8 neighbours are cleared and the springs
    # restore the linear field there; every pass also clears the border, which
    # the springs refill flat, and the later passes then clear the nodes next
    # to two of the corners, where that flat border departs most
    xn = np.arange(20.0, 20.0 + 16 * 14, 16.0)
    yn = np.arange(20.0, 20.0 + 16 * 12, 16.0)
    X, Y = np.meshgrid(xn, yn, indexing="ij")
    u = 0.5 + 0.01 * X - 0.02 * Y
    v = -1.0 + 0.03 * X + 0.005 * Y
    us = u.copy()
    us[6, 5] += 5.0
    out = Init(us, v, None, X.T, Y.T, 1)
    ou = out[0::2].reshape(u.shape)
    ov = out[1::2].reshape(u.shape)
    np.testing.assert_allclose(ou[3:-3, 3:-3], u[3:-3, 3:-3], atol=1e-8)
    np.testing.assert_allclose(ov[3:-3, 3:-3], v[3:-3, 3:-3], atol=1e-8)
    np.testing.assert_allclose(ou[1:-1, 1:-1][3:9, 2:8], u[1:-1, 1:-1][3:9, 2:8], atol=1e-8)
    assert np.abs(ou[0, 1:-1] - u[1, 1:-1]).max() < 1.0


@pytest.mark.gpu
def test_front_half_on_reference_frames(ctx, ref_frames):
    """The driver's chain on the reference's own frames (search g -> f,
    thin-plate RBF, regularizeNd, MeshSetUp, Init, mask NaNs), then the
    quadtree mesh and the transfer of U0 onto it, then LocalICGNQuadtree:
    the result lands on the reference's stored answer, as it does from the
    rounded stored answer."""
    from staq_dic_amd import LocalICGNQuadtree, classify_local
    from staq_dic_amd.initguess import uniform_initial_guess
    from staq_dic_amd.mesh import GenerateQuadtreeMesh, transfer_initial_guess
    r = ref_frames
    H, W = r["f"].shape
    P = {"winsize": 32, "winstepsize": 16, "SizeOfFFTSearchRegion": 6,
         "gridxyROIRange": {"gridx": [1, H], "gridy": [1, W]}, "ImgRefMask": r["mask"]}
    print("uniform initial guess on the reference frames", flush=True)
    P2, mesh0, U0, (x0t, y0t, u, v), cc = uniform_initial_guess(r["f"], r["g"], P, r["mask"],
                                                               r["mask"], ctx=ctx)
    c0 = mesh0["coordinatesFEM"]
    assert c0.shape == (529, 2) and U0.shape == (1058,)
    np.testing.assert_array_equal(c0, r["coords"][:529])
    hole = r["mask"][c0[:, 0].astype(int) - 1, c0[:, 1].astype(int) - 1] == 0
    assert np.isnan(U0[0::2][hole]).all() and np.isfinite(U0[0::2][~hole]).all()
    # off the hole the smoothed search result is within a tenth of a pixel of
    # the reference's stored answer at its first 529 (uniform) nodes
    U1 = r["U1"][:1058]
    d0 = np.hypot(U0[0::2] - U1[0::2], U0[1::2] - U1[1::2])[~hole]
    assert np.median(d0) < 0.1 and np.percentile(d0, 90) < 0.25, np.percentile(d0, [50, 90])
    c, E, irr, mark = GenerateQuadtreeMesh(c0, mesh0["elementsFEM"], r["mask"], 16, 32)
    np.testing.assert_array_equal(c, r["coords"])
    Uq = transfer_initial_guess(U0, c0, c, r["mask"])
    print("LocalICGNQuadtree from the transferred guess", flush=True)
    DICpara = {"winsize": r["winsize"], "winsize_List": r["winsize_List"]}
    Uf, F, Hp, t, conv, nbad, _ = LocalICGNQuadtree(Uq, c, {"ImgRefMask": r["mask"]}, r["f"],
                                                    r["g"], DICpara, "GaussNewton", 1e-2, ctx=ctx)
    bad, _ = classify_local(conv)
    d = np.hypot(Uf[0::2] - r["U1"][0::2], Uf[1::2] - r["U1"][1::2])
    assert np.median(d[~bad]) < 5e-3 and np.percentile(d[~bad], 90) < 0.06


@pytest.mark.gpu
def test_al_frame_on_reference_frames(ctx, ref_frames):
    """Sections 4-6 of the incremental driver on the reference's own frames
    and mesh (local 6-DOF step, thin-plate refit, beta tuning, global step
    and FE nodal strain, two rounds of the 2-DOF step): the result stays
    close to the reference's stored post-ADMM answer, which its v2 driver
    reached with the RBF nodal strain instead of the FE one."""
    from staq_dic_amd.aldic import ALDICQuadtreeFrame
    r = ref_frames
    print("AL frame on the reference frames", flush=True)
    out = ALDICQuadtreeFrame(np.round(r["U1"]), r["mesh"], {"ImgRefMask": r["mask"]}, r["f"],
                             r["g"], {"winsize": 32, "winstepsize": 16}, tol=1e-2, ctx=ctx)
    U2 = out["USubpb2"]
    assert np.isfinite(U2).all() and np.isfinite(out["FSubpb2"]).all()
    assert out["ALSolveStep"] == 3
    bl = np.array([1e-3, 1e-2, 1e-1]) * 256 * 1e-3
    assert bl[0] <= out["beta"] <= bl[-1]
    d = np.hypot(U2[0::2] - r["U1"][0::2], U2[1::2] - r["U1"][1::2])
    assert np.median(d) < 0.05 and np.percentile(d, 90) < 0.15, np.percentile(d, [50, 90])


@pytest.mark.gpu
def test_front_half_with_multigrid_search(ctx, ref_frames):
    """The same chain with the multigrid search (InitFFTSearchMethod 0, the
    incremental driver's choice): the uniform mesh and the hole NaNs as with
    the whole-field search, and a guess within a tenth of a pixel (median) of
    the stored answer off the hole, which the local step then takes to the
    stored answer."""
    from staq_dic_amd import LocalICGNQuadtree, classify_local
    from staq_dic_amd.initguess import uniform_initial_guess
    from staq_dic_amd.mesh import GenerateQuadtreeMesh, transfer_initial_guess
    r = ref_frames
    H, W = r["f"].shape
    P = {"winsize": 32, "winstepsize": 16, "InitFFTSearchMethod": 0,
         "gridxyROIRange": {"gridx": [1, H], "gridy": [1, W]}, "ImgRefMask": r["mask"]}
    print("uniform initial guess with the multigrid search", flush=True)
    P2, mesh0, U0, _, cc = uniform_initial_guess(r["f"], r["g"], P, r["mask"], r["mask"], ctx=ctx)
    c0 = mesh0["coordinatesFEM"]
    np.testing.assert_array_equal(c0, r["coords"][:529])
    hole = r["mask"][c0[:, 0].astype(int) - 1, c0[:, 1].astype(int) - 1] == 0
    assert np.isnan(U0[0::2][hole]).all() and np.isfinite(U0[0::2][~hole]).all()
    U1 = r["U1"][:1058]
    d0 = np.hypot(U0[0::2] - U1[0::2], U0[1::2] - U1[1::2])[~hole]
    print("guess distance percentiles 50/90/100:", np.percentile(d0, [50, 90, 100]), flush=True)
    assert np.median(d0) < 0.1 and np.percentile(d0, 90) < 0.25
    c, E, irr, mark = GenerateQuadtreeMesh(c0, mesh0["elementsFEM"], r["mask"], 16, 32)
    Uq = transfer_initial_guess(U0, c0, c, r["mask"])
    DICpara = {"winsize": r["winsize"], "winsize_List": r["winsize_List"]}
    Uf, F, Hp, t, conv, nbad, _ = LocalICGNQuadtree(Uq, c, {"ImgRefMask": r["mask"]}, r["f"],
                                                    r["g"], DICpara, "GaussNewton", 1e-2, ctx=ctx)
    bad, _ = classify_local(conv)
    d = np.hypot(Uf[0::2] - r["U1"][0::2], Uf[1::2] - r["U1"][1::2])
    assert np.median(d[~bad]) < 5e-3 and np.percentile(d[~bad], 90) < 0.06


def test_mask_nan_nodes_try_catch_rule():
    """~fMask(x,y) || ~gMask(floor(x+u), floor(y+v)) inside a try/catch
    (main_ALDIC_Quadtree_v2.m:133-143), node by node on a hand-made case."""
    from staq_dic_amd.initguess import mask_nan_nodes
    fm = np.ones((6, 6), dtype=bool)
    gm = np.ones((6, 6), dtype=bool)
    fm[1, 1] = False          # node (2, 2) sits on a hole of f
    gm[3, 4] = False          # g's hole at (4, 5)
    X = np.array([[2.0, 2.0, 2.0], [4.0, 4.0, 4.0], [2.5, 6.0, 7.0]])
    Y = np.array([[2.0, 3.0, 4.0], [3.0, 4.0, 5.0], [2.0, 6.0, 2.0]])
    u = np.array([[0.0, 0.0, 2.2], [0.9, 0.0, 0.3], [0.0, 1.0, 0.0]])
    v = np.array([[0.0, 0.0, 1.1], [2.5, np.nan, 0.0], [0.0, 0.0, 0.0]])
    hole = mask_nan_nodes(X, Y, u, v, fm, gm)
    expect = np.array([
        [True, False, True],     # f hole; clear; g(floor 4.2, floor 5.1) = (4, 5) hole
        [True, False, True],     # g(4, 5) hole; NaN index skipped; g(4, 5) hole
        [False, False, False],   # x not an integer: skipped; g(7, 6) outside: skipped; x outside f: skipped
    ])
    np.testing.assert_array_equal(hole, expect)
    np.testing.assert_array_equal(mask_nan_nodes(X, Y, u, v, fm, None),
                                  np.array([[True, False, False], [False] * 3, [False] * 3]))


@pytest.mark.gpu
def test_al_frame_with_resident_local_steps(ctx, ref_frames):
    """The same AL frame with the local steps on a device-resident batch
    (captured graphs): the result equals the one-shot entry points' up to the
    FP32 summation order of the kernel forms, and stays as close to the
    stored answer."""
    from staq_dic_amd.aldic import ALDICQuadtreeFrame, ResidentLocalSteps
    r = ref_frames
    args = (np.round(r["U1"]), r["mesh"], {"ImgRefMask": r["mask"]}, r["f"], r["g"],
            {"winsize": 32, "winstepsize": 16})
    one = ALDICQuadtreeFrame(*args, tol=1e-2, ctx=ctx)
    res = ResidentLocalSteps(ctx, r["coords"], 32, r["f"], r["g"], r["mask"])
    try:
        out = ALDICQuadtreeFrame(*args, tol=1e-2, ctx=ctx, resident=res)
    finally:
        res.close()
    U2, U2o = out["USubpb2"], one["USubpb2"]
    d = np.hypot(U2[0::2] - U2o[0::2], U2[1::2] - U2o[1::2])
    print(f"resident vs one-shot local steps: median {np.median(d):.2e} px, max {d.max():.2e} px",
          flush=True)
    assert np.median(d) < 1e-4 and d.max() < 1e-2
    ds = np.hypot(U2[0::2] - r["U1"][0::2], U2[1::2] - r["U1"][1::2])
    assert np.median(ds) < 0.05 and np.percentile(ds, 90) < 0.15
