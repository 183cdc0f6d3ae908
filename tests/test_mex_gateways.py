"""The MEX gateways (staq_dic_amd/mex) built against the test stand-in MEX API
(tests/mexshim) and called the way MATLAB calls them: argument checks and
error texts on the CPU (ba_interp2.cpp:259-335 for ba_interp2), results
against the library's own entry points on the GPU."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "mexshim"))
import build_gateways  # noqa: E402

DOUBLE, SINGLE = 6, 7


class MexError(RuntimeError):
    pass


def _global_library():
    """The gateways leave the staq_* symbols to the process (a MATLAB build
    links -lstaqdic): load libstaqdic.so globally first."""
    from staq_dic_amd import _lib
    _lib.load()
    C.CDLL(_lib.LIB_PATH, mode=C.RTLD_GLOBAL)


class Gateway:
    """One gateway library: numpy arrays in, numpy arrays out (column-major,
    at least two dimensions, 1-D arrays as columns)."""

    def __init__(self, name):
        _global_library()
        so = build_gateways.path(name)
        if not os.path.exists(so):
            build_gateways.build(force=False)
        L = C.CDLL(so)
        vp = C.c_void_p
        L.fakemx_array.restype = vp
        L.fakemx_array.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_long), vp]
        for f, r in (("fakemx_class", C.c_int), ("fakemx_ndim", C.c_int), ("fakemx_numel", C.c_long)):
            getattr(L, f).restype = r
            getattr(L, f).argtypes = [vp]
        L.fakemx_dim.restype = C.c_long
        L.fakemx_dim.argtypes = [vp, C.c_int]
        L.fakemx_data.restype = vp
        L.fakemx_data.argtypes = [vp]
        L.fakemx_free.argtypes = [vp]
        L.fakemx_call.restype = C.c_int
        L.fakemx_call.argtypes = [C.c_int, C.POINTER(vp), C.c_int, C.POINTER(vp), C.c_char_p, C.c_int]
        self.L = L

    def _in(self, a):
        a = np.asarray(a)
        if a.dtype not in (np.float64, np.float32):
            a = a.astype(np.float64)
        cls = SINGLE if a.dtype == np.float32 else DOUBLE
        shape = a.shape if a.ndim >= 2 else ((a.size, 1) if a.ndim == 1 else (1, 1))
        buf = np.asfortranarray(a.reshape(shape, order="F"))
        dims = (C.c_long * len(shape))(*shape)
        return self.L.fakemx_array(cls, len(shape), dims, buf.ctypes.data)

    def _out(self, p):
        L = self.L
        nd = L.fakemx_ndim(p)
        dims = tuple(L.fakemx_dim(p, k) for k in range(nd))
        dt = np.float32 if L.fakemx_class(p) == SINGLE else np.float64
        n = L.fakemx_numel(p)
        raw = np.ctypeslib.as_array((C.c_char * (n * np.dtype(dt).itemsize)).from_address(L.fakemx_data(p)))
        return np.frombuffer(raw.tobytes(), dtype=dt).reshape(dims, order="F")

    def __call__(self, nlhs, *args):
        prhs = (C.c_void_p * max(len(args), 1))(*[self._in(a) for a in args])
        plhs = (C.c_void_p * max(nlhs, 1))()
        err = C.create_string_buffer(512)
        rc = self.L.fakemx_call(nlhs, plhs, len(args), prhs, err, 512)
        for k in range(len(args)):
            self.L.fakemx_free(prhs[k])
        if rc != 0:
            raise MexError(err.value.decode())
        out = []
        for k in range(max(nlhs, 1)):
            if plhs[k]:
                out.append(self._out(plhs[k]))
                self.L.fakemx_free(plhs[k])
        return out


@pytest.fixture(scope="module")
def gw():
    build_gateways.build(force=False)
    return {n: Gateway(n) for n in build_gateways.GATEWAYS}


def test_gateways_build_with_the_stand_in_api():
    out = build_gateways.build(force=True)
    _global_library()
    for name, so in out.items():
        L = C.CDLL(so)
        assert hasattr(L, "mexFunction") and hasattr(L, "fakemx_call"), name


def test_ba_interp2_argument_errors(gw):
    """The reference gateway's checks and messages, raised before any device
    work (ba_interp2.cpp:266-282, :321-334)."""
    g = gw["ba_interp2"]
    F = np.zeros((8, 9))
    X = np.ones((3, 4))
    with pytest.raises(MexError, match=r"Wrong number of input arguments for Z = ba_interp2\(F, X, Y, \[method\]\)"):
        g(1, F, X)
    with pytest.raises(MexError, match="Wrong number of output arguments"):
        g(2, F, X, X)
    with pytest.raises(MexError, match="only double or single arguments"):
        g(1, F.astype(np.float32), X, X)
    with pytest.raises(MexError, match="X, Y should have the same size"):
        g(1, F, X, np.ones((4, 3)))
    with pytest.raises(MexError, match="X, Y should have the same size"):
        g(1, F, X, np.ones((3, 4, 2)))


def test_local_step_and_search_argument_errors(gw):
    li, sp, se = gw["staq_local_icgn_mex"], gw["staq_subpb1_mex"], gw["staq_integer_search_mex"]
    coords = np.ones((5, 2))
    img = np.zeros((40, 40))
    with pytest.raises(MexError, match="7 or 8 inputs expected"):
        li(4, np.zeros(10), coords, 16, img, img)
    with pytest.raises(MexError, match="coordinatesFEM must be an N x 2"):
        li(4, np.zeros(10), np.ones((5, 3)), 16, img, img, img, 1e-3)
    with pytest.raises(MexError, match="U0 must hold 2N"):
        li(4, np.zeros(9), coords, 16, img, img, img, 1e-3)
    with pytest.raises(MexError, match="winsize must be a scalar or N-vector"):
        li(4, np.zeros(10), coords, np.array([16, 16]), img, img, img, 1e-3)
    with pytest.raises(MexError, match="at most 4 outputs"):
        li(5, np.zeros(10), coords, 16, img, img, img, 1e-3)
    with pytest.raises(MexError, match="FOld must hold 4N"):
        sp(2, np.zeros(10), np.zeros(19), np.zeros(10), coords, 16, img, img, img, 1e-3, 1e-3)
    with pytest.raises(MexError, match="winsize_List must be N x 2"):
        sp(2, np.zeros(10), np.zeros(20), np.zeros(10), coords, np.ones((5, 3)), img, img, img,
           1e-3, 1e-3)
    with pytest.raises(MexError, match="r must be a scalar or"):
        se(5, img, img, np.array([1, 2, 3]), [1, 40], [1, 40], 16, 8)
    with pytest.raises(MexError, match="gridx and gridy must be"):
        se(5, img, img, 4, [1, 40, 3], [1, 40], 16, 8)


@pytest.mark.gpu
def test_ba_interp2_gateway_matches_library(gw):
    """Z = ba_interp2(F, X, Y, 'cubic') through the gateway equals the
    library's own entry point, double and single, with channels: the output
    takes X's dimensions followed by F's third."""
    from staq_dic_amd import ba_interp2
    g = gw["ba_interp2"]
    rng = np.random.default_rng(7)
    F = rng.random((40, 30, 3))
    X = rng.uniform(-2, 33, (5, 7, 2))
    Y = rng.uniform(-2, 43, (5, 7, 2))
    (Z,) = g(1, F, X, Y, np.array([99.0]))
    assert Z.shape == (5, 7, 2, 3) and Z.dtype == np.float64
    ref = np.stack([ba_interp2(F[:, :, c], X, Y) for c in range(3)], axis=-1)
    np.testing.assert_array_equal(Z, ref)
    (Zs,) = g(1, F.astype(np.float32), X.astype(np.float32), Y.astype(np.float32))
    assert Zs.shape == (5, 7, 2, 3) and Zs.dtype == np.float32
    np.testing.assert_allclose(Zs, ref, rtol=0, atol=2e-5)
    (Z2,) = g(1, F[:, :, 0], X[:, :, 0], Y[:, :, 0])
    assert Z2.shape == (5, 7)
    np.testing.assert_array_equal(Z2, ref[:, :, 0, 0])


@pytest.mark.gpu
def test_local_step_gateways_match_library(gw, plain_cfg):
    """staq_local_icgn_mex and staq_subpb1_mex (MATLAB layouts in and out)
    equal the library's batch entry points on the same inputs."""
    from staq_dic_amd import local_icgn_raw, subpb1_raw, synth
    fr, coords, win = plain_cfg
    sel = np.arange(0, coords.shape[0], 4)
    coords = np.asfortranarray(coords[sel])
    N = coords.shape[0]
    truth = synth.truth_at(fr.field, coords)
    U0 = np.round(truth[:, 4:6]).reshape(-1)
    mask = np.ones_like(fr.f) if fr.mask is None else fr.mask
    U, F, Hp, st = gw["staq_local_icgn_mex"](4, U0, coords, 32.0, fr.f, fr.g, mask, 1e-3)
    Ur, Fr, Hr, sr = local_icgn_raw(U0, coords, np.full(N, 32), fr.f, fr.g, mask, tol=1e-3)
    assert U.shape == (2 * N, 1) and Hp.shape == (N, 21) and st.shape == (N, 1)
    np.testing.assert_array_equal(st[:, 0], sr)
    np.testing.assert_array_equal(U[:, 0], Ur)
    np.testing.assert_array_equal(F[:, 0], Fr)
    np.testing.assert_array_equal(Hp, Hr)
    vd = np.zeros(2 * N)
    U2, st2 = gw["staq_subpb1_mex"](2, Ur, Fr, vd, coords, np.full((N, 2), 32.0), fr.f, fr.g,
                                     mask, 1e-3, 1e-3)
    U2r, s2r = subpb1_raw(Ur, Fr, vd, coords, np.full(N, 32), np.full(N, 32), 1e-3, fr.f, fr.g,
                          mask, tol=1e-3)
    np.testing.assert_array_equal(U2[:, 0], U2r)
    np.testing.assert_array_equal(st2[:, 0], np.where(s2r == -1, 0, s2r))


@pytest.mark.gpu
def test_integer_search_gateway_matches_library(gw, plain_cfg):
    from staq_dic_amd import funIntegerSearch
    fr, _, _ = plain_cfg
    H, W = fr.f.shape
    x, y, u, v, cc = gw["staq_integer_search_mex"](5, fr.f, fr.g, 10.0, [1, H], [1, W], 32.0, 16.0)
    xr, yr, ur, vr, ccr = funIntegerSearch(fr.f, fr.g, 10, [1, H], [1, W], 32, 16)
    assert x.shape == xr.shape == (26, 26)
    for a, b in ((x, xr), (y, yr), (u, ur), (v, vr), (cc, ccr["max"])):
        np.testing.assert_array_equal(a, b)
    # clearing the MEX file runs its mexAtExit handler (the context is destroyed)
    gw["staq_integer_search_mex"].L.fakemx_clear()
