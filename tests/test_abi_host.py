"""CPU tests: the C ABI library loads and exports every declared symbol, the
Python binding covers the header, and the host-side bookkeeping mirrors the
reference (bad-subset sets, hole nodes, parameter defaults).  No GPU calls."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "staq_dic.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\**\s*(staq_\w+)\s*\(", text, flags=re.M)))


@pytest.fixture(scope="module")
def lib():
    from staq_dic_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        from staq_dic_amd import build
        build.build()
    return _lib.load()


def test_header_declares_the_path():
    names = _declared()
    for must in ("staq_ctx_create", "staq_set_frame", "staq_ba_interp2", "staq_img_gradient",
                 "staq_integer_search", "staq_local_icgn", "staq_subpb1", "staq_znssd_fwd",
                 "staq_znssd_bwd", "staq_batch_create", "staq_batch_icgn6", "staq_batch_icgn2",
                 "staq_batch_graph_capture", "staq_batch_autotune"):
        assert must in names


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header(lib):
    from staq_dic_amd import _lib
    assert sorted(_lib.SIGNATURES) == _declared()


def test_params_default_host_only(lib):
    from staq_dic_amd._lib import StaqParams
    p = StaqParams()
    lib.staq_params_default(C.byref(p))
    assert p.tol == 1e-2 and p.mu == 1e-3  # LocalICGNQuadtree / Subpb1 driver defaults
    assert p.max_iter in (0, 100) and p.plain == 0 and p.block_size == 0 and p.persistent == 0


def test_null_handles_are_rejected(lib):
    assert lib.staq_batch_icgn6(None, None) != 0
    assert lib.staq_batch_reorder(None) != 0
    assert lib.staq_batch_graph_launch(None) != 0


# ---- bad-subset classification (host logic around the kernels) ----------

def _reference_local_bad(st, plain):
    """LocalICGNQuadtree.m:151-169 / LocalICGN.m:117-131 written as set algebra."""
    n = st.size
    row1 = {i for i in range(n) if st[i] < 0}
    row2 = {i for i in range(n) if st[i] > 99}
    row3 = {i for i in range(n) if st[i] == 102}
    bad = row1 | row2
    nbad = len(bad) - len(row3)
    good = sorted(set(range(n)) - bad)
    g = st[good].astype(float)
    mean, std = g.mean(), g.std(ddof=1)
    thr = mean + 0.15 * std if plain else max(mean + std, 6.0)
    bad |= {i for i in range(n) if st[i] > thr}
    return bad, nbad


@pytest.mark.parametrize("plain", [False, True])
def test_classify_local(plain):
    from staq_dic_amd.api import classify_local
    rng = np.random.default_rng(3)
    for _ in range(20):
        st = rng.choice([-1, 1, 2, 3, 4, 5, 6, 7, 9, 12, 30, 101, 102, 103], size=200)
        bad, nbad = classify_local(st, plain=plain)
        rb, rn = _reference_local_bad(st, plain)
        assert set(np.flatnonzero(bad).tolist()) == rb and nbad == rn


def test_classify_subpb1():
    from staq_dic_amd.api import classify_subpb1
    st = np.array([0, 1, 2, 3, 2, 1, 50, 101, 102, 2, 3, 1])
    bad, nbad = classify_subpb1(st)
    # Subpb1Quadtree.m:173-186: <0 or >99 is bad, 102 not counted; then
    # > max(mean + 0.25 std, 10) over the rest
    good = st[[0, 1, 2, 3, 4, 5, 6, 9, 10, 11]].astype(float)
    thr = max(good.mean() + 0.25 * good.std(ddof=1), 10)
    expect = (st > 99) | (st > thr)
    np.testing.assert_array_equal(bad, expect)
    assert nbad == 1
    bad_p, n_p = classify_subpb1(st, plain=True)  # Subpb1.m:167-169: no threshold step
    np.testing.assert_array_equal(bad_p, st > 99)
    assert n_p == 2


def test_hole_nodes():
    from staq_dic_amd.api import _hole_nodes
    f = np.ones((64, 64))
    mask = np.ones((64, 64))
    mask[30:34, 30:34] = 0
    coords = np.array([[32.0, 32.0], [10.0, 10.0], [3.0, 40.0], [50.0, 50.0]])
    win = np.array([16, 8, 16, 32])
    # node 1 covers the hole, node 2 is clear, node 3 leaves the image,
    # node 4's 33 x 33 window (34..66) leaves the 64-pixel image too
    np.testing.assert_array_equal(_hole_nodes(coords, win, f, mask), [1, 3, 4])


def test_bench_gpus_flag_exits_cleanly():
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 2 and "one GPU" in r.stdout


def test_bad_arguments_without_a_device(lib):
    """Calls with a null context or batch, or impossible sizes, return
    STAQ_EINVAL (-1) before any device work; the host-only interpolation
    entry point rejects too few sites and a query triangle out of range."""
    import ctypes as C

    import numpy as np
    L = lib
    null = C.c_void_p()
    assert L.staq_local_icgn(null, None, 0, None, None, None, None, None, None, None, None) == -1
    assert L.staq_batch_icgn6(null, None) == -1
    assert L.staq_batch_icgn2(null, None) == -1
    assert L.staq_batch_graph_launch(null) == -1
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
    sites = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]])
    tri = np.array([[0, 1, 2]], dtype=np.int32)
    nbr = np.full((1, 3), -1, dtype=np.int32)
    ptr = np.array([0, 2, 4, 6], dtype=np.int64)
    adj = np.array([1, 2, 0, 2, 0, 1], dtype=np.int32)
    vals = np.array([1.0, 2.0, 3.0])
    q = np.array([[0.25, 0.25]])
    out = np.zeros(1)
    lp = ptr.ctypes.data_as(C.POINTER(C.c_long))
    assert L.staq_natural_interp(2, dp(sites), 1, ip(tri), ip(nbr), lp, ip(adj), 1, dp(vals), 1,
                                 dp(q), ip(np.array([0], dtype=np.int32)), dp(out)) == -1
    assert L.staq_natural_interp(3, dp(sites), 1, ip(tri), ip(nbr), lp, ip(adj), 1, dp(vals), 1,
                                 dp(q), ip(np.array([5], dtype=np.int32)), dp(out)) == -1
    # a lone triangle: the call succeeds and reproduces the linear field
    # through its corners (Sibson's coordinates have linear precision)
    assert L.staq_natural_interp(3, dp(sites), 1, ip(tri), ip(nbr), lp, ip(adj), 1, dp(vals), 1,
                                 dp(q), ip(np.array([0], dtype=np.int32)), dp(out)) == 0
    assert abs(out[0] - (1.0 + 0.25 + 2 * 0.25)) < 1e-9
    # the weights form: a bad triangle index and too small a cap are
    # rejected (the count is still reported); with room, the weights are the
    # barycentric ones and reproduce the value above
    idx = np.zeros(3, dtype=np.int32)
    w = np.zeros(3)
    cnt = np.zeros(1, dtype=np.int32)
    assert L.staq_natural_weights(3, dp(sites), 1, ip(tri), ip(nbr), lp, ip(adj), 1, dp(q),
                                  ip(np.array([5], dtype=np.int32)), 3, ip(idx), dp(w),
                                  ip(cnt)) == -1
    assert L.staq_natural_weights(3, dp(sites), 1, ip(tri), ip(nbr), lp, ip(adj), 1, dp(q),
                                  ip(np.array([0], dtype=np.int32)), 2, ip(idx), dp(w),
                                  ip(cnt)) == -1
    assert cnt[0] == 3
    assert L.staq_natural_weights(3, dp(sites), 1, ip(tri), ip(nbr), lp, ip(adj), 1, dp(q),
                                  ip(np.array([0], dtype=np.int32)), 3, ip(idx), dp(w),
                                  ip(cnt)) == 0
    assert cnt[0] == 3 and abs(w.sum() - 1.0) < 1e-12
    assert abs(w @ vals[idx] - out[0]) < 1e-12
