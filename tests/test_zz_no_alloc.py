"""staq_local_icgn_noalloc (the 6-DOF local step that allocates no device
memory) against staq_local_icgn on the same subsets, through the C ABI: two
calls in a row on one context, the second with fewer subsets, each to the
IC-GN parity bar (codes agree, |dU| < 1e-3 px and |dF| < 1e-5 on every
subset converged in both)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(ctx, p, coords, win, U0, wmax):
    import torch
    dev = torch.device("cuda", 0)
    N = len(coords)
    t_U0 = torch.tensor(U0, dtype=torch.float64, device=dev)
    t_co = torch.tensor(np.asfortranarray(coords).ravel(order="F"), dtype=torch.float64, device=dev)
    t_w = torch.tensor(win, dtype=torch.int32, device=dev)
    U = torch.full((2 * N,), np.nan, dtype=torch.float64, device=dev)
    F = torch.full((4 * N,), np.nan, dtype=torch.float64, device=dev)
    H = torch.full((21 * N,), np.nan, dtype=torch.float64, device=dev)
    st = torch.full((N,), -7, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    vp = C.c_void_p
    rc = ctx.L.staq_local_icgn_noalloc(ctx.h, C.byref(p), 0, N, wmax, vp(t_U0.data_ptr()),
                                       vp(t_co.data_ptr()), vp(t_w.data_ptr()), vp(U.data_ptr()),
                                       vp(F.data_ptr()), vp(H.data_ptr()), vp(st.data_ptr()), None)
    ctx.check(rc, "staq_local_icgn_noalloc")
    torch.cuda.synchronize()
    return U.cpu().numpy(), F.cpu().numpy(), H.cpu().numpy().reshape(21, N).T, st.cpu().numpy()


def test_noalloc_matches_local_icgn(ctx):
    from staq_dic_amd import params, synth
    from staq_dic_amd.api import local_icgn_raw
    from test_gpu_parity import _cmp_icgn
    fr, coords, win = synth.config_hole(H=384, W=384, n_speckles=1700, max_disp=3.0, radius=60,
                                        margin=24, band=(96, 40))
    truth = synth.truth_at(fr.field, coords)
    U0 = np.round(truth[:, 4:6]).reshape(-1)
    ctx.set_frame(fr.f, fr.g, fr.mask, slot=0)
    p = params(tol=1e-3)
    for n in (len(coords), len(coords) // 3):
        c, w, u0 = coords[:n], win[:n], U0[:2 * n]
        print("no-alloc local step, hole 384^2, N =", n, flush=True)
        Ug, Fg, Hg, sg = _noalloc(ctx, p, c, w, u0, int(w.max()))
        Ur, Fr, Hr, sr = local_icgn_raw(u0, c, w, fr.f, fr.g, fr.mask, tol=1e-3, ctx=ctx)
        assert not np.any(sg == -7)
        _cmp_icgn(sg, sr, Ug, Ur, Fg, Fr, name="no-alloc vs staq_local_icgn")
        ok = (sg >= 1) & (sg <= 100) & (sr >= 1) & (sr <= 100)
        scale = np.abs(Hr[ok]).max(axis=1, keepdims=True)
        assert np.max(np.abs(Hg[ok] - Hr[ok]) / scale) < 1e-5


def test_noalloc_rejects_sizes_beyond_its_limits(ctx):
    from staq_dic_amd import params
    p = params(tol=1e-3)
    vp = C.c_void_p
    one = vp(8)  # never dereferenced: the size checks come first
    for N, wmax in (((1 << 20) + 1, 32), (0, 32), (10, 65), (10, 1)):
        rc = ctx.L.staq_local_icgn_noalloc(ctx.h, C.byref(p), 0, N, wmax, one, one, one, one, one,
                                           one, one, None)
        assert rc != 0, (N, wmax)
