"""Correlation quality factors (compute_qFactor, func/funIntegerSearch.m:
211-264): the oracle's statistics against closed forms and invariants, then
the device entry point staq_xcorr_qfactors against the oracle on the boxes of
the whole-field search, of the multigrid search's last level and on boxes too
large for LDS.

Pins of the oracle (oracle/pyoracle.py qfactor_stats): a single spike gives
pce = n and the entropy of an (n-1, 1) split; a map spread evenly over the
bins gives the entropy of that split over exactly 30 bins in natural log; pce
and ppe do not change when a constant is added to the map (they see
cc - min(cc)) or when it is scaled.  MATLAB's histcounts(x, 30) edge rule is
not in the reference's files; the bins here are 30 equal ones over
[0, max(cc - min(cc))] (ppe's binning is parity unpinned against MATLAB)."""
import numpy as np
import pytest


def test_qfactor_spike(oracle):
    n = 7 * 9
    cc = np.zeros((7, 9))
    cc[3, 4] = 0.8
    pce, ppe = oracle.qfactor_stats(cc)
    assert pce == pytest.approx(n, rel=1e-12)
    ent = (n - 1) / n * np.log(n / (n - 1)) + np.log(n) / n
    assert ppe == pytest.approx(1.0 / ent, rel=1e-12)


def test_qfactor_even_spread_uses_30_natural_log_bins(oracle):
    x = np.concatenate([[0.0], (np.arange(30) + 0.5) / 30.0, [1.0]])
    p = np.array([2] + [1] * 28 + [2]) / 32.0
    ent = float(np.sum(p * np.log(1.0 / p)))
    pce, ppe = oracle.qfactor_stats(x)
    assert ppe == pytest.approx(1.0 / ent, rel=1e-12)
    assert pce == pytest.approx(1.0 / np.mean(x ** 2), rel=1e-12)


def test_qfactor_shift_and_scale_invariance(oracle):
    rng = np.random.default_rng(3)
    cc = rng.uniform(-0.4, 0.9, (40, 35))
    a = oracle.qfactor_stats(cc)
    b = oracle.qfactor_stats(cc + 0.3)
    c = oracle.qfactor_stats(2.5 * cc)
    np.testing.assert_allclose(a, b, rtol=1e-9)
    np.testing.assert_allclose(a, c, rtol=1e-9)


def test_oracle_box_failures(oracle):
    f = np.random.default_rng(1).normal(size=(64, 64))
    f[:20, :20] = 0.5
    q, ok = oracle.xcorr_qfactors(f, f, [[2, 12, 2, 12], [30, 40, 30, 40], [50, 60, 50, 70]],
                                  [[1, 15, 1, 15], [28, 42, 28, 42], [48, 62, 48, 72]])
    np.testing.assert_array_equal(ok, [0, 1, 0])  # constant template; box off the image
    assert np.isnan(q[0]).all() and np.isfinite(q[1]).all()


def _boxes_whole_field(H, W, w, r, step):
    ii = np.arange(1 + r + 1, H - w - r - 1, step)
    jj = np.arange(1 + r + 1, W - w - r - 1, step)
    I, J = np.meshgrid(ii, jj, indexing="ij")
    I, J = I.ravel(), J.ravel()
    bf = np.stack([I, I + w, J, J + w], 1)
    bg = np.stack([I - r, I + w + r, J - r, J + w + r], 1)
    return bf, bg


@pytest.mark.gpu
def test_device_qfactors_against_oracle(ctx, oracle, plain_cfg):
    from staq_dic_amd.api import xcorr_qfactors
    fr, _, _ = plain_cfg
    f = fr.f.copy()
    f[:40, :40] = 0.25  # a constant template region
    r32 = lambda a: a.astype(np.float32).astype(np.float64)
    fo, go = r32(f), r32(fr.g)
    ctx.set_frame(f, fr.g, None, slot=0)
    H, W = f.shape
    sets = {"whole field w32 r6": _boxes_whole_field(H, W, 32, 6, 40),
            "whole field w32 r10": _boxes_whole_field(H, W, 32, 10, 44)}
    rng = np.random.default_rng(2)
    x0 = rng.integers(20, H - 60, 60)
    y0 = rng.integers(20, W - 60, 60)
    sets["multigrid last level (33 px, +-5)"] = (
        np.stack([x0, x0 + 32, y0, y0 + 32], 1),
        np.stack([x0 - 5 + 1, x0 + 32 + 5 + 1, y0 - 5 - 1, y0 + 32 + 5 - 1], 1))
    sets["large (HBM path)"] = (np.array([[200, 264, 200, 264], [150, 214, 220, 284]]),
                                np.array([[136, 328, 136, 328], [86, 278, 156, 348]]))
    sets["failures"] = (np.array([[2, 34, 2, 34], [500, 532, 10, 42], [100, 140, 100, 132]]),
                        np.array([[1, 40, 1, 40], [495, 537, 5, 47], [100, 130, 100, 140]]))
    for name, (bf, bg) in sets.items():
        print(name, len(bf), flush=True)
        q = xcorr_qfactors(ctx, 0, bf, bg, fail=np.nan, normalise=False)
        qo, oko = oracle.xcorr_qfactors(fo, go, bf, bg)
        okg = np.isfinite(q[:, 0])
        np.testing.assert_array_equal(okg, oko.astype(bool), err_msg=name)
        if not okg.any():
            continue
        np.testing.assert_allclose(q[okg, 0], qo[okg, 0], rtol=1e-9, err_msg=name)
        # a map value on a bin edge may fall on either side under a different
        # summation order: one count moves, a relative change of order 1/n
        rel = np.abs(q[okg, 1] / qo[okg, 1] - 1.0)
        print(f"  pce max rel {np.max(np.abs(q[okg, 0] / qo[okg, 0] - 1)):.2e}, ppe max rel "
              f"{rel.max():.2e}, ppe differing > 1e-9: {(rel > 1e-9).sum()}", flush=True)
        assert rel.max() < 2e-3 and (rel > 1e-9).mean() <= 0.05
    assert not np.isfinite(xcorr_qfactors(ctx, 0, *sets["failures"], fail=np.nan,
                                          normalise=False)[[0, 1, 2], 0]).any()


@pytest.mark.gpu
def test_integer_search_qfactors(ctx, oracle, plain_cfg):
    """funIntegerSearch's cc.qfactors: one row per search point in the
    reference's order, normalised to [0, 1] per column, equal to the oracle's
    statistics of the same boxes."""
    from staq_dic_amd.api import funIntegerSearch
    fr, _, _ = plain_cfg
    H, W = fr.f.shape
    x, y, u, v, cc = funIntegerSearch(fr.f, fr.g, 6, [1, H], [1, W], 32, 16, 0, ctx=ctx)
    q = cc["qfactors"]
    assert q.shape == (x.size, 2)
    np.testing.assert_allclose(np.nanmin(q, 0), 0.0)
    np.testing.assert_allclose(np.nanmax(q, 0), 1.0)
    ii = np.rint(x.ravel() - 16).astype(int)
    jj = np.rint(y.ravel() - 16).astype(int)
    bf = np.stack([ii, ii + 32, jj, jj + 32], 1)
    bg = np.stack([ii - 6, ii + 38, jj - 6, jj + 38], 1)
    r32 = lambda a: a.astype(np.float32).astype(np.float64)
    qo, _ = oracle.xcorr_qfactors(r32(fr.f), r32(fr.g), bf, bg)
    qo = (qo - qo.min(0)) / (qo - qo.min(0)).max(0)
    np.testing.assert_allclose(q[:, 0], qo[:, 0], atol=1e-9)
    assert np.abs(q[:, 1] - qo[:, 1]).max() < 5e-3
