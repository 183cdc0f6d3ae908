"""Integer-search outlier filter (funRemoveOutliers, inpaint_nans 3 / 4, the
3-value median test) against closed forms and hand-computed cases."""
import numpy as np
import pytest

from staq_dic_amd.outliers import funRemoveOutliers, inpaint_nans, med_filt2


def _plane(n, m):
    r, c = np.meshgrid(np.arange(n), np.arange(m), indexing="ij")
    return 0.7 + 0.05 * r - 0.03 * c


@pytest.mark.parametrize("method", [3, 4])
def test_inpaint_reproduces_a_plane_in_the_interior(method):
    A = _plane(20, 17)
    B = A.copy()
    B[5:9, 4:7] = np.nan
    B[12, 10] = np.nan
    np.testing.assert_allclose(inpaint_nans(B, method), A, atol=1e-10)


def test_inpaint_springs_average_of_neighbours():
    """A single NaN with springs to its four neighbours takes their mean."""
    B = np.arange(25, dtype=float).reshape(5, 5) ** 2
    B[2, 2] = np.nan
    want = (B[1, 2] + B[3, 2] + B[2, 1] + B[2, 3]) / 4
    assert inpaint_nans(B, 4)[2, 2] == pytest.approx(want, abs=1e-12)


def test_median_of_self_below_and_diagonal():
    V = np.array([[1.0, 5.0, 2.0], [4.0, 3.0, 9.0], [7.0, 8.0, 6.0]])
    got = med_filt2(V)
    # (element, below, below-right); the last row / column mirror themselves
    want = np.array([[np.median([1, 4, 3]), np.median([5, 3, 9]), np.median([2, 9, 9])],
                     [np.median([4, 7, 8]), np.median([3, 8, 6]), np.median([9, 6, 6])],
                     [np.median([7, 7, 8]), np.median([8, 8, 6]), np.median([6, 6, 6])]])
    np.testing.assert_array_equal(got, want.astype(np.float32))
    assert got.dtype == np.float32


def test_remove_outliers_spike_and_holes():
    u = _plane(15, 15) * 3
    v = -_plane(15, 15)
    u[7, 8] += 40.0  # a wrong integer peak
    u[2, 2] = v[2, 2] = np.nan  # a hole point
    u2, v2, _, _, _, lst = funRemoveOutliers(u, v, {}, 0, 100)
    assert 7 + 8 * 15 + 1 in lst.tolist()
    ref = _plane(15, 15) * 3
    assert u2[7, 8] == pytest.approx(ref[7, 8], abs=1e-9)
    assert np.isnan(u2[2, 2]) and np.isnan(v2[2, 2])
    ok = np.ones_like(u, dtype=bool)
    ok[2, 2] = False
    np.testing.assert_allclose(u2[ok], ref[ok], atol=1e-9)
    np.testing.assert_allclose(v2[ok], (-_plane(15, 15))[ok], atol=1e-12)


def test_remove_outliers_keeps_a_smooth_field():
    u = _plane(12, 10) + 0.01 * np.sin(np.arange(120).reshape(12, 10))
    v = 2 * _plane(12, 10)
    u2, v2, _, _, _, lst = funRemoveOutliers(u, v, {}, 0, 100)
    assert lst.size == 0
    np.testing.assert_array_equal(u2, u)
