"""Host helpers of the realigned dataset (CPU)."""
import numpy as np

from simu_rca_ab_amd.dataset import normalize_rf, realignment_grids


def test_realignment_grid_closed_form():
    lam = 0.2e-3
    bub = np.array([[1e-3, 2e-3], [-1e-3, 0.5e-3], [10e-3, 12e-3]])
    X, Y, Z = realignment_grids(bub, lam)
    assert X.shape == Y.shape == Z.shape == (17, 2)
    np.testing.assert_allclose(X[:, 1], 2e-3 + lam / 4)
    np.testing.assert_allclose(Y[:, 0], -1e-3 + lam / 4)
    # z = (1:17) * lam/4 + z_bubble - 2 lam: the column spans the bubble
    np.testing.assert_allclose(Z[:, 0], np.arange(1, 18) * lam / 4 + 10e-3 - 2 * lam)
    assert Z[7, 0] == 10e-3   # 8th depth (1-based) is the bubble


def test_normalize_rf():
    rf = np.array([[-3.0, 1.5], [0.75, 0.0]])
    np.testing.assert_allclose(normalize_rf(rf), rf / 3 * 32768)
    z = np.zeros(4)
    np.testing.assert_array_equal(normalize_rf(z), z)
