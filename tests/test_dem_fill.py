"""Depression filling (host preprocessing in libhp2d): Priority-Flood+epsilon."""
import numpy as np
import pytest

from hydropol2d_amd import synthetic


def _lib_ok():
    try:
        from hydropol2d_amd import _lib
        _lib.load()
        return True
    except Exception:
        return False


pytestmark = pytest.mark.skipif(not _lib_ok(), reason="libhp2d.so not built")


def test_bowl_fills_to_its_spill_point():
    # a 5x5 bowl whose lowest rim cell (1.0 m) is on the seeded last column
    z = np.full((5, 5), 3.0)
    z[1:4, 1:4] = 0.5
    z[2, 4] = 1.0
    seed = np.zeros((5, 5), np.uint8)
    seed[:, 4] = 1
    f = synthetic.fill_sinks(z, seed, epsilon=0.0)
    assert np.allclose(f[1:4, 1:4], 1.0)
    assert f[0, 0] == 3.0 and f[2, 4] == 1.0


def test_no_pits_remain_and_nothing_is_lowered():
    rng = np.random.default_rng(0)
    z = synthetic.spectral_dem(64, 80, 5.0, 0.8, 20.0, seed=4) + rng.normal(0, 0.05, (64, 80))
    seed = np.zeros(z.shape, np.uint8)
    seed[:, -1] = 1
    f = synthetic.fill_sinks(z, seed, epsilon=1e-4)
    assert (f >= z).all()
    ny, nx = z.shape
    for i in range(ny):
        for j in range(nx - 1):
            nb = [f[a, b] for a in range(max(i - 1, 0), min(i + 2, ny))
                  for b in range(max(j - 1, 0), min(j + 2, nx)) if (a, b) != (i, j)]
            assert min(nb) < f[i, j], (i, j)
    # cells that already drained keep their elevation
    assert (f == z).mean() > 0.2


def test_nan_cells_stay_outside():
    z = synthetic.spectral_dem(32, 32, 5.0, 0.8, 10.0, seed=1)
    z[:5, :5] = np.nan
    f = synthetic.fill_sinks(z)
    assert np.isnan(f[:5, :5]).all() and np.isfinite(f[5:, 5:]).all()
