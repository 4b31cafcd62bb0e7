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


def test_flow_accumulation_on_a_plane_and_a_valley():
    from hydropol2d_amd import synthetic
    # a plane falling along the columns: every row drains straight to the last column
    z = np.asfortranarray(100.0 - 0.05 * np.arange(12)[None, :] + 0.0 * np.arange(7)[:, None])
    a = synthetic.flow_accumulation(z)
    np.testing.assert_array_equal(a, np.tile(np.arange(1, 13, dtype=float), (7, 1)))
    # a V valley draining along its axis: the axis cell of the last column
    # collects the whole grid, the side cells collect only themselves
    i = np.arange(9)[:, None]
    j = np.arange(15)[None, :]
    z = np.asfortranarray(0.1 * np.abs(i - 4) + 0.01 * (14 - j))
    a = synthetic.flow_accumulation(z)
    assert a[4, 14] == z.size
    assert a[0, 0] == 1 and a[8, 7] == 1


def test_flow_accumulation_counts_every_cell_once():
    from hydropol2d_amd import synthetic
    z = synthetic.fill_sinks(synthetic.spectral_dem(48, 64, 5.0, 0.8, 20.0, seed=4))
    z[:5, :7] = np.nan
    a = synthetic.flow_accumulation(z)
    assert np.all(a[np.isnan(z)] == 0)
    # the cells that pass their count on are those with a lower D8 neighbour;
    # the counts kept at the others (the drainage points) add up to the valid cells
    ny, nx = z.shape
    keeps = np.zeros_like(a, dtype=bool)
    for ii in range(ny):
        for jj in range(nx):
            if np.isnan(z[ii, jj]):
                continue
            nb = z[max(ii - 1, 0):ii + 2, max(jj - 1, 0):jj + 2]
            keeps[ii, jj] = not np.any(nb < z[ii, jj])
    assert a[keeps].sum() == np.count_nonzero(~np.isnan(z))
    assert np.all(a[~np.isnan(z)] >= 1)


def test_d8_slope_closed_forms():
    from hydropol2d_amd import synthetic
    # plane falling 0.05 m per column at dx = 2: steepest drop along the row, 0.025;
    # the last column has no lower neighbour (0)
    z = 100.0 - 0.05 * np.arange(10)[None, :] + 0.0 * np.arange(6)[:, None]
    s = synthetic.d8_slope(z, 2.0)
    np.testing.assert_allclose(s[:, :-1], 0.025, rtol=1e-12)
    assert np.all(s[:, -1] == 0.0)
    # a cone: the centre is a pit (0); the cells next to it fall straight in
    i, j = np.meshgrid(np.arange(7), np.arange(7), indexing="ij")
    z = np.hypot(i - 3.0, j - 3.0)
    s = synthetic.d8_slope(z, 1.0)
    assert s[3, 3] == 0.0
    assert s[3, 4] == pytest.approx(1.0) and s[4, 4] == pytest.approx(1.0)
    z[0, 0] = np.nan
    assert np.isnan(synthetic.d8_slope(z, 1.0)[0, 0])
