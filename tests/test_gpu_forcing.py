"""GPU parity of the spatial forcing against the FP64 oracle:

* hp2d_idw_interpolate (Rainfall_Interpolator / idw_function) on random gauges;
* the device loop with rainfall maps interpolated from gauges (time-varying,
  with a gauge missing in one interval) and with inflow hydrographs on two
  gauges, FP64 to 1e-9 (the loop parity tolerance), FP32 within 1e-3 m.
"""
import numpy as np
import pytest

from hydropol2d_amd import synthetic
from hydropol2d_amd.runspec import Inflow
from oracle import oracle

pytestmark = pytest.mark.gpu


def _close(a, b, name, rtol=1e-9):
    a, b = np.asarray(a, float), np.asarray(b, float)
    np.testing.assert_array_equal(np.isnan(a), np.isnan(b), err_msg=name)
    scale = np.nanmax(np.abs(b)) if np.isfinite(b).any() else 1.0
    np.testing.assert_allclose(a, b, rtol=rtol, atol=rtol * (scale + 1e-300), equal_nan=True,
                               err_msg=name)


def test_idw_dropin_matches_oracle():
    from hydropol2d_amd import forcing
    rng = np.random.default_rng(4)
    for n_obs, ny, nx in ((1, 7, 9), (5, 120, 90), (700, 64, 80)):
        print(f"idw n_obs={n_obs} grid={ny}x{nx}", flush=True)
        x_grid = 500.0 + 30.0 * np.arange(1, nx + 1)
        y_grid = 9000.0 - 30.0 * np.arange(1, ny + 1)
        xo = rng.uniform(x_grid[0], x_grid[-1], n_obs)
        yo = rng.uniform(y_grid[-1], y_grid[0], n_obs)
        xo[0], yo[0] = x_grid[2], y_grid[3]        # one gauge on a grid point
        fo = rng.uniform(0.0, 80.0, n_obs)
        got, X, Y = forcing.Rainfall_Interpolator(xo, yo, fo, x_grid, y_grid)
        ref = oracle.idw(xo, yo, fo, x_grid, y_grid)
        np.testing.assert_allclose(got, ref, rtol=1e-13, atol=0)
        if n_obs < 700:
            assert got[3, 2] == pytest.approx(fo[0], rel=1e-12)
        assert X.shape == Y.shape == (ny, nx)


def _maps_spec(ny=48, nx=56, precision="fp64"):
    from hydropol2d_amd import forcing
    spec = synthetic.tilted_plane_c1(ny, nx, routing_time_min=10.0, precision=precision)
    dx = spec.resolution
    x_grid = 1000.0 + dx * np.arange(1, nx + 1)
    y_grid = 5000.0 - dx * np.arange(1, ny + 1)
    xg = np.array([1010.0, 1200.0, 1100.0, 1260.0])
    yg = np.array([4990.0, 4800.0, 4900.0, 4770.0])
    # three 4-min intervals of four gauges, one gauge missing in the second
    R = np.array([[80.0, 20.0, 50.0, 5.0],
                  [10.0, np.nan, 90.0, 40.0],
                  [30.0, 30.0, 0.0, 60.0]])
    spec.rain_maps_mmh = forcing.spatial_gauge_rain_maps(R, xg, yg, x_grid, y_grid)
    spec.rain_map_time_min = np.array([4.0, 8.0, 12.0])
    return spec


def _inflow_spec(ny=40, nx=50, precision="fp64"):
    spec = synthetic.tilted_plane_c1(ny, nx, routing_time_min=10.0, precision=precision)
    cells = np.zeros((ny, nx, 2), np.uint8)
    cells[5, 3, 0] = 1
    cells[20:23, 10, 1] = 1
    spec.inflow = Inflow(time_min=np.array([3.0, 6.0, 1e4]),
                         rate_mmh=np.array([[900.0, 100.0], [0.0, 400.0], [300.0, 300.0]]),
                         cells=cells)
    return spec


@pytest.mark.parametrize("case", ["maps", "inflow"])
def test_loop_forcing_matches_oracle(case):
    from hydropol2d_amd.simulation import Simulation
    spec = _maps_spec() if case == "maps" else _inflow_spec()
    print(f"loop forcing {case}", flush=True)
    with Simulation(spec) as s:
        s.run()
        rep, st, dts = s.report(), s.state(infiltration=True), s.dt_history()
    o = oracle.run(spec, dt_follow=dts)
    orep = o["report"]
    assert o["n_dt_mismatch"] == 0
    assert rep["steps"] == orep["steps"] and rep["n_reports"] == orep["n_reports"]
    _close(dts, o["dt_hist"][-len(dts):], "dt")
    for key in ("d_t", "outflow", "I_t", "dmax"):
        _close(st[key], o[key], key)
    for key in ("cum_rain_m3", "cum_inflow_m3", "cum_infiltration_m3", "outflow_volume_m3",
                "inflow_volume_m3", "t_min"):
        assert rep[key] == pytest.approx(orep[key], rel=1e-9, abs=1e-12), key
    if case == "maps":
        assert rep["cum_rain_m3"] > 0
    else:
        assert rep["cum_inflow_m3"] > 0
    scale = max(1.0, orep["storage_m3"] + orep["cum_infiltration_m3"])
    assert abs(rep["volume_error_m3"] - orep["volume_error_m3"]) <= 1e-9 * scale


@pytest.mark.parametrize("case", ["maps", "inflow"])
def test_loop_forcing_fp32_within_1mm(case):
    from hydropol2d_amd.simulation import Simulation
    spec64 = _maps_spec() if case == "maps" else _inflow_spec()
    spec32 = _maps_spec(precision="fp32") if case == "maps" else _inflow_spec(precision="fp32")
    with Simulation(spec32) as s:
        s.run()
        st, dts = s.state(), s.dt_history()
    o = oracle.run(spec64, dt_follow=dts, dt_force=True)
    err = float(np.nanmax(np.abs(st["d_t"] - o["d_t"]))) / 1000.0
    print(f"fp32 {case}: depth error {err:.2e} m", flush=True)
    assert err <= 1e-3
