"""Pins of the oracle's loop with cellular-automata routing in place of the
local-inertial step (flags.flag_CA = 1: CA_Routing at
HydroPol2D_Main_While.m:461-465, the inflow gather and mass balance of
:631-660, the CA branch of refreshing_timestep.m).  CA_Routing and the gather
are pinned by their own tests (test_oracle_d8.py, test_oracle_li.py); these
pin the loop around them: the first step equals the per-call routines on the
post-rain depth, the closed basin keeps its water, a lake at rest stays at
rest with the dt at its maximum, and a mirror-symmetric valley stays
symmetric."""
import numpy as np
import pytest

import loop_cases
from hydropol2d_amd import synthetic
from oracle import oracle


@pytest.mark.parametrize("outlet_type,slope", [(2, 0.0), (1, 0.02)])
def test_first_step_is_the_per_call_routing(outlet_type, slope):
    spec = loop_cases.ca_plane(infiltration=None, outlet_type=outlet_type, slope=slope, h0=True)
    ny, nx = spec.shape
    o = oracle.run(spec, max_steps=1)
    # the first step runs at min_dt with the plane's rain rate
    dt_min = spec.min_time_step_s / 60.0
    rain = oracle.rain_avg(spec.rain_time_min, spec.rain_intensity_mmh, 0.0, dt_min)[0] * dt_min / 60.0
    d_tot = spec.d0 + rain
    rows, cols = np.nonzero(spec.outlet_mask.T)
    r = oracle.ca_routing(spec.elevation, d_tot, spec.roughness, spec.resolution ** 2, dt_min,
                          spec.ca_h0_mm, spec.resolution, np.zeros((ny, nx)), outlet_type, slope,
                          cols + 1, rows + 1)
    q = np.stack([r[k] for k in ("qout_left", "qout_right", "qout_up", "qout_down",
                                 "outlet_flow")], axis=2)
    _, d = oracle.gather(q, dt_min, r["d_t"])
    np.testing.assert_array_equal(o["d_t"], d)
    np.testing.assert_array_equal(o["outflow"][:, :, 0], r["qout_right"])
    np.testing.assert_array_equal(o["outflow"][:, :, 1], r["qout_up"])
    np.testing.assert_array_equal(o["outflow"][:, :, 2], r["outlet_flow"])
    assert (r["qout_right"] > 0).any()


@pytest.mark.parametrize("infiltration", [None, "ga", "vg"])
def test_closed_basin_keeps_its_water(infiltration):
    spec = loop_cases.ca_plane(outlets=False, infiltration=infiltration, T=3.0)
    if infiltration == "ga":  # no drainage out of the soil (k_out = 0)
        spec.infiltration.theta_sat_minus_r = np.array([0.0])
    ny, nx = spec.shape
    A = spec.resolution ** 2
    o = oracle.run(spec)
    rep = o["report"]
    I0 = 0.0
    if infiltration == "ga":
        I0 = spec.infiltration.I_min_mm * ny * nx
    S0 = (np.sum(spec.d0) + I0) / 1000 * A
    S1 = (np.sum(o["d_t"]) + (np.sum(o["I_t"]) if infiltration else 0.0)) / 1000 * A
    assert rep["steps"] > 10
    assert S1 - S0 == pytest.approx(rep["cum_rain_m3"], rel=1e-11)


def test_lake_at_rest_stays_at_rest():
    ny, nx = 12, 15
    rng = np.random.default_rng(4)
    z = rng.uniform(0.0, 0.5, (ny, nx))
    spec = synthetic.tilted_plane_c1(ny, nx, routing_time_min=1.0)
    spec.elevation = np.asfortranarray(z)
    spec.d0 = np.asfortranarray((1.0 - z) * 1000.0)  # level water surface 1 m
    spec.outlet_mask = np.zeros((ny, nx), np.uint8, order="F")
    spec.rain_intensity_mmh = np.zeros_like(spec.rain_intensity_mmh)
    spec.infiltration = None
    spec.routing = "ca"
    o = oracle.run(spec)
    np.testing.assert_allclose(o["d_t"], spec.d0, rtol=0, atol=1e-9)
    dts = o["dt_hist"][:o["report"]["steps"]]
    # no cell moves: every cell is stagnant and the dt goes to its maximum
    assert dts[-1] == spec.max_time_step_s


def test_mirror_valley_stays_symmetric():
    ny, nx = 20, 21
    x = np.abs(np.arange(nx) - nx // 2) * 0.05
    y = np.arange(ny)[:, None] * 0.02
    z = np.asfortranarray(x[None, :] + y + 10.0)
    spec = synthetic.tilted_plane_c1(ny, nx, routing_time_min=3.0)
    spec.elevation = z
    spec.roughness = np.full((ny, nx), 0.03, order="F")
    out = np.zeros((ny, nx), np.uint8, order="F")
    out[0, nx // 2] = 1
    spec.outlet_mask = out
    spec.infiltration = None
    spec.routing = "ca"
    o = oracle.run(spec)
    d = o["d_t"]
    assert o["report"]["outflow_volume_m3"] > 0
    np.testing.assert_allclose(d, d[:, ::-1], rtol=1e-12, atol=1e-12)


def test_ca_refuses_wash_off():
    spec = loop_cases.ca_plane()
    spec.water_quality = loop_cases.plane().water_quality
    with pytest.raises(RuntimeError):
        oracle.run(spec, max_steps=1)


@pytest.mark.parametrize("adaptive", [1, 0])
def test_ca_time_step_of_a_uniform_sheet(adaptive):
    # every cell passes q = 360 mm/h to its right-hand neighbour at depth D = 20 mm
    # on a 5 m grid: u = q / 3.6e6 * R / (D / 1000) = 0.025 m/s (the CA components
    # are plain differences, no 1/2), c = sqrt(9.81 * 0.02) m/s; the new dt is the
    # Courant candidate snapped down to 1 ms
    ny, nx, R, q, D = 6, 7, 5.0, 360.0, 20.0
    q4 = np.zeros((ny, nx, 4))
    q4[:, :, 1] = q
    tcr, vmax, raster = oracle.refresh_dt_ca(q4, np.full((ny, nx), D), R,
                                             flag_adaptive_timestepping=adaptive, max_time_step=1e4)
    u = q / 3.6e6 * R / (D / 1000.0)
    c = np.sqrt(9.81 * D / 1000.0)
    assert vmax == pytest.approx(u, rel=1e-14)
    np.testing.assert_allclose(raster, u, rtol=1e-14)
    dt_adv = R / u * 0.7
    dt_wave = 0.7 * R / c if adaptive else 0.7 * R / (u + c)
    assert tcr == pytest.approx(np.floor(min(dt_adv, dt_wave) / 0.001) * 0.001, abs=1e-9)
