"""GPU parity of the device loop with cellular-automata routing
(flags.flag_CA = 1: CA_Routing at HydroPol2D_Main_While.m:461-465, the inflow
gather and mass balance of :631-660, the CA branch of refreshing_timestep.m)
against the FP64 oracle loop: with Green-Ampt, the van Genuchten module, the
groundwater module, rainfall maps and inflow hydrographs, both outlet types,
the critical-velocity cap, LULC h_0, cells outside the domain, an injected
rejected step (which keeps I_tot_end_cell, as the reference's catch block
does), eager against graph launches, and FP32 within bounds.

Tolerance (FP64): 1e-9 of each field's scale (the device's pow and hypot may
differ from the host library's in the last place); step counts exactly."""
import numpy as np
import pytest

import loop_cases
from hydropol2d_amd import forcing
from hydropol2d_amd.runspec import Inflow
from oracle import oracle

pytestmark = pytest.mark.gpu


def _close(a, b, name, rtol=1e-9):
    a, b = np.asarray(a, float), np.asarray(b, float)
    np.testing.assert_array_equal(np.isnan(a), np.isnan(b), err_msg=name)
    scale = np.nanmax(np.abs(b)) if np.isfinite(b).any() else 1.0
    np.testing.assert_allclose(a, b, rtol=rtol, atol=rtol * (scale + 1e-300), equal_nan=True,
                               err_msg=name)


def _with_maps_and_inflow(spec):
    ny, nx = spec.shape
    dx = spec.resolution
    x_grid = 1000.0 + dx * np.arange(1, nx + 1)
    y_grid = 5000.0 - dx * np.arange(1, ny + 1)
    xg = np.array([1010.0, 1100.0, 1060.0])
    yg = np.array([4990.0, 4900.0, 4880.0])
    R = np.array([[80.0, 20.0, 50.0], [10.0, 90.0, 40.0]])
    spec.rain_maps_mmh = forcing.spatial_gauge_rain_maps(R, xg, yg, x_grid, y_grid)
    spec.rain_map_time_min = np.array([2.0, 6.0])
    cells = np.zeros((ny, nx), np.uint8)
    cells[5, 3] = 1
    cells[20:23, 10] = 2
    spec.inflow = Inflow(time_min=np.array([3.0, 1e4]),
                         rate_mmh=np.array([[900.0, 100.0], [0.0, 400.0]]), cells=cells)
    return spec


CASES = {
    "ca_ga": dict(),
    "ca_no_infiltration_h0": dict(infiltration=None, h0=True),
    "ca_normal_slope_critical": dict(outlet_type=1, slope=0.02, critical=1),
    "ca_vg_nan_corner": dict(infiltration="vg", nan_corner=True),
    "ca_groundwater": dict(gw=True, every=3),
    "ca_maps_inflow": dict(forcing=True),
    "ca_closed_basin": dict(outlets=False),
}


def _spec(name, **over):
    kw = dict(CASES[name])
    kw.update(over)
    forc = kw.pop("forcing", False)
    spec = loop_cases.ca_plane(**kw)
    return _with_maps_and_inflow(spec) if forc else spec


@pytest.mark.parametrize("name", list(CASES))
def test_ca_loop_matches_oracle(name):
    from hydropol2d_amd.simulation import Simulation
    spec = _spec(name)
    gw = spec.groundwater is not None
    inf = spec.infiltration is not None
    print(f"CA loop case {name} {CASES[name]}", flush=True)
    with Simulation(spec) as s:
        s.run()
        rep, st, dts = s.report(), s.state(infiltration=inf, groundwater=gw), s.dt_history()
    o = oracle.run(spec, dt_follow=dts)
    orep = o["report"]
    print(f"  {rep['steps']} steps, {o['n_dt_ties']} dt ties, outflow {rep['outflow_volume_m3']:.4g} m3",
          flush=True)
    assert o["n_dt_mismatch"] == 0
    for key in ("steps", "rejected_steps", "n_reports", "gw_updates"):
        assert rep[key] == orep[key], (key, rep[key], orep[key])
    _close(dts, o["dt_hist"][-len(dts):], "dt")
    keys = ["d_t", "outflow", "dmax"] + (["I_t"] if inf else []) + (["gw_h"] if gw else [])
    for key in keys:
        _close(st[key], o[key], key)
    for key in ("t_min", "cum_rain_m3", "cum_inflow_m3", "cum_infiltration_m3", "storage_m3",
                "outflow_volume_m3", "inflow_volume_m3"):
        assert rep[key] == pytest.approx(orep[key], rel=1e-9, abs=1e-12), key
    scale = max(1.0, orep["storage_m3"] + orep["cum_infiltration_m3"])
    assert abs(rep["volume_error_m3"] - orep["volume_error_m3"]) <= 1e-9 * scale
    hy = None
    with Simulation(spec) as s:
        s.run()
        hy = s.hydrograph()
    np.testing.assert_allclose(hy["hydro_q"], o["hydro_q"][:len(hy["hydro_q"])], rtol=1e-9,
                               atol=1e-12)


def test_ca_rejected_step_keeps_the_limiter():
    from hydropol2d_amd.simulation import Simulation
    spec = _spec("ca_ga")
    o = oracle.run(spec, inject_fail_step=12)
    with Simulation(spec) as s:
        s.inject_failure(12)
        s.run()
        rep, st = s.report(), s.state(infiltration=True)
    assert rep["rejected_steps"] == o["report"]["rejected_steps"] == 1
    assert rep["steps"] == o["report"]["steps"]
    _close(st["d_t"], o["d_t"], "d_t")
    _close(st["I_t"], o["I_t"], "I_t")


def test_ca_eager_and_graph_paths_agree():
    from hydropol2d_amd.simulation import Simulation
    # (a module call every 4 steps: the timing path checks the cadence before
    # each launch, so a call due after the last of the 40 steps would run there)
    spec = _spec("ca_groundwater", every=4)
    with Simulation(spec) as a, Simulation(spec) as b:
        a.run(40)
        b.run_fixed(40, graph_steps=0)
        sa = a.state(infiltration=True, groundwater=True)
        sb = b.state(infiltration=True, groundwater=True)
        assert a.report()["gw_updates"] == b.report()["gw_updates"] > 0
    for key in ("d_t", "I_t", "gw_h", "outflow"):
        np.testing.assert_array_equal(sa[key], sb[key], err_msg=key)


def test_ca_fp32_loop_within_bounds_of_fp64_oracle():
    from hydropol2d_amd.simulation import Simulation
    o = oracle.run(_spec("ca_maps_inflow"))
    with Simulation(_spec("ca_maps_inflow", precision="fp32")) as s:
        s.run()
        st, rep = s.state(infiltration=True), s.report()
    err = np.nanmax(np.abs(st["d_t"] - o["d_t"]))
    print(f"fp32 CA loop: depth {err:.2e} mm of {np.nanmax(o['d_t']):.3g} mm, "
          f"{rep['steps']} / {o['report']['steps']} steps", flush=True)
    assert err <= 1.0
    assert rep["outflow_volume_m3"] == pytest.approx(o["report"]["outflow_volume_m3"], rel=2e-3)


def test_ca_map_records_and_human_instability_match_oracle():
    """CA runs with the map stacks (save_output_maps.m:24-27, 148-258) and the
    Human_Instability_Module (HydroPol2D_Main_While.m:718) on the CA branch's
    velocities.  (A 40 mm wet start makes this plane's CA run chaotic: the
    oracle's own dt sequence changes at step 99 when the roughness moves by one
    unit in the last place; the 5 mm start moves it by 1e-13 mm.)"""
    from hydropol2d_amd.simulation import Simulation
    spec = _spec("ca_ga", T=6.0, d0=5.0)
    spec.human_instability = {"weight_person": 20.0}
    spec.record_time_maps_min = 1.0
    with Simulation(spec) as s:
        s.run()
        rec, mr, dts = s.map_records(), s.max_risk(), s.dt_history()
    o = oracle.run(spec, dt_follow=dts, n_map_records=64)
    assert o["n_dt_mismatch"] == 0
    assert len(rec["t_min"]) == len(o["map_record_t"]) >= 5
    np.testing.assert_allclose(rec["t_min"], o["map_record_t"], rtol=1e-12)
    for name, arr in o["map_records"].items():
        if np.all(np.isnan(arr)):
            assert name not in rec, name
            continue
        _close(rec[name], arr, name)
    _close(mr, o["max_risk"], "max_risk")
    assert np.nanmax(mr) > 0 and np.nanmax(rec["velocity"]) > 0


def test_ca_refuses_what_it_does_not_run():
    from hydropol2d_amd._lib import HP2DError
    from hydropol2d_amd.simulation import Ensemble, Simulation
    spec = _spec("ca_ga")
    spec.water_quality = loop_cases.plane().water_quality
    with pytest.raises(HP2DError):
        Simulation(spec)
    with pytest.raises(HP2DError):
        Ensemble([_spec("ca_ga")])
    with Simulation(_spec("ca_ga")) as s:
        with pytest.raises(HP2DError):
            s.autotune()
