"""GPU parity of the device-resident step loop (rain + Green-Ampt + LI + dt
control + hydrograph + mass balance) against the FP64 oracle driver.

FP64 mode: depths, fluxes, dt sequence, hydrograph and volume_error within
1e-9 relative.  FP32 mode: depths within 1e-3 m of the FP64 oracle.
"""
import numpy as np
import pytest

from hydropol2d_amd import synthetic
from oracle import oracle

pytestmark = pytest.mark.gpu


def _sim(spec):
    from hydropol2d_amd.simulation import Simulation
    return Simulation(spec)


def _close(a, b, rtol=1e-9, atol_scale=1e-9, name=""):
    a, b = np.asarray(a, float), np.asarray(b, float)
    np.testing.assert_array_equal(np.isnan(a), np.isnan(b), err_msg=name)
    scale = np.nanmax(np.abs(b)) if b.size and np.isfinite(np.nanmax(np.abs(b))) else 1.0
    np.testing.assert_allclose(a, b, rtol=rtol, atol=atol_scale * (scale + 1e-300), equal_nan=True,
                               err_msg=name)


@pytest.mark.parametrize("shape", [(64, 64), (61, 130), (150, 47)])
def test_sim_fp64_full_run_matches_oracle(shape):
    spec = synthetic.tilted_plane_c1(*shape, routing_time_min=20.0)
    print(f"sim fp64 shape={shape}", flush=True)
    o = oracle.run(spec)
    with _sim(spec) as s:
        done = s.run()
        assert done
        rep, st, hy, dts = s.report(), s.state(infiltration=True), s.hydrograph(), s.dt_history()
    orep = o["report"]
    assert rep["steps"] == orep["steps"]
    _close(dts, o["dt_hist"][-len(dts):], name="dt")
    _close(st["d_t"], o["d_t"], name="d_t")
    _close(st["outflow"], o["outflow"], name="outflow")
    _close(st["I_t"], o["I_t"], name="I_t")
    _close(st["dmax"], o["dmax"], name="dmax")
    _close(hy["hydro_q"], o["hydro_q"], name="hydrograph")
    assert hy["n_hydro"] == o["n_hydro"]
    for key in ("t_min", "time_step_s", "max_velocity", "outflow_volume_m3", "inflow_volume_m3",
                "outlet_runoff_volume_mm", "cum_infiltration_m3", "cum_rain_m3", "storage_m3"):
        assert rep[key] == pytest.approx(orep[key], rel=1e-9, abs=1e-9), key
    assert rep["n_reports"] == orep["n_reports"]
    # volume_error is a difference of large storages: compare against the storage scale
    assert abs(rep["volume_error_m3"] - orep["volume_error_m3"]) <= 1e-9 * max(1.0, orep["storage_m3"] + orep["cum_infiltration_m3"])


def test_sim_fp64_partial_runs_compose():
    spec = synthetic.tilted_plane_c1(40, 52, routing_time_min=10.0)
    o = oracle.run(spec, max_steps=37)
    with _sim(spec) as s:
        s.run(20)
        s.run(17)
        rep, st = s.report(), s.state()
    assert rep["steps"] == 37
    _close(st["d_t"], o["d_t"], name="d_t")


def test_sim_fp32_within_1mm_of_fp64_oracle():
    spec = synthetic.tilted_plane_c1(96, 80, routing_time_min=15.0)
    o = oracle.run(spec)
    spec32 = synthetic.tilted_plane_c1(96, 80, routing_time_min=15.0, precision="fp32")
    with _sim(spec32) as s:
        s.run()
        st = s.state()
    err = np.nanmax(np.abs(st["d_t"] - o["d_t"])) / 1000.0
    assert err <= 1e-3, err


def test_sim_rollback_matches_oracle():
    spec = synthetic.tilted_plane_c1(48, 48, routing_time_min=8.0)
    o = oracle.run(spec, inject_fail_step=30)
    with _sim(spec) as s:
        s.inject_failure(30)
        s.run()
        rep, st = s.report(), s.state()
    assert rep["rejected_steps"] == 1 == o["report"]["rejected_steps"]
    assert rep["catch_index"] == o["report"]["catch_index"]
    assert rep["steps"] == o["report"]["steps"]
    _close(st["d_t"], o["d_t"], name="d_t")


def test_sim_is_deterministic():
    spec = synthetic.tilted_plane_c1(130, 190, routing_time_min=5.0, precision="fp32")
    outs = []
    for _ in range(2):
        with _sim(spec) as s:
            s.run()
            outs.append((s.state()["d_t"], s.report()))
    np.testing.assert_array_equal(outs[0][0], outs[1][0])
    assert outs[0][1] == outs[1][1]


def test_closed_basin_conserves_volume_at_full_size():
    # 4096 x 4096 FP32, no rain / outlet: every step moves water but the
    # total volume is a size-independent invariant.
    ny = nx = 4096
    rng = np.random.default_rng(7)
    z = synthetic.spectral_dem(ny, nx, 2.0, seed=9) * 0.05
    d0 = np.asfortranarray(rng.uniform(0, 100.0, (ny, nx)))
    from hydropol2d_amd.runspec import RunSpec
    spec = RunSpec(elevation=np.asfortranarray(z), roughness=np.full((ny, nx), 0.03, order="F"),
                   resolution=2.0, d0=d0, precision="fp32", routing_time_min=1e6)
    v0 = d0.sum(dtype=np.float64)
    with _sim(spec) as s:
        s.run(20)
        d = s.state()["d_t"]
        rep = s.report()
    assert rep["steps"] == 20
    assert d.min() >= 0
    assert abs(d.sum(dtype=np.float64) - v0) / v0 < 1e-5


def test_fp32_error_per_output_field():
    # The FP32 state path against the FP64 oracle on the C1 plane (GA, rain,
    # critical outlet) at 128 x 128 over 20 min: the error of every output
    # field, printed as a table, with the bound each one is held to.
    spec = synthetic.tilted_plane_c1(128, 128, routing_time_min=20.0)
    o = oracle.run(spec)
    spec32 = synthetic.tilted_plane_c1(128, 128, routing_time_min=20.0, precision="fp32")
    with _sim(spec32) as s:
        s.run()
        st = s.state(infiltration=True)
        hy = s.hydrograph()
        rep = s.report()
    orep = o["report"]
    n_h = min(hy["n_hydro"], o["n_hydro"])
    qo, qg = o["hydro_q"][:n_h], hy["hydro_q"][:n_h]
    rows = [
        ("d_t [m]", np.nanmax(np.abs(st["d_t"] - o["d_t"])) / 1000, 1e-3),
        ("dmax [m]", np.nanmax(np.abs(st["dmax"] - o["dmax"])) / 1000, 1e-3),
        ("I_t [mm]", np.nanmax(np.abs(st["I_t"] - o["I_t"])), 0.05),
        ("outflow x [mm/h], rel. to max", np.nanmax(np.abs(st["outflow"][:, :, 0] - o["outflow"][:, :, 0]))
         / np.nanmax(np.abs(o["outflow"][:, :, 0])), 2e-2),
        ("outflow y [mm/h], rel. to max", np.nanmax(np.abs(st["outflow"][:, :, 1] - o["outflow"][:, :, 1]))
         / np.nanmax(np.abs(o["outflow"][:, :, 1])), 2e-2),
        ("outlet hydrograph [m3/s], rel. to peak", np.max(np.abs(qg - qo)) / max(np.max(qo), 1e-12), 1e-2),
        ("outflow volume [m3], rel.", abs(rep["outflow_volume_m3"] - orep["outflow_volume_m3"])
         / max(orep["outflow_volume_m3"], 1e-12), 1e-2),
        # the FP64 ledgers integrate over the FP32 run's own dt sequence, whose
        # last step ends at a slightly different time
        ("cum. rain [m3], rel.", abs(rep["cum_rain_m3"] - orep["cum_rain_m3"]) / orep["cum_rain_m3"], 1e-4),
        ("cum. infiltration [m3], rel.", abs(rep["cum_infiltration_m3"] - orep["cum_infiltration_m3"])
         / orep["cum_infiltration_m3"], 1e-3),
        ("storage [m3], rel.", abs(rep["storage_m3"] - orep["storage_m3"]) / orep["storage_m3"], 1e-3),
        ("max velocity [m/s], rel.", abs(rep["max_velocity"] - orep["max_velocity"])
         / orep["max_velocity"], 2e-2),
        ("steps, rel.", abs(rep["steps"] - orep["steps"]) / orep["steps"], 1e-2),
        # mass_balance_check's volume_error (a step's fluxes against the storage
        # change since the last report) is large by construction; it is compared
        # with the oracle's, relative to the storage
        ("volume error [m3], rel. to storage", abs(rep["volume_error_m3"] - orep["volume_error_m3"])
         / orep["storage_m3"], 1e-3),
    ]
    print("\nFP32 vs FP64 oracle, 128x128 C1 plane, 20 min:")
    for name, err, bound in rows:
        print(f"  {name:45s} {err:.3e}  (bound {bound:.1e})")
    for name, err, bound in rows:
        assert err <= bound, (name, err, bound)
