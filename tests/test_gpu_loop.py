"""GPU parity of the device-resident loop with pollutant wash-off
(HydroPol2D_Main_While.m:683-695, 724-731, 919-922, build_up_wash_off.m) and
the groundwater module every N accepted steps, against the FP64 oracle.

Tolerance (FP64 mode): the wash-off power law and the van Genuchten recharge
use the device's pow, which may differ from the host library's in the last
place, so fields are compared to 1e-9 relative (of the field's scale); step
counts, rejections, splits and groundwater calls must agree exactly.

The dt of a wash-off run is snapped down to the increments by a floor(), and
a cell washed out to zero that refills gives tmin_wq = dt up to rounding, so
the snap regularly sits on a tie at the level of the last bits.  The oracle is
therefore run on the device's dt sequence, which it adopts only at such ties
(its own pre-snap dt within 1e-9 of an increment boundary and the device's dt
one increment away); any other dt difference fails the test.
"""
import numpy as np
import pytest

import loop_cases
from oracle import oracle

pytestmark = pytest.mark.gpu


def _close(a, b, name, rtol=1e-9):
    a, b = np.asarray(a, float), np.asarray(b, float)
    np.testing.assert_array_equal(np.isnan(a), np.isnan(b), err_msg=name)
    scale = np.nanmax(np.abs(b)) if np.isfinite(b).any() else 1.0
    np.testing.assert_allclose(a, b, rtol=rtol, atol=rtol * (scale + 1e-300), equal_nan=True,
                               err_msg=name)


# wq_rtol: the stress case washes a cell out within a fraction of a step for
# hundreds of steps (rates ~ 1/dt); its internal sub-steps amplify last-bit
# differences of pow in the pollutant masses to a few 1e-9, while the
# hydraulic state stays within 1e-9
CASES = {
    "split_in_graph": dict(C3=3e4),
    "long_splits_mass_lost": dict(C3=3e5, min_Bt=5.0),
    "wash_off_errors": dict(C3=1e6, min_Bt=5.0, max_steps=300, wq_rtol=1e-7),
    "cells_outside": dict(C3=3e4, nan_corner=True, every=7),
    "rating_curve": dict(C3=500.0, model=1, Bmin=2.0, Bmax=6.0),
    "no_groundwater": dict(C3=6e4, gw=False),
    "groundwater_only": dict(wq=False, every=3, h0_offset=-0.05),
}


@pytest.mark.parametrize("name", list(CASES))
def test_loop_matches_oracle(name):
    from hydropol2d_amd.simulation import Simulation
    kw = dict(CASES[name])
    max_steps = kw.pop("max_steps", 0)
    wq_rtol = kw.pop("wq_rtol", 1e-9)
    spec = loop_cases.plane(**kw)
    wq, gw = spec.water_quality is not None, spec.groundwater is not None
    print(f"loop case {name} {kw}", flush=True)
    with Simulation(spec) as s:
        s.run(max_steps)
        rep, st = s.report(), s.state(infiltration=True, water_quality=wq, groundwater=gw)
        dts = s.dt_history()
    assert len(dts) == rep["steps"] < 4096
    o = oracle.run(spec, max_steps=max_steps, dt_follow=dts)
    print(f"  {rep['steps']} steps, {o['n_dt_ties']} dt ties", flush=True)
    assert o["n_dt_mismatch"] == 0
    orep = o["report"]
    for key in ("steps", "rejected_steps", "catch_index", "n_reports", "wq_substep_steps",
                "gw_updates"):
        assert rep[key] == orep[key], (key, rep[key], orep[key])
    _close(dts, o["dt_hist"][-len(dts):], "dt")
    _close(st["d_t"], o["d_t"], "d_t")
    _close(st["outflow"], o["outflow"], "outflow")
    _close(st["I_t"], o["I_t"], "I_t")
    _close(st["dmax"], o["dmax"], "dmax")
    for key in ("t_min", "time_step_s", "cum_rain_m3", "cum_infiltration_m3", "storage_m3",
                "outflow_volume_m3"):
        assert rep[key] == pytest.approx(orep[key], rel=1e-9, abs=1e-12), key
    if wq:
        _close(st["B_t"], o["B_t"], "B_t", wq_rtol)
        _close(st["P_conc"], o["P_conc"], "P_conc", wq_rtol)
        _close(st["Tot_Washed"], o["Tot_Washed"], "Tot_Washed", wq_rtol)
        for key in ("wq_mass_kg", "wq_mass_lost_kg", "wq_out_conc_mgl", "wq_tmin_s",
                    "wq_mass_outlet_mg", "wq_vol_outlet"):
            assert rep[key] == pytest.approx(orep[key], rel=wq_rtol, abs=1e-12), key
    if gw:
        _close(st["gw_h"], o["gw_h"], "gw_h")
        for key in ("gw_storage_m3", "gw_total_recharge_m3", "gw_total_exfiltration_m3"):
            assert rep[key] == pytest.approx(orep[key], rel=1e-9, abs=1e-9), key
    scale = max(1.0, orep["storage_m3"] + orep["cum_infiltration_m3"] + orep.get("gw_storage_m3", 0))
    assert abs(rep["volume_error_m3"] - orep["volume_error_m3"]) <= 1e-9 * scale


def test_eager_and_graph_paths_agree():
    # the per-launch timing path (holds and groundwater calls between launches)
    # reaches the same state as the captured-graph path
    from hydropol2d_amd.simulation import Simulation
    spec = loop_cases.plane(C3=3e5, min_Bt=5.0, every=4)
    with Simulation(spec) as a, Simulation(spec) as b:
        a.run(40)
        b.run_fixed(40, graph_steps=0)
        sa = a.state(infiltration=True, water_quality=True, groundwater=True)
        sb = b.state(infiltration=True, water_quality=True, groundwater=True)
        ra, rb = a.report(), b.report()
    assert ra["steps"] == rb["steps"] == 40
    for key in ("d_t", "I_t", "B_t", "P_conc", "Tot_Washed", "gw_h"):
        np.testing.assert_array_equal(sa[key], sb[key], err_msg=key)


def test_partial_runs_compose():
    from hydropol2d_amd.simulation import Simulation
    spec = loop_cases.plane(C3=3e5, min_Bt=5.0, every=6)
    with Simulation(spec) as s:
        for n in (5, 11, 17):
            s.run(n)
        st = s.state(water_quality=True, groundwater=True)
        assert s.report()["steps"] == 33
        dts = s.dt_history()
    o = oracle.run(spec, max_steps=33, dt_follow=dts)
    assert o["n_dt_mismatch"] == 0
    _close(st["B_t"], o["B_t"], "B_t")
    _close(st["gw_h"], o["gw_h"], "gw_h")
    _close(st["d_t"], o["d_t"], "d_t")


def test_fp32_loop_within_bounds_of_fp64_oracle():
    from hydropol2d_amd.simulation import Simulation
    spec = loop_cases.plane(C3=3e4, every=5)
    o = oracle.run(spec)
    with Simulation(loop_cases.plane(C3=3e4, every=5, precision="fp32")) as s:
        s.run()
        st = s.state(water_quality=True, groundwater=True)
        rep = s.report()
    err_d = np.nanmax(np.abs(st["d_t"] - o["d_t"])) / 1000.0
    # the head moves by little per call; its change since the start is what
    # the FP32 run's groundwater passes (two-pass sub-steps with reciprocal
    # multiplies) compute, so it is held to the FP64 change tightly
    dh_o = o["gw_h"] - spec.groundwater.h0
    dh = st["gw_h"] - spec.groundwater.h0
    err_dh = np.nanmax(np.abs(dh - dh_o)) / np.nanmax(np.abs(dh_o))
    print(f"fp32 loop: depth {err_d:.2e} m, B_t {np.nanmax(np.abs(st['B_t'] - o['B_t'])):.2e} kg, "
          f"head {np.nanmax(np.abs(st['gw_h'] - o['gw_h'])):.2e} m, head change {err_dh:.2e} of "
          f"{np.nanmax(np.abs(dh_o)):.2e} m", flush=True)
    assert err_d <= 1e-3
    assert np.nanmax(np.abs(st["gw_h"] - o["gw_h"])) <= 1e-3
    assert np.nanmax(np.abs(dh_o)) > 0 and err_dh <= 1e-3
    assert rep["wq_mass_kg"] == pytest.approx(o["report"]["wq_mass_kg"], rel=1e-3)


def test_ensemble_refuses_map_stack_members():
    # members with wash-off and groundwater are batched (tests/test_gpu_ensemble.py);
    # map stacks are not
    from hydropol2d_amd._lib import HP2DError
    from hydropol2d_amd.simulation import Ensemble
    spec = loop_cases.plane(C3=3e4, gw=False)
    spec.record_time_maps_min = 1.0
    with pytest.raises(HP2DError):
        Ensemble([spec])
