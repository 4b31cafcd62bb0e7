"""Pins of the FP64 oracle's step loop with pollutant wash-off
(HydroPol2D_Main_While.m:683-695, 724-731, 919-922) and the groundwater module
called every N accepted steps: closed-basin pollutant and water budgets, the
dry-start dt lock that tmin_wq imposes, the call count and interval of the
groundwater module, and the error paths of the wash-off."""
import numpy as np
import pytest

import loop_cases
from oracle import oracle


def test_closed_basin_pollutant_mass_is_conserved():
    # no outlet: every kg that leaves a cell enters a neighbour, so the only
    # change of the total is the clamp of negative masses (mass_lost)
    spec = loop_cases.plane(C3=2e5, min_Bt=5.0, gw=False, outlets=False)
    r = oracle.run(spec)
    rep = r["report"]
    m0 = loop_cases.initial_mass(spec)
    assert rep["wq_substep_steps"] > 0 and rep["wq_mass_lost_kg"] < 0
    total = np.sum(r["B_t"]) + rep["wq_mass_lost_kg"]
    assert total == pytest.approx(m0, rel=1e-12, abs=spec.shape[0] * spec.shape[1] * 1e-12)
    assert rep["wq_mass_kg"] == pytest.approx(np.sum(r["B_t"]), rel=1e-14)


def test_initial_buildup_matches_the_closed_form():
    spec = loop_cases.plane(gw=False, T=0.001)
    r = oracle.run(spec, max_steps=1)
    # one step of wash-off moves mass between cells and out of the outlet only
    assert np.sum(r["B_t"]) <= loop_cases.initial_mass(spec) * (1 + 1e-12)
    assert np.sum(r["B_t"]) > 0.9 * loop_cases.initial_mass(spec)


def test_dry_start_locks_dt_at_the_minimum():
    # with no water anywhere min(dW) >= 0, so tmin_wq = large_timestep = the
    # current dt (build_up_wash_off.m:61-62) and refreshing_timestep.m:329-332
    # keeps the next dt there; without the wash-off the dt grows
    spec = loop_cases.plane(d0=0.0, gw=False, T=0.5)
    spec.rain_intensity_mmh = np.array([0.0, 0.0])
    r = oracle.run(spec, max_steps=200)
    assert np.all(r["dt_hist"] == spec.min_time_step_s)
    spec.water_quality = None
    r2 = oracle.run(spec, max_steps=200)
    assert r2["dt_hist"][-1] == spec.max_time_step_s


def test_dt_never_exceeds_the_previous_tmin_wq():
    spec = loop_cases.plane(C3=6e4, gw=False)
    n = 12
    tmins, dts = [], []
    for k in range(1, n + 1):
        r = oracle.run(spec, max_steps=k)
        tmins.append(r["report"]["wq_tmin_s"])
        dts.append(r["dt_hist"][-1])
    # the dt chosen after step k is at most that step's tmin_wq
    assert all(dts[i] <= tmins[i] + 1e-12 for i in range(n))
    assert any(abs(dts[i] - np.floor(tmins[i] / spec.time_step_increments_s) *
                   spec.time_step_increments_s) < 1e-12 for i in range(n))


def test_groundwater_called_every_n_accepted_steps():
    spec = loop_cases.plane(wq=False, every=7)
    r = oracle.run(spec)
    steps = r["report"]["steps"]
    assert r["report"]["gw_updates"] == (steps - 1) // 7
    r10 = oracle.run(spec, max_steps=21)
    assert r10["report"]["gw_updates"] == 2  # at the starts of steps 8 and 15


def test_closed_basin_water_balance_with_groundwater():
    # no outlet, no Green-Ampt drainage sink (k_out = 0), a shallow water table
    # (5 cm; above the ground the recharge bucket of a cell without unsaturated
    # zone drops its storage, simulate_groundwater_recharge.m): rain = change of
    # surface + soil + aquifer storage
    spec = loop_cases.plane(wq=False, outlets=False, every=5, h0_offset=-0.05, seed=3)
    spec.infiltration.theta_sat_minus_r = np.array([0.0])
    r = oracle.run(spec)
    rep = r["report"]
    ny, nx = spec.shape
    A = spec.resolution ** 2
    gw = spec.groundwater
    S0 = (np.sum(spec.d0) + spec.infiltration.I_min_mm * ny * nx) / 1000 * A + \
        np.sum(A * gw.Sy * (gw.h0 - (spec.elevation - gw.soil_depth)))
    S1 = (np.nansum(r["d_t"]) + np.nansum(r["I_t"])) / 1000 * A + rep["gw_storage_m3"]
    S_gw0 = np.sum(A * gw.Sy * (gw.h0 - (spec.elevation - gw.soil_depth)))
    assert rep["gw_updates"] > 5 and rep["gw_storage_m3"] != pytest.approx(S_gw0, rel=1e-9)
    assert S1 - S0 == pytest.approx(rep["cum_rain_m3"], rel=1e-11)


def test_wash_off_errors_reject_steps():
    # a wash-off strong enough to clamp large negative masses trips the
    # 1.2 x initial-mass check and the tmin_wq check: the catch block runs
    # (each rejection shrinks the next dt by 1/catch_index and the splits grow:
    # the run is cut after 300 accepted steps)
    spec = loop_cases.plane(C3=1e6, min_Bt=5.0, gw=True)
    r = oracle.run(spec, max_steps=300)
    rep = r["report"]
    assert rep["rejected_steps"] > 0
    assert rep["catch_index"] == 1 + rep["rejected_steps"]


def test_runaway_split_is_refused():
    # tmin_wq -> 0 would need more than 1e7 internal sub-steps (the reference
    # would not terminate): the oracle refuses the run
    spec = loop_cases.plane(C3=1e6, min_Bt=0.0, gw=False)
    with pytest.raises(RuntimeError):
        oracle.run(spec)


def test_empty_cells_put_tmin_wq_on_the_dt_grid():
    # a cell washed out to zero that then receives a steady inflow has
    # B_t_mid = |dW| dt / 3600, so its tmin_wq candidate is the step itself up
    # to rounding: the floor() snap of the next dt sits on a tie at the level of
    # the last bits, which any two implementations may break differently
    spec = loop_cases.plane(C3=3e5, min_Bt=5.0)
    inc = spec.time_step_increments_s
    near = 0
    for k in range(18, 30):
        t = oracle.run(spec, max_steps=k)["report"]["wq_tmin_s"]
        u = t / inc
        near += abs(u - round(u)) < 1e-9
    assert near >= 2


def test_dt_follow_adopts_only_ties():
    spec = loop_cases.plane(C3=3e5, min_Bt=5.0)
    own = oracle.run(spec)
    same = oracle.run(spec, dt_follow=own["dt_hist"])
    assert same["n_dt_ties"] == 0 and same["n_dt_mismatch"] == 0
    np.testing.assert_array_equal(same["B_t"], own["B_t"])
    bad = own["dt_hist"].copy()
    bad[3] += 0.5  # not a tie: refused and counted
    r = oracle.run(spec, dt_follow=bad)
    assert r["n_dt_mismatch"] == 1 and r["n_dt_ties"] == 0
    np.testing.assert_array_equal(r["dt_hist"], own["dt_hist"])


def test_dt_force_adopts_every_dt():
    # dt_force: the oracle runs on a given dt sequence (an FP32 run's) and
    # counts where its own refreshed dt differed
    spec = loop_cases.plane(C3=3e5, min_Bt=5.0)
    own = oracle.run(spec, max_steps=40)
    given = own["dt_hist"].copy()
    given[5:] = np.maximum(given[5:] - 2 * spec.time_step_increments_s, spec.min_time_step_s)
    r = oracle.run(spec, max_steps=40, dt_follow=given, dt_force=True)
    np.testing.assert_array_equal(r["dt_hist"], given)
    assert r["n_dt_mismatch"] >= 1
