"""Pins of the FP64 CPU oracle's local-inertial step (A1-A8) on closed forms,
invariants, the reference's validation cases and its golden fixtures.

Each check is chosen so that a plausible restatement mistake (a dropped term,
a wrong sign or index, a wrong exponent) fails it.
"""
import csv
import math
import os

import numpy as np
import pytest

import li_cases
from oracle import oracle

GOLD = os.path.join(os.path.dirname(__file__), "golden")
G = 9.81


def _read_csv(name):
    with open(os.path.join(GOLD, name)) as f:
        return list(csv.DictReader(f))


# ---------------------------------------------------------------- A1 Bates ----

def test_bates_frictionless_is_explicit_momentum():
    rng = np.random.default_rng(0)
    q_p, Hf, S = rng.normal(size=(3, 4)), rng.uniform(0.1, 2, (3, 4)), rng.normal(0, 0.01, (3, 4))
    dt = 0.5
    q = oracle.inertial_solver(q_p, dt, Hf, S, np.zeros((3, 4)))
    np.testing.assert_allclose(q, q_p - G * Hf * dt * S, rtol=1e-15, atol=1e-15)


def test_bates_normal_flow_is_fixed_point():
    # Manning normal flow q_n = Hf^(5/3) sqrt(|S|)/n is a fixed point of the
    # semi-implicit update (any other friction exponent breaks this).
    Hf = np.array([[0.05, 0.3, 1.7]])
    S = -np.array([[1e-3, 5e-3, 2e-2]])      # downhill towards +x
    n = np.array([[0.02, 0.035, 0.1]])
    qn = Hf ** (5 / 3) * np.sqrt(-S) / n
    for dt in (0.1, 1.0, 30.0):
        q = oracle.inertial_solver(qn, dt, Hf, S, n ** 2)
        np.testing.assert_allclose(q, qn, rtol=1e-12)


def test_bates_dry_face_gives_nonfinite_then_zero():
    q = oracle.inertial_solver(np.array([[0.0, 0.3]]), 1.0, np.zeros((1, 2)), np.zeros((1, 2)),
                               np.full((1, 2), 0.03 ** 2))
    assert np.isnan(q[0, 0])           # 0/0, zeroed by Local_Inertial_Model_D4.m:264
    assert q[0, 1] == 0.0              # q_p / Inf


# ---------------------------------------------------- A2-A7 invariants ----

def _li(z, d_mm, n, dt_s, dx, outflow=None, rows=(), cols=(), outlet_type=2, slope=0.0,
        critical=0):
    ny, nx = z.shape
    if outflow is None:
        outflow = np.zeros((ny, nx, 3))
    return oracle.li_d4(z, d_mm, n, n ** 2, dx * dx, dt_s / 60, dx, outlet_type, slope,
                        np.asarray(rows), np.asarray(cols), outflow, critical)


def test_lake_at_rest_is_exact():
    # z on a 1/8 m lattice so that d/1000 and z + d/1000 are exact: the
    # water surface is then exactly flat and every slope is exactly zero.
    rng = np.random.default_rng(1)
    z = rng.integers(0, 8, (12, 9)) / 8.0
    wse = 1.5
    d = (wse - z) * 1000
    n = np.full(z.shape, 0.03)
    out = _li(z, d, n, 2.0, 5.0)
    assert np.all(out["outflow"] == 0)
    np.testing.assert_array_equal(out["d_t"], d)


def test_closed_domain_conserves_volume():
    rng = np.random.default_rng(2)
    ny, nx, dx = 20, 15, 4.0
    z = rng.uniform(0, 0.3, (ny, nx))
    d = rng.uniform(0, 200, (ny, nx))
    d[5:8, 3:6] = 0.0      # dry patch: exercises the draining limiter
    n = np.full((ny, nx), 0.025)
    state = np.zeros((ny, nx, 3))
    v0 = d.sum()
    for _ in range(200):
        out = _li(z, d, n, 0.2, dx, outflow=state)
        d, state = out["d_t"], out["outflow"]
        assert d.min() >= 0
    assert abs(d.sum() - v0) / v0 < 1e-12


def test_draining_limiter_never_creates_mass():
    # a thin film next to a steep drop with a huge inherited flux: without the
    # donor limiter the cell would go negative and the d<0 clamp would add mass.
    ny, nx, dx = 3, 4, 1.0
    z = np.array([[1.0, 0.0, 0.0, 0.0]] * 3)
    d = np.array([[5.0, 0.0, 0.0, 0.0]] * 3)
    n = np.full((ny, nx), 0.01)
    prev = np.zeros((ny, nx, 3))
    prev[:, 0, 0] = 5e6     # mm/h, far beyond the available 5 mm
    out = _li(z, d, n, 5.0, dx, outflow=prev)
    assert abs(out["d_t"].sum() - d.sum()) < 1e-9
    assert out["d_t"][:, 0].min() >= 0


def test_flux_sign_conventions():
    # x-face flux positive from (i,j) to (i,j+1); y-face (stored at row i)
    # positive from (i,j) to (i-1,j) (Local_Inertial_Model_D4.m:123-127,313-315)
    z = np.zeros((3, 3))
    d = np.zeros((3, 3))
    d[1, 1] = 100.0
    n = np.full((3, 3), 0.03)
    out = _li(z, d, n, 0.5, 2.0)
    ox, oy = out["outflow"][:, :, 0], out["outflow"][:, :, 1]
    assert ox[1, 1] > 0 and ox[1, 0] < 0          # spreads right and left
    assert oy[1, 1] > 0 and oy[2, 1] < 0          # spreads up (to row 0) and down
    np.testing.assert_allclose(out["qout_left"][1, 1], -ox[1, 0])
    np.testing.assert_allclose(out["qout_down"][1, 1], -oy[2, 1])
    # symmetric spreading
    np.testing.assert_allclose(ox[1, 1], -ox[1, 0], rtol=1e-14)
    np.testing.assert_allclose(oy[1, 1], -oy[2, 1], rtol=1e-14)


def test_outlet_critical_closed_form():
    # single cell, no faces: Q = dx*sqrt(g)*h^(3/2) (critical-depth outlet)
    for h_mm in (1.0, 50.0, 800.0):
        z = np.zeros((1, 1))
        d = np.array([[h_mm]])
        n = np.array([[0.05]])
        dt_s, dx = 0.01, 10.0
        out = _li(z, d, n, dt_s, dx, rows=[1], cols=[1], outlet_type=2)
        h = h_mm / 1000
        q_mmh = dx * math.sqrt(G) * h ** 1.5 / dx ** 2 * 1000 * 3600
        q_mmh = min(q_mmh, h_mm / (dt_s / 3600))
        np.testing.assert_allclose(out["outlet_flow"][0, 0], q_mmh, rtol=1e-12)
        np.testing.assert_allclose(out["d_t"][0, 0], h_mm - q_mmh * dt_s / 3600, rtol=1e-12,
                                   atol=1e-12)
        assert out["Hf"][0, 0, 2] == h


def test_outlet_normal_closed_form_and_cap():
    z = np.zeros((1, 1))
    n = np.array([[0.03]])
    dx, S = 5.0, 0.002
    out = _li(z, np.array([[200.0]]), n, 0.1, dx, rows=[1], cols=[1], outlet_type=1, slope=S)
    h = 0.2
    q = (1 / 0.03) * dx * h ** (5 / 3) * math.sqrt(S) / dx ** 2 * 1000 * 3600
    np.testing.assert_allclose(out["outlet_flow"][0, 0], q, rtol=1e-12)
    # cap: never drains below zero
    out = _li(z, np.array([[0.5]]), n, 600.0, dx, rows=[1], cols=[1], outlet_type=1, slope=0.5)
    assert out["d_t"][0, 0] == pytest.approx(0.0, abs=1e-12)


def test_nan_cells_are_walls():
    z = np.zeros((4, 4))
    z[:, 0] = np.nan
    d = np.full((4, 4), 100.0)
    d[:, 0] = np.nan
    d[1, 2] = 300.0
    out = _li(z, d, np.full((4, 4), 0.03), 0.5, 2.0)
    assert np.all(np.isnan(out["d_t"][:, 0]))
    assert np.all(out["outflow"][:, 0, 0] == 0)     # face between NaN column and column 1
    np.testing.assert_allclose(np.nansum(out["d_t"]), np.nansum(d), rtol=1e-13)


def test_critical_cap_bounds_flux():
    z = np.array([[0.0, -5.0]])
    d = np.array([[400.0, 0.0]])
    n = np.zeros((1, 2))
    out = _li(z, d, n, 1.0, 1.0, critical=1)
    Hf = out["Hf"][0, 0, 0]
    q = out["outflow"][0, 0, 0] / 1000 / 3600 * 1.0   # m2/s (dx^2/dx = 1)
    assert q <= Hf * math.sqrt(G * Hf) * (1 + 1e-12)


# ------------------------------------------------------------- A8 dt -------

def test_refresh_dt_hand_computed_1x3():
    # one row: x-velocities via the unshifted vel_left quirk (refreshing_timestep.m:73)
    dx = 2.0
    qr = np.array([[3600.0, 7200.0, 0.0]])          # mm/h
    qu = np.zeros((1, 3))
    Hf = np.zeros((1, 3, 3))
    Hf[0, :, 0] = [0.1, 0.2, 0.0]
    tcr, vmax, raster = oracle.refresh_dt(qr, qu, Hf, dx, alfa_min=0.5, alfa_max=0.7,
                                          max_time_step=10.0, min_time_step=0.01,
                                          time_step_increments=0.001, depth_tolerance_mm=1.0)
    vr = qr[0] / 1000 / 3600 * dx / np.array([0.1, 0.2, 1e9])   # Hf=0 -> 1e12 mm
    # right component: j=1 -> 0.5*vr (vel_left = 0), j>=2 -> vr (vel_left = -vr, unshifted)
    comp = np.array([0.5 * vr[0], vr[1], vr[2]])
    assert vmax == pytest.approx(comp.max(), rel=1e-15)
    dt_adv = dx / comp.max() * max(0.7 - 2 * max(comp.max() - 5, 0), 0.5)
    # wave: non-stagnant cells only (|v| >= 0.01)
    c = np.sqrt(G * Hf[0, :, 0])
    moving = comp >= 0.01
    dt_wave = np.min(0.5 * dx / c[moving])
    expect = max(0.001 * math.floor(min(dt_adv, dt_wave) / 0.001), 0.01)
    assert tcr == pytest.approx(min(expect, 10.0), rel=1e-15)
    np.testing.assert_allclose(raster[0], np.abs(vr))


def test_refresh_dt_all_stagnant_uses_max_dt():
    tcr, vmax, _ = oracle.refresh_dt(np.zeros((2, 2)), np.zeros((2, 2)), np.zeros((2, 2, 3)), 5.0,
                                     max_time_step=7.5)
    assert vmax == 0.0 and tcr == 7.5


# ---------------------------------------------------------- A18 rain ------

def test_rain_average_integrates_exactly():
    times = np.array([30.0, 60.0, 90.0])
    vals = np.array([50.0, 10.0, 0.0])
    total, cur, t = 0.0, 1, 0.0
    rng = np.random.default_rng(3)
    while t < 100:
        t1 = t + rng.uniform(0.01, 7.0)
        avg, cur = oracle.rain_avg(times, vals, t, t1, cur)
        total += avg * (t1 - t) / 60
        t = t1
    exact = (50 * 30 + 10 * 30) / 60
    assert total == pytest.approx(exact, rel=1e-12)


# ----------------------------------- reference validation cases (LI) -------

def test_plane_case_meets_reference_thresholds():
    r = li_cases.plane_case(li_cases.oracle_li_step())
    # run_...validation.m:112-116
    assert r["mass_abs_pct"] < 0.1
    assert r["nse"] > 0.70
    assert r["rel_l2"] < 0.45


def test_vtilted_case_meets_reference_thresholds():
    r = li_cases.vtilted_case(li_cases.oracle_li_step())
    # run_...validation.m:195-197
    assert r["mass_abs_pct"] < 0.1
    assert r["symmetry_max"] < 1e-8
    assert r["channel_final"] > 0


def test_stage_case_meets_reference_thresholds():
    r = li_cases.stage_case(li_cases.oracle_li_step())
    assert r["max_stage_error"] < 1e-12
    assert r["mass_abs_pct"] < 0.1
    assert r["response_amp"] > 1e-4


def test_ritter_case_conserves_mass_and_reports_error():
    # Registered diagnostic_failed for LI (Case_Config.csv:4): gate on mass only.
    r = li_cases.ritter_case(li_cases.oracle_li_step())
    assert r["mass_abs_pct"] < 0.1
    assert np.isfinite(r["rmse"]) and r["rmse"] < 0.5


# ------------------------------------------------------ golden fixtures ----

def test_kinematic_wave_fixture_matches_closed_form():
    rows = _read_csv("plane_kinematic_wave_hydrograph.csv")
    rain = 100 / 1000 / 3600
    for r in rows:
        q = li_cases.kinematic_plane_q(float(r["time_sec"]), rain, 2000.0, 1000.0, 0.015, 0.01)
        assert q == pytest.approx(float(r["Q_analytic_m3s"]), rel=1e-9, abs=1e-9)


def test_ritter_fixture_matches_closed_form():
    rows = _read_csv("ritter_analytical_profiles.csv")
    x = np.array([float(r["x_m"]) for r in rows])
    for t in (10, 20, 30, 40, 50, 60):
        h = np.array([float(r[f"h_t{t:03d}s_m"]) for r in rows])
        np.testing.assert_allclose(li_cases.ritter_depth(x, float(t), 1.0, 250.0), h,
                                   rtol=1e-9, atol=1e-12)


@pytest.mark.slow
def test_kinematic_wave_fixture_plane_on_oracle():
    # the reference's own benchmark plane (compare_flows.m:10-33), 2000 m long
    # x 1000 m wide, at dx = 50 m; the LI oracle's outlet hydrograph against the
    # golden kinematic-wave hydrograph.
    rows = _read_csv("plane_kinematic_wave_hydrograph.csv")
    t_ref = np.array([float(r["time_sec"]) for r in rows])
    q_ref = np.array([float(r["Q_analytic_m3s"]) for r in rows])
    dx, S0, nman = 50.0, 0.01, 0.015
    ny, nx = int(2000 / dx), int(1000 / dx)       # flow down the rows
    i = (np.arange(ny) + 0.5) * dx
    z = np.tile((2000 - i)[:, None] * S0, (1, nx))
    mask, rows_o, cols_o = li_cases.outlet_cells(ny, nx, "down", np.ones(nx, dtype=bool))
    step = li_cases.oracle_li_step()
    h = np.zeros((ny, nx))
    state = None
    rain = 100 / 1000 / 3600
    dt = 1.0
    q_model = {0: 0.0}
    for it in range(1, int(t_ref.max()) + 1):
        h = h + rain * dt
        d_t, oflow, state = step(z, h * 1000, np.full((ny, nx), nman), dt, dx, state,
                                 (mask, rows_o, cols_o), S0)
        h = np.maximum(d_t / 1000, 0)
        if it % 600 == 0:
            q_model[it] = float(np.sum(oflow)) / 1000 / 3600 * dx * dx
    qm = np.array([q_model[int(t)] for t in t_ref])
    assert li_cases.nse(qm[1:], q_ref[1:]) > 0.9


def test_duplicate_outlets_collapse_like_vectorised_assignment():
    # d_t(outlet_sub) = d_t(outlet_sub) - q_out*dt_h with a repeated index
    # subtracts once (Local_Inertial_Model_D4.m:496)
    z = np.zeros((1, 3))
    d = np.array([[0.0, 0.0, 120.0]])
    n = np.full((1, 3), 0.04)
    once = _li(z, d, n, 0.3, 2.0, rows=[1], cols=[3])
    twice = _li(z, d, n, 0.3, 2.0, rows=[1, 1], cols=[3, 3])
    np.testing.assert_array_equal(once["d_t"], twice["d_t"])
    np.testing.assert_array_equal(once["outlet_flow"], twice["outlet_flow"])


# ------------------------------------------ A16 CA routing (CA_Routing.m) ----

def test_ca_plane_and_vtilted_meet_reference_thresholds():
    step = li_cases.oracle_ca_step()
    r = li_cases.plane_case(step)
    # run_...validation.m:112-116 (CA branch): mass < 0.1 % and some outflow
    assert r["mass_abs_pct"] < 0.1 and r["q_model"].max() > 0
    r = li_cases.vtilted_case(step)
    assert r["mass_abs_pct"] < 0.1
    assert r["symmetry_max"] < 1e-8
    assert r["channel_final"] > 0


def test_ca_stage_and_ritter_meet_reference_thresholds():
    step = li_cases.oracle_ca_step()
    r = li_cases.stage_case(step)
    assert r["max_stage_error"] < 1e-12 and r["mass_abs_pct"] < 0.1 and r["response_amp"] > 1e-4
    r = li_cases.ritter_case(step, n_manning=0.003)
    assert r["mass_abs_pct"] < 0.1 and r["rmse"] < 0.50


def test_ca_flat_pool_does_not_move():
    # equal water surfaces: every depth difference is below h_min -> no transfer
    z = np.array([[0.0, 0.25, 0.5], [0.5, 0.0, 0.25]])
    d = (1.0 - z) * 1000
    n = np.full(z.shape, 0.03)
    out = oracle.ca_routing(z, d, n, 4.0, 1 / 60, None, 2.0, np.zeros(z.shape), 1, 0.0, [], [], 0)
    assert np.all(out["I_tot_end_cell"] == 0)
    np.testing.assert_array_equal(out["d_t"], d)


def test_ca_weights_split_by_available_volume():
    # single wet cell with two lower neighbours: the outflow splits in
    # proportion to the depth differences (weights V_k / (sum V + V_min))
    z = np.array([[0.0, 0.0, 0.0]])
    d = np.array([[0.0, 200.0, 100.0]])
    n = np.full(z.shape, 0.03)
    out = oracle.ca_routing(z, d, n, 1.0, 0.5 / 60, None, 1.0, np.zeros(z.shape), 1, 0.0, [], [], 0)
    ql, qr = out["qout_left"][0, 1], out["qout_right"][0, 1]
    assert ql > 0 and qr > 0
    assert ql / qr == pytest.approx(0.2 / 0.1, rel=1e-12)
