"""Pins of the FP64 CPU oracle's Boussinesq groundwater solver (A14) and the
vadose recharge bucket (A13) on the reference's hillslope closed forms,
invariants and the sub-step schedule.

A plausible restatement mistake fails one of them: a wrong face average or
donor, a lost Sy, a divergence sign or boundary slip (steady Dupuit
equilibrium to 1e-5 m, linearised decay rate, toe flux), a mis-scheduled
sub-step (sub-step counts), a broken mass ledger (closed-box conservation).
"""
import math

import numpy as np
import pytest

import gw_cases
from oracle import oracle


def test_steady_dupuit_hillslope():
    # run_hillslope_analytical_validation.m:42-110 thresholds (:99-103)
    res, ratio, passed = gw_cases.steady_dupuit(gw_cases.oracle_solver)
    fine = res[-1]
    assert fine["eq_err"] < 1e-5, fine["eq_err"]
    assert fine["rmse"] < 1e-2 and fine["max_err"] < 2e-2
    assert fine["q_err_pct"] < 0.1
    assert ratio > 2
    assert passed


def test_predictor_omits_recharge():
    # The predictor of Boussinesq_2D_explicit.m:113 moves h by -div only (no
    # recharge), so the discrete Dupuit equilibrium (div = R) is not an exact
    # fixed point: after 60 s it drifts by ~4e-7 m at 161 columns.  A predictor
    # that included R would keep it to round-off (< 1e-12).
    res, _, _ = gw_cases.steady_dupuit(gw_cases.oracle_solver, resolutions=(161,))
    assert 1e-8 < res[0]["eq_err"] < 1e-5, res[0]["eq_err"]


def test_transient_linearized_decay():
    # :112-196 thresholds (:189-190): max RMSE < 0.02 m, max error < 0.05 m,
    # toe-flux NSE > 0.99, peak flux error < 5 %
    s, passed = gw_cases.transient_linearized(gw_cases.oracle_solver)
    assert s["rmse"].max() < 0.02
    assert s["max_err"].max() < 0.05
    assert s["nse"] > 0.99, s["nse"]
    assert s["peak_err_pct"] < 5
    assert passed


def test_transient_decay_rate():
    # the cosine mode decays as exp(-D lambda^2 t), D = K H / Sy: over 5 days
    # the model's mode amplitude follows it to 2 %
    cfg = gw_cases.TRANSIENT
    s, _ = gw_cases.transient_linearized(gw_cases.oracle_solver, duration_s=5 * 86400)
    q = s["q_model"]
    lam = math.pi / (2 * cfg["length_m"])
    D = cfg["K_m_s"] * cfg["base_head_m"] / cfg["Sy"]
    ratio_model = q[-1] / q[0]
    ratio_exact = math.exp(-D * lam ** 2 * 5 * 86400)
    assert abs(ratio_model / ratio_exact - 1) < 0.02


def _box(ny=12, nx=15, seed=0):
    rng = np.random.default_rng(seed)
    z0 = rng.uniform(0, 1.0, (ny, nx))
    h = z0 + rng.uniform(0.5, 3.0, (ny, nx))
    K = rng.uniform(1e-5, 5e-4, (ny, nx))
    Sy = rng.uniform(0.1, 0.3, (ny, nx))
    return z0, h, K, Sy


def test_closed_box_conserves_storage():
    # no recharge, no Dirichlet cells, deep soil: sum(Sy (h - z0)) A is
    # conserved by the flux form whatever the number of sub-steps
    z0, h, K, Sy = _box()
    dx, dy = 5.0, 7.0
    r = oracle.boussinesq_2d(100.0, dx, dy, h, z0, Sy, 0.0, K, 0.25, 100.0)
    s0 = np.sum(Sy * (h - z0))
    s1 = np.sum(Sy * np.maximum(r["h_t"] - z0, 0))
    np.testing.assert_allclose(s1, s0, rtol=1e-12)
    assert abs(r["error"]) < 1e-9 * s0 * dx * dy


def test_single_substep_recharge_ledger():
    # one sub-step (T below the Courant limit): error = recharge - exfiltration
    # - storage change closes to round-off
    z0, h, K, Sy = _box(seed=1)
    R = np.full(h.shape, 2e-6)
    r = oracle.boussinesq_2d(10.0, 5.0, 5.0, h, z0, Sy, R, K * 1e-3, 0.25, 100.0)
    assert r["n_sub"] == 1
    s = np.sum(Sy * (h - z0)) * 25
    assert abs(r["error"]) < 1e-12 * s
    # and the head rose by R dt / Sy on average
    np.testing.assert_allclose(np.sum(Sy * (r["h_t"] - h)), np.sum(R * 10.0), rtol=1e-9)


def _dt_gw(h, K, dx, dy, courant):
    ux = -(0.5 * (K[:, :-1] + K[:, 1:])) * ((h[:, 1:] - h[:, :-1]) / dx)
    uy = -(0.5 * (K[:-1, :] + K[1:, :])) * ((h[1:, :] - h[:-1, :]) / dy)
    with np.errstate(divide="ignore"):
        c = np.concatenate([(courant * dx / np.abs(ux)).ravel(),
                            (courant * dy / np.abs(uy)).ravel()])
    c = c[np.isfinite(c) & (c > 0)]
    return c.min()


@pytest.mark.parametrize("ratio,expected", [(0.5, 1), (1.5, 2), (2.4, 5), (2.6, 4), (3.5, 4),
                                            (5.5, 6)])
def test_substep_schedule(ratio, expected):
    # Every cell a Dirichlet cell keeps the head, hence the velocities and
    # dt_GW, fixed, so only the loop logic sets the count for T = ratio*dt_GW:
    # ceil(ratio) sub-steps, except when ceil(T/dt) == 3 on the second pass
    # (2 < ratio <= 3): the trim sets the *next* dt to T - 2 dt and t advances
    # by it, so ratio 2.6 takes dt_GW, dt_GW, 0.6, 0.6 (4 sub-steps, 3.2 dt_GW
    # of integration) and ratio 2.4 takes 1, 1, 0.4, 0.4, 0.4 (5).
    ny, nx = 4, 30
    x = np.arange(nx) * 10.0
    h = np.tile(50.0 - 0.01 * x, (ny, 1))
    z0 = np.zeros((ny, nx))
    K = np.full((ny, nx), 1e-3)
    dtg = _dt_gw(h, K, 10.0, 10.0, 0.25)
    dm = np.ones((ny, nx), dtype=bool)
    r = oracle.boussinesq_2d(ratio * dtg, 10.0, 10.0, h, z0, 0.2, 0.0, K, 0.25, 100.0, None, dm, h)
    assert r["n_sub"] == expected
    np.testing.assert_array_equal(r["h_t"], h)


def test_exfiltration_and_surface_clamp():
    z0, h, K, Sy = _box(seed=2)
    h_soil = np.full(h.shape, 1.5)
    hs = z0 + h_soil
    r = oracle.boussinesq_2d(1.0, 5.0, 5.0, h, z0, Sy, 0.0, K * 1e-4, 0.25, h_soil)
    assert r["n_sub"] == 1
    above = h > hs + 1e-3
    assert above.any()
    np.testing.assert_array_less(r["h_t"], hs + 1e-12)
    # q_exf = (h - h_surf) Sy / dt where the head was above the surface
    np.testing.assert_allclose(r["q_exf"][above], (h[above] - hs[above]) * Sy[above] / 1.0,
                               rtol=1e-6)
    assert (r["q_exf"][h < hs - 1e-3] == 0).all()


def test_dirichlet_and_catchment_mask():
    z0, h, K, Sy = _box(seed=3)
    dm = np.zeros(h.shape, dtype=bool)
    dm[:, -1] = True
    hd = np.full(h.shape, 1.25)
    cm = np.ones(h.shape, dtype=bool)
    cm[0, :] = False
    r = oracle.boussinesq_2d(600.0, 5.0, 5.0, h, z0, Sy, 0.0, K, 0.25, 100.0, cm, dm, hd)
    assert np.isnan(r["h_t"][0, :]).all() and np.isnan(r["q_exf"][0, :]).all()
    np.testing.assert_array_equal(r["h_t"][1:, -1], np.maximum(1.25, z0[1:, -1]))
    assert np.isnan(r["u_x"][:, -1]).all() and np.isnan(r["u_y"][-1, :]).all()
    # a head of NaN next to a cell gives that cell a NaN flux, and max(NaN, z0)
    # resets it to the aquifer base (cells bordering the masked row)
    r2 = oracle.boussinesq_2d(600.0, 5.0, 5.0, r["h_t"], z0, Sy, 0.0, K, 0.25, 100.0, cm, dm, hd)
    np.testing.assert_array_equal(r2["h_t"][1, :-1], z0[1, :-1])


def test_recharge_bucket_closed_forms():
    # simulate_groundwater_recharge.m: saturated bucket (Se = 1) drains at Ks
    # with unit gradient; a dry bucket at Se floor gives the capped upward
    # gradient -1; no unsaturated zone gives no exchange; mass balance
    # R_mb = f - dS/dt holds in every case
    Ks = 36.0                                   # mm/h = 1e-5 m/s
    ts, tr, nv, a = 0.45, 0.05, 1.6, 3.0
    max_S = np.array([0.4, 0.4, 0.0, 0.4])      # zwt = 1 m for the UZ cells
    S0 = np.array([0.4, 0.0, 0.0, 0.1])
    f = np.array([0.0, 0.0, 0.0, 1e-6])
    dt = 60.0
    R, S, cum = oracle.gw_recharge(f, S0, dt, 0.0, max_S, ts, tr, nv, a, Ks)
    np.testing.assert_allclose(R[0], 1e-5, rtol=1e-12)                 # Kr = 1, grad = 1
    assert R[1] == 0.0 and S[1] == 0.0                                    # Kr ~ 0 at the floor
    assert R[2] == 0.0 and S[2] == 0.0
    np.testing.assert_allclose(R, f - (S - S0) / dt, rtol=1e-12, atol=1e-18)
    np.testing.assert_allclose(cum, np.maximum(R, 0) * dt * 1000, rtol=1e-15)
    # van Genuchten head of the half-full bucket: h = -(1/a)((Se^(-1/m) - 1)^(1/n))
    m = 1 - 1 / nv
    Se = (0.1 / 1.0) / (ts - tr)
    h = -(1 / a) * (Se ** (-1 / m) - 1) ** (1 / nv)
    Kr = Se ** 0.5 * (1 - (1 - Se ** (1 / m)) ** m) ** 2
    grad = min(max(1 + h / 0.5, -1), 2)
    R3 = 1e-5 * Kr * grad
    np.testing.assert_allclose(R[3], f[3] - (max(min(0.1 + dt * (f[3] - R3), 0.4), 0) - 0.1) / dt,
                               rtol=1e-12)
