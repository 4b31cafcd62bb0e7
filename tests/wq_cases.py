"""Replay of the reference's wash-off validation driver with an injected routine.

Restates the case set-ups and closed forms of
Validation/WaterQuality/BuildUpWashOff_MassBalance/run_water_quality_washoff_validation.m
(parameters :51-76, dry build-up :78-103, single cell :105-180, two cells
:182-252, the routine calls :254-276) so the same harness drives the CPU
oracle and the HIP library.

washoff(C3, C4, ql, qr, qu, qd, qo, B_t, time_step_min, cell_area, nan_mask,
        flag_wq_model, mass_lost, Tot_Washed, Bmin, Bmax, min_Bt) -> dict with
        keys B_t, P_conc, Out_Conc, tmin_wq, tot_W_out, mass_lost, Tot_Washed
"""
from __future__ import annotations

import math

import numpy as np


def base_parameters():
    """run_water_quality_washoff_validation.m:51-76."""
    P = dict(area_m2=400.0, runoff_mm_h=30.0, washoff_exponent=1.0, max_buildup_kg_ha=56.04,
             buildup_rate_per_day=1.0, antecedent_dry_days=5.0,
             target_initial_concentration_mg_L=120.0, dt_min=5e-4, duration_min=15.0,
             record_dt_min=0.5, min_Bt_g_m2=0.0, Bmin_g_m2=0.0, Bmax_g_m2=0.0)
    P["initial_buildup_kg"] = (P["max_buildup_kg_ha"] *
                               (1 - math.exp(-P["buildup_rate_per_day"] * P["antecedent_dry_days"])) *
                               P["area_m2"] / 10000)
    P["runoff_m3_s"] = P["runoff_mm_h"] / 1000 / 3600 * P["area_m2"]
    P["lambda_h"] = (P["target_initial_concentration_mg_L"] * (P["runoff_mm_h"] * P["area_m2"]) /
                     1e6 / P["initial_buildup_kg"])
    P["washoff_coefficient"] = P["lambda_h"] / (P["runoff_m3_s"] ** P["washoff_exponent"])
    return P


def oracle_washoff(*args, **kw):
    from oracle import oracle
    return oracle.washoff(*args, **kw)


def single_cell(washoff, P=None, duration_min=None):
    """:105-180; returns (t_min, model_mass, reference_mass, out_conc, residual)."""
    P = P or base_parameters()
    dur = P["duration_min"] if duration_min is None else duration_min
    n_steps = round(dur / P["dt_min"])
    rec_every = max(1, round(P["record_dt_min"] / P["dt_min"]))
    B = np.array([[P["initial_buildup_kg"]]])
    ml, tw = 0.0, np.zeros((1, 1))
    z = np.zeros((1, 1))
    qo = np.full((1, 1), P["runoff_mm_h"])
    rows = [(0.0, B[0, 0], P["target_initial_concentration_mg_L"])]
    for it in range(1, n_steps + 1):
        r = washoff(P["washoff_coefficient"], P["washoff_exponent"], z, z, z, z, qo, B, P["dt_min"],
                    P["area_m2"], None, 0, ml, tw, P["Bmin_g_m2"], P["Bmax_g_m2"], P["min_Bt_g_m2"])
        B, ml, tw = r["B_t"], r["mass_lost"], r["Tot_Washed"]
        if it % rec_every == 0:
            rows.append((it * P["dt_min"], B[0, 0], r["Out_Conc"]))
    t = np.array([x[0] for x in rows])
    m = np.array([x[1] for x in rows])
    ref = P["initial_buildup_kg"] * np.exp(-P["lambda_h"] * t / 60)
    export = np.maximum(P["initial_buildup_kg"] - m, 0)
    residual = m + export - P["initial_buildup_kg"]
    return t, m, ref, np.array([x[2] for x in rows]), residual


def two_cell(washoff, P=None, duration_min=None):
    """:182-252; returns (t_min, B1, B2, ref1, ref2, export, ref_export, residual)."""
    P = P or base_parameters()
    dur = P["duration_min"] if duration_min is None else duration_min
    n_steps = round(dur / P["dt_min"])
    rec_every = max(1, round(P["record_dt_min"] / P["dt_min"]))
    B0 = P["initial_buildup_kg"]
    B = np.array([[B0, B0]])
    init_total = B.sum()
    ml, tw = 0.0, np.zeros((1, 2))
    z = np.zeros((1, 2))
    qr = np.array([[P["runoff_mm_h"], 0.0]])
    qo = np.array([[0.0, P["runoff_mm_h"]]])
    C3 = P["washoff_coefficient"] * np.ones((1, 2))
    C4 = P["washoff_exponent"] * np.ones((1, 2))
    rows = [(0.0, B[0, 0], B[0, 1])]
    for it in range(1, n_steps + 1):
        r = washoff(C3, C4, z, qr, z, z, qo, B, P["dt_min"], P["area_m2"], None, 0, ml, tw,
                    P["Bmin_g_m2"], P["Bmax_g_m2"], P["min_Bt_g_m2"])
        B, ml, tw = r["B_t"], r["mass_lost"], r["Tot_Washed"]
        if it % rec_every == 0:
            rows.append((it * P["dt_min"], B[0, 0], B[0, 1]))
    t = np.array([x[0] for x in rows])
    b1 = np.array([x[1] for x in rows])
    b2 = np.array([x[2] for x in rows])
    th = t / 60
    ref1 = B0 * np.exp(-P["lambda_h"] * th)
    ref2 = np.exp(-P["lambda_h"] * th) * (B0 + P["lambda_h"] * B0 * th)
    export = np.maximum(init_total - (b1 + b2), 0)
    ref_export = init_total - ref1 - ref2
    residual = b1 + b2 + export - init_total
    return t, b1, b2, ref1, ref2, export, ref_export, residual


def fixture_buildup_washoff(washoff, rows, dt_s=1.0, area=400.0):
    """Continuous build-up b [kg/s] and wash-off k*Q*B [kg/s] of
    Validation/Reference_Outputs/Phase1/washoff/P1-WQ-001_reference.csv,
    driven through the routine: one cell, mass model, C4 = 1, outlet
    discharge Q = runoff_reference [m3/s] expressed in mm/h over `area`,
    C3 = 3600*k so W = 3600*k*Q*B [kg/h].  Build-up b*dt is added after each
    call.  The split Euler step B <- B(1 - kQ dt) + b dt has exactly the
    analytic steady state b/(kQ).  Returns the model mass at each row time.
    """
    k = float(rows[0]["washoff_coefficient"])
    Q = float(rows[0]["runoff_reference"])
    b = float(rows[0]["build_rate_kg_s"])
    B = np.array([[float(rows[0]["surface_mass_kg"])]])
    q_mmh = Q * 3.6e6 / area
    z = np.zeros((1, 1))
    qo = np.full((1, 1), q_mmh)
    times = [float(r["time_s"]) for r in rows]
    out, t, ml, tw = [], 0.0, 0.0, np.zeros((1, 1))
    for target in times:
        while t < target - 1e-9:
            r = washoff(3600.0 * k, 1.0, z, z, z, z, qo, B, dt_s / 60.0, area, None, 0, ml, tw,
                        0.0, 0.0, 0.0)
            B, ml, tw = r["B_t"] + b * dt_s, r["mass_lost"], r["Tot_Washed"]
            t += dt_s
        out.append(B[0, 0])
    return np.array(out)


def substep_chain_closed_form(B0, lam_h, T_s):
    """One call on a 1-row chain of cells: cell c washes off at lam_h[c] [1/h]
    into cell c+1 (the last one into the outlet), time step T_s [s].

    tmin_wq (build_up_wash_off.m:61-76) only counts cells that gain mass (dW <
    0); if none does it is the whole step.  When B - dW*T < 0 somewhere or
    tmin < T, the call takes floor(n)+1 explicit sub-steps, all of T/steps
    but the last, which is (n - floor(n))*tmin with n = T/tmin (:80-90).  Each
    sub-step is B <- (I - dt/3600 M) B with the chain's rate matrix M.
    Returns (B after the call before the negative clamp, steps, outflow mass
    per cell over the step).
    """
    B0 = np.asarray(B0, dtype=float)
    lam = np.asarray(lam_h, dtype=float)
    n = B0.size
    M = np.diag(lam) - np.diag(lam[:-1], -1)
    dW = M @ B0
    if dW.min() >= 0:
        tmin = T_s
    else:
        dneg = np.minimum(dW, 0)
        tmin = 3600.0 * np.min(B0 / (np.abs(dneg) + 1e-16))
    mid = B0 - dW * T_s / 3600.0
    if mid.min() >= 0 and tmin >= T_s:
        return mid, 0, lam * B0 * T_s / 3600.0
    n_dec = T_s / tmin
    n_floor = math.floor(n_dec)
    steps = n_floor + 1
    B, out = B0.copy(), np.zeros(n)
    for s in range(1, steps + 1):
        dt = (n_dec - n_floor) * tmin if s == steps else T_s / steps
        out += lam * B * dt / 3600.0
        B = (np.eye(n) - dt / 3600.0 * M) @ B
    return B, steps, out
