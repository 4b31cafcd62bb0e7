"""Replay of the reference's hillslope groundwater validation with an injected solver.

Restates the set-ups, closed forms and metrics of
Validation/Groundwater/Hillslope_Analytical/run_hillslope_analytical_validation.m
(steady Dupuit :42-110, transient linearised decay :112-196, discrete
steady profile :198-205, toe flux :207-210, state records :212-236,
configurations :282-310) so the same harness drives the CPU oracle and the
HIP library.

solver(dt, dx, dy, h, z0, Sy, R, K, courant, h_soil, catchment_mask,
       dirichlet_mask, h_dirichlet) -> dict with at least h_t
"""
from __future__ import annotations

import math

import numpy as np

STEADY = dict(length_m=100.0, n_rows=3, dy_m=20.0, K_m_s=1e-4, Sy=0.25, recharge_m_s=1e-7,
              drain_head_m=0.50, soil_depth_m=5.0, check_dt_s=60.0, courant=0.25)
TRANSIENT = dict(length_m=100.0, n_col=81, n_rows=3, dy_m=20.0, K_m_s=1e-4, Sy=0.25,
                 base_head_m=2.0, amplitude_m=0.01, soil_depth_m=5.0, dt_s=600.0,
                 record_dt_s=86400.0, duration_s=20 * 86400.0, courant=0.25)


def oracle_solver(*args, **kw):
    from oracle import oracle
    return oracle.boussinesq_2d(*args, **kw)


def steady_discrete_profile(cfg, n_col, dx):
    """:198-205: the scheme's exact equilibrium, H^2 summed from the drain."""
    H2 = np.zeros(n_col)
    H2[-1] = cfg["drain_head_m"] ** 2
    for j in range(n_col - 1, 0, -1):           # MATLAB j = n_col-1 .. 1
        H2[j - 1] = H2[j] + 2 * cfg["recharge_m_s"] * j * dx ** 2 / cfg["K_m_s"]
    return np.sqrt(H2)


def toe_flux_per_width(H, K, dx):
    """:207-210."""
    H_face = 0.5 * (H[-2] + H[-1])
    return -K * H_face * (H[-1] - H[-2]) / dx


def steady_dupuit(solver, resolutions=(41, 81, 161)):
    """:42-110; returns per-resolution metrics and the pass flag of :99-103."""
    cfg = STEADY
    out = []
    for n_col in resolutions:
        dx = cfg["length_m"] / (n_col - 1)
        x = np.arange(n_col) * dx
        H_cont = np.sqrt(cfg["drain_head_m"] ** 2 + (cfg["recharge_m_s"] / cfg["K_m_s"]) *
                         (cfg["length_m"] ** 2 - x ** 2))
        H_disc = steady_discrete_profile(cfg, n_col, dx)
        nr = cfg["n_rows"]
        h0 = np.tile(H_disc, (nr, 1))
        R = np.full((nr, n_col), cfg["recharge_m_s"])
        R[:, -1] = 0
        dmask = np.zeros((nr, n_col), dtype=bool)
        dmask[:, -1] = True
        hd = np.zeros((nr, n_col))
        hd[:, -1] = cfg["drain_head_m"]
        r = solver(cfg["check_dt_s"], dx, cfg["dy_m"], h0, np.zeros((nr, n_col)),
                   np.full((nr, n_col), cfg["Sy"]), R, np.full((nr, n_col), cfg["K_m_s"]),
                   cfg["courant"], np.full((nr, n_col), cfg["soil_depth_m"]),
                   np.ones((nr, n_col), dtype=bool), dmask, hd)
        H_model = np.nanmean(r["h_t"], axis=0)
        e_cont = H_model - H_cont
        e_eq = H_model - H_disc
        q_model = toe_flux_per_width(H_model, cfg["K_m_s"], dx)
        q_exp = cfg["recharge_m_s"] * (n_col - 1) * dx
        out.append(dict(n_col=n_col, rmse=float(np.sqrt(np.nanmean(e_cont ** 2))),
                        max_err=float(np.nanmax(np.abs(e_cont))),
                        eq_err=float(np.nanmax(np.abs(e_eq))),
                        q_err_pct=abs(q_model - q_exp) / max(q_exp, 2.2e-16) * 100,
                        H_model=H_model))
    ratio = out[0]["rmse"] / max(out[-1]["rmse"], 2.2e-16)
    f = out[-1]
    passed = (f["rmse"] < 1e-2 and f["max_err"] < 2e-2 and f["eq_err"] < 1e-5 and
              f["q_err_pct"] < 0.1 and ratio > 2)
    return out, ratio, passed


def transient_linearized(solver, duration_s=None):
    """:112-196; returns (series dict, pass flag of :189-190)."""
    cfg = TRANSIENT
    n_col, nr = cfg["n_col"], cfg["n_rows"]
    dx = cfg["length_m"] / (n_col - 1)
    x = np.arange(n_col) * dx
    lam = math.pi / (2 * cfg["length_m"])
    D = cfg["K_m_s"] * cfg["base_head_m"] / cfg["Sy"]
    z = np.zeros((nr, n_col))
    Sy = np.full((nr, n_col), cfg["Sy"])
    K = np.full((nr, n_col), cfg["K_m_s"])
    hs = np.full((nr, n_col), cfg["soil_depth_m"])
    cm = np.ones((nr, n_col), dtype=bool)
    dm = np.zeros((nr, n_col), dtype=bool)
    dm[:, -1] = True
    hd = np.zeros((nr, n_col))
    hd[:, -1] = cfg["base_head_m"]
    h = np.tile(cfg["base_head_m"] + cfg["amplitude_m"] * np.cos(lam * x), (nr, 1))
    h[:, -1] = cfg["base_head_m"]
    dur = cfg["duration_s"] if duration_s is None else duration_s
    n_steps = round(dur / cfg["dt_s"])
    rec_every = round(cfg["record_dt_s"] / cfg["dt_s"])
    series = dict(t=[], q_model=[], q_ref=[], rmse=[], max_err=[])

    def record(t_s, h):
        Hm = np.nanmean(h, axis=0)
        decay = math.exp(-D * lam ** 2 * t_s)
        Ha = cfg["base_head_m"] + cfg["amplitude_m"] * np.cos(lam * x) * decay
        Ha[-1] = cfg["base_head_m"]
        err = Hm - Ha
        series["t"].append(t_s)
        series["q_model"].append(toe_flux_per_width(Hm, cfg["K_m_s"], dx))
        series["q_ref"].append(cfg["K_m_s"] * cfg["base_head_m"] * cfg["amplitude_m"] * lam * decay)
        series["rmse"].append(float(np.sqrt(np.nanmean(err ** 2))))
        series["max_err"].append(float(np.nanmax(np.abs(err))))

    record(0.0, h)
    for it in range(1, n_steps + 1):
        h = solver(cfg["dt_s"], dx, cfg["dy_m"], h, z, Sy, np.zeros((nr, n_col)), K,
                   cfg["courant"], hs, cm, dm, hd)["h_t"]
        if it % rec_every == 0:
            record(it * cfg["dt_s"], h)
    s = {k: np.array(v) for k, v in series.items()}
    qm, qr = s["q_model"], s["q_ref"]
    nse = 1 - np.sum((qm - qr) ** 2) / max(np.sum((qr - qr.mean()) ** 2), 2.2e-16)
    peak_err = abs(qm.max() - qr.max()) / max(qr.max(), 2.2e-16) * 100
    s["nse"], s["peak_err_pct"] = nse, peak_err
    passed = s["rmse"].max() < 0.02 and s["max_err"].max() < 0.05 and nse > 0.99 and peak_err < 5
    return s, passed
