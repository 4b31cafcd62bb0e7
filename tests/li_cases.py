"""Replay of the reference's local-inertial / CA validation drivers.

Restates the case set-ups and metrics of
Validation/Hydrodynamics/LocalInertial_CA_Phase1/run_local_inertial_ca_hydrodynamics_validation.m
(plane :44-134, V-tilted :136-217, Ritter :219-296, stage BC :298-384,
routing_step / local_inertial_step / ca_step :386-440, closed forms :523-548)
with the routing step injected as a callable, so the same harness drives the
CPU oracle and the HIP library.

step(z, d_tot_mm, roughness, dt_s, dx, state, outlet) -> (d_t_mm, outlet_flow_mmh, state)
"""
from __future__ import annotations

import math

import numpy as np


def kinematic_plane_q(t_s, rain_m_s, length_m, width_m, n_manning, slope):
    """run_...validation.m:523-538 (closed form)."""
    if t_s <= 0 or rain_m_s <= 0:
        return 0.0
    m = 5.0 / 3.0
    alpha = (1.0 / n_manning) * math.sqrt(slope)
    q_unit_steady = rain_m_s * length_m
    tc_s = (length_m / (alpha * rain_m_s ** (m - 1))) ** (1 / m)
    if t_s < tc_s:
        q_unit = alpha * (rain_m_s * t_s) ** m
    else:
        q_unit = q_unit_steady
    return width_m * q_unit


def ritter_depth(x, t_s, h0, dam_x, g=9.81):
    """run_...validation.m:540-548 (closed form)."""
    c0 = math.sqrt(g * h0)
    xi = np.asarray(x, dtype=float) - dam_x
    h = np.zeros_like(xi)
    h[xi <= -c0 * t_s] = h0
    fan = (xi > -c0 * t_s) & (xi < 2 * c0 * t_s)
    h[fan] = (2 * c0 - xi[fan] / t_s) ** 2 / (9 * g)
    return h


def outlet_cells(ny, nx, side, mask_in):
    """run_...validation.m:450-474; returns (mask, rows1, cols1) in find() order."""
    mask = np.zeros((ny, nx), dtype=bool)
    m = np.asarray(mask_in, dtype=bool).ravel()
    if side == "right":
        mask[:, nx - 1] = m
    elif side == "left":
        mask[:, 0] = m
    elif side == "down":
        mask[ny - 1, :] = m
    elif side == "up":
        mask[0, :] = m
    lin = np.flatnonzero(mask.ravel(order="F"))
    rows = lin % ny + 1
    cols = lin // ny + 1
    return mask, rows, cols


def nse(sim, obs):
    sim, obs = np.asarray(sim), np.asarray(obs)
    return 1.0 - np.sum((sim - obs) ** 2) / np.sum((obs - obs.mean()) ** 2)


def rmse(e):
    e = np.asarray(e)
    e = e[np.isfinite(e)]
    return float(np.sqrt(np.mean(e ** 2))) if e.size else float("nan")


def plane_case(step, n_steps_limit=None):
    """run_plane_case, :44-134."""
    dx, nx, ny, slope, n_man, rain_mm_h = 20.0, 50, 5, 0.005, 0.035, 40.0
    duration_s, dt_s, record_dt_s = 2 * 3600.0, 1.0, 60.0
    x = (np.arange(1, nx + 1) - 0.5) * dx
    z = np.tile(x.max() * slope - x * slope, (ny, 1))
    h = np.zeros((ny, nx))
    n = n_man * np.ones((ny, nx))
    outlet = outlet_cells(ny, nx, "right", np.ones(ny, dtype=bool))
    rain_m_s = rain_mm_h / 1000 / 3600
    n_steps = int(round(duration_s / dt_s))
    if n_steps_limit:
        n_steps = min(n_steps, n_steps_limit)
    record_every = max(1, int(round(record_dt_s / dt_s)))
    area, width, length = dx ** 2, ny * dx, nx * dx
    state = None
    cum_out = 0.0
    q_out = 0.0
    t_min, q_model, q_ref, mass_err = [0.0], [0.0], [0.0], [0.0]
    for it in range(1, n_steps + 1):
        h = h + rain_m_s * dt_s
        d_t, oflow, state = step(z, h * 1000, n, dt_s, dx, state, outlet, slope)
        h = np.maximum(d_t / 1000, 0)
        q_out = float(np.sum(oflow)) / 1000 / 3600 * area
        cum_out += q_out * dt_s
        if it % record_every == 0:
            t_now = it * dt_s
            t_min.append(t_now / 60)
            q_model.append(q_out)
            q_ref.append(kinematic_plane_q(t_now, rain_m_s, length, width, n_man, slope))
            stored = h.sum() * area
            rain_vol = rain_m_s * t_now * nx * ny * area
            mass_err.append(100 * (stored + cum_out - rain_vol) / max(rain_vol, 2.2e-16))
    t_min, q_model, q_ref = map(np.asarray, (t_min, q_model, q_ref))
    valid = t_min > 0
    err = q_model[valid] - q_ref[valid]
    return {
        "nse": nse(q_model[valid], q_ref[valid]),
        "rel_l2": float(np.linalg.norm(err) / max(np.linalg.norm(q_ref[valid]), 2.2e-16)),
        "mass_abs_pct": float(np.max(np.abs(mass_err))),
        "q_model": q_model,
        "q_ref": q_ref,
        "h": h,
    }


def vtilted_case(step, n_steps_limit=None):
    """run_vtilted_case, :136-217."""
    dx, nx, ny = 20.0, 41, 31
    side_slope, downslope, n_man, rain_mm_h = 0.02, 0.004, 0.04, 30.0
    duration_s, dt_s, record_dt_s = 90 * 60.0, 1.0, 60.0
    cc, rr = np.meshgrid(np.arange(1, nx + 1), np.arange(1, ny + 1))
    center = math.ceil(nx / 2)
    z = side_slope * np.abs(cc - center) * dx + downslope * (ny - rr) * dx
    h = np.zeros((ny, nx))
    n = n_man * np.ones((ny, nx))
    omask = np.zeros(nx, dtype=bool)
    omask[center - 2:center + 1] = True
    outlet = outlet_cells(ny, nx, "down", omask)
    rain_m_s = rain_mm_h / 1000 / 3600
    n_steps = int(round(duration_s / dt_s))
    if n_steps_limit:
        n_steps = min(n_steps, n_steps_limit)
    record_every = max(1, int(round(record_dt_s / dt_s)))
    area = dx ** 2
    state = None
    cum_out = 0.0
    sym, mass_err, qs = [], [], []
    for it in range(1, n_steps + 1):
        h = h + rain_m_s * dt_s
        d_t, oflow, state = step(z, h * 1000, n, dt_s, dx, state, outlet, downslope)
        h = np.maximum(d_t / 1000, 0)
        q_out = float(np.sum(oflow)) / 1000 / 3600 * area
        cum_out += q_out * dt_s
        if it % record_every == 0:
            t_now = it * dt_s
            stored = h.sum() * area
            rain_vol = rain_m_s * t_now * nx * ny * area
            mass_err.append(100 * (stored + cum_out - rain_vol) / max(rain_vol, 2.2e-16))
            left = h[:, : center - 1]
            right = np.fliplr(h[:, center:])
            sym.append(rmse((left - right).ravel()))
            qs.append(q_out)
    return {
        "symmetry_max": float(np.max(sym)) if sym else 0.0,
        "mass_abs_pct": float(np.max(np.abs(mass_err))) if mass_err else 0.0,
        "channel_final": float(np.mean(h[:, center - 1])),
        "q": np.asarray(qs),
        "h": h,
    }


def ritter_case(step, n_manning=0.0):
    """n_manning: 0 for local inertial, 0.003 for CA (run_...validation.m:226-230)."""
    return _ritter_case(step, n_manning)


def _ritter_case(step, n_manning):
    """run_ritter_case, :219-296 (registered diagnostic_failed for LI: report only)."""
    dx, nx, ny, dam_x, h0, g = 2.0, 251, 5, 180.0, 1.0, 9.81
    duration_s, dt_s = 20.0, 0.05
    compare_steps = [int(round(t / dt_s)) for t in (5.0, 10.0, 15.0, 20.0)]
    x = (np.arange(1, nx + 1) - 0.5) * dx
    z = np.zeros((ny, nx))
    h = np.tile((x < dam_x).astype(float) * h0, (ny, 1))
    n = n_manning * np.ones((ny, nx))
    outlet = outlet_cells(ny, nx, "none", np.zeros((ny, nx), dtype=bool))
    V0 = h.sum() * dx ** 2
    state = None
    rm, rel, mass = [], [], []
    n_steps = int(round(duration_s / dt_s))
    for it in range(1, n_steps + 1):
        d_t, _, state = step(z, h * 1000, n, dt_s, dx, state, outlet, 0.001)
        h = np.maximum(d_t / 1000, 0)
        if it in compare_steps:
            t_s = it * dt_s
            prof = h.mean(axis=0)
            ref = ritter_depth(x, t_s, h0, dam_x, g)
            err = prof - ref
            active = (x >= dam_x - math.sqrt(g * h0) * t_s - 5 * dx) & (
                x <= dam_x + 2 * math.sqrt(g * h0) * t_s + 5 * dx)
            rm.append(rmse(err[active]))
            rel.append(np.linalg.norm(err[active]) / max(np.linalg.norm(ref[active]), 2.2e-16))
            mass.append(100 * (h.sum() * dx ** 2 - V0) / V0)
    return {"rmse": max(rm), "rel_l2": max(rel), "mass_abs_pct": max(abs(m) for m in mass), "h": h}


def stage_case(step, n_steps_limit=None):
    """run_stage_hydrograph_case, :298-384."""
    dx, nx, ny, n_man, slope = 10.0, 80, 5, 0.03, 0.0005
    base, amp, period_s, duration_s, dt_s, record_dt_s = 0.25, 0.10, 1800.0, 3600.0, 1.0, 60.0
    x = (np.arange(1, nx + 1) - 0.5) * dx
    z = np.tile(x.max() * slope - x * slope, (ny, 1))
    h = base * np.ones((ny, nx))
    n = n_man * np.ones((ny, nx))
    outlet = outlet_cells(ny, nx, "right", np.ones(ny, dtype=bool))
    area = dx ** 2
    n_steps = int(round(duration_s / dt_s))
    if n_steps_limit:
        n_steps = min(n_steps, n_steps_limit)
    record_every = max(1, int(round(record_dt_s / dt_s)))
    V0 = h.sum() * area
    cum_b, cum_out = 0.0, 0.0
    state = None
    stage_err, mass_err, down = [0.0], [0.0], [float(np.mean(h[:, -1]))]
    for it in range(1, n_steps + 1):
        t_s = it * dt_s
        stage_now = base + amp * math.sin(2 * math.pi * t_s / period_s)
        old = h[:, 0].copy()
        h[:, 0] = stage_now
        imm = float(np.max(np.abs(h[:, 0] - stage_now)))
        cum_b += float(np.sum(h[:, 0] - old)) * area
        d_t, oflow, state = step(z, h * 1000, n, dt_s, dx, state, outlet, slope)
        h = np.maximum(d_t / 1000, 0)
        q_out = float(np.sum(oflow)) / 1000 / 3600 * area
        cum_out += q_out * dt_s
        if it % record_every == 0:
            stage_err.append(imm)
            stored = h.sum() * area
            mass_err.append(100 * (stored + cum_out - (V0 + cum_b)) / max(V0 + cum_b, 2.2e-16))
            down.append(float(np.mean(h[:, -1])))
    return {
        "max_stage_error": float(np.max(stage_err)),
        "mass_abs_pct": float(np.max(np.abs(mass_err))),
        "response_amp": float(max(down) - min(down)),
        "h": h,
    }


def oracle_li_step():
    """routing_step / local_inertial_step (:386-415) on the CPU oracle."""
    from oracle import oracle

    def step(z, d_tot, n, dt_s, dx, state, outlet, outlet_slope):
        ny, nx = z.shape
        if state is None:
            state = np.zeros((ny, nx, 3), order="F")
        _, rows, cols = outlet
        out = oracle.li_d4(z, d_tot, n, n ** 2, dx ** 2, dt_s / 60, dx, 1, outlet_slope,
                           rows, cols, state, 0)
        return out["d_t"], out["outlet_flow"], out["outflow"]

    return step


def ca_gather(qleft, qright, qup, qdown):
    """Outflows become inflows (run_...validation.m:431-436, Main_While.m:631-645)."""
    ny, nx = qleft.shape
    qin = np.zeros((ny, nx))
    qin[:, 1:nx] += qright[:, 0:nx - 1]
    qin[:, 0:nx - 1] += qleft[:, 1:nx]
    qin[1:ny, :] += qdown[0:ny - 1, :]
    qin[0:ny - 1, :] += qup[1:ny, :]
    return qin


def oracle_ca_step():
    """ca_step (:417-440) on the CPU oracle."""
    from oracle import oracle

    def step(z, d_tot, n, dt_s, dx, state, outlet, outlet_slope):
        ny, nx = z.shape
        if state is None:
            state = np.zeros((ny, nx), order="F")
        _, rows, cols = outlet
        out = oracle.ca_routing(z, d_tot, n, dx ** 2, dt_s / 60, np.zeros((ny, nx)), dx, state, 1,
                                outlet_slope, rows, cols, 0)
        qin = ca_gather(out["qout_left"], out["qout_right"], out["qout_up"], out["qout_down"])
        d_t = out["d_t"] + qin * (dt_s / 60) / 60
        return d_t, out["outlet_flow"], out["I_tot_end_cell"]

    return step
