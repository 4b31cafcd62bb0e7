"""Pins of the oracle's loop with the van Genuchten Infiltration_Module in the
infiltration slot of Hydrological_Model.m:24 (flag_infiltration = 2 of the
device loop).  The closure itself is pinned by closed forms in
test_oracle_inf.py; these tests pin its wiring into the loop: the first step's
unsaturated storage equals the per-call module applied to the post-rain depth
with the run's head, classes and impervious mask (the routing leaves I_t
alone), the closed basin keeps its water, and the groundwater module's head
reaches the closure."""
import numpy as np
import pytest

import loop_cases
from oracle import oracle


def _per_call(spec, d_post, I, h):
    v = spec.infiltration
    cls = v.soil_class
    A = spec.resolution ** 2
    return oracle.infiltration(
        time_step_min=spec.min_time_step_s / 60.0, resolution=spec.resolution, d_t=d_post,
        I_t=I, elevation=spec.elevation, h_t=h, soil_depth=v.soil_depth,
        theta_sat=v.theta_sat[cls], theta_r=v.theta_r[cls], n_vg=v.n_vg[cls],
        alpha_vg=v.alpha_vg[cls], ksat_mmh=v.ksat_mmh[cls],
        idx_imp=None if v.impervious is None else v.impervious.astype(bool), C_a=A,
        l_vg=v.l_vg, Ltop=v.Ltop, dh_max=v.dh_max, Ksurf_weight=v.Ksurf_weight)


@pytest.mark.parametrize("I0", [None, 30.0])
def test_first_step_is_the_per_call_module(I0):
    spec = loop_cases.vg_plane(I0=I0, d0=0.0, rain_mmh=3600.0)
    ny, nx = spec.shape
    o = oracle.run(spec, max_steps=1)
    # the first step runs at min_dt; its rain depth is rate * dt
    rain = 3600.0 * spec.min_time_step_s / 3600.0
    I_start = np.zeros((ny, nx)) if I0 is None else np.full((ny, nx), I0)
    r = _per_call(spec, np.full((ny, nx), rain), I_start, spec.infiltration.h0)
    np.testing.assert_array_equal(o["I_t"], r["I_t"])
    assert o["report"]["cum_infiltration_m3"] == pytest.approx(r["inf_volume"], rel=1e-14)
    # impervious cells take nothing (their storage is only capped at the
    # unsaturated zone's capacity), nor do cells with the water table at the surface
    v = spec.infiltration
    sd = 2.0
    zwt = np.maximum(sd - (v.h0 - (spec.elevation - sd)), 0.0)
    cap = zwt * (v.theta_sat[v.soil_class] - v.theta_r[v.soil_class]) * 1000.0
    imp = v.impervious.astype(bool)
    np.testing.assert_array_equal(o["I_t"][imp], np.minimum(I_start, cap)[imp])
    assert (o["I_t"][ny // 2, :] == 0.0).all()
    low = np.minimum(I_start, cap)
    assert (o["I_t"][~imp] >= low[~imp]).all() and (o["I_t"] > I_start).any()


def test_closed_basin_keeps_its_water():
    # (an initial storage above a cell's capacity would be clipped by the
    # closure's cap, I_t = min(I_t + f dt, uz_max): the run starts from I_t = 0)
    spec = loop_cases.vg_plane(outlets=False, T=3.0, impervious=False)
    ny, nx = spec.shape
    A = spec.resolution ** 2
    o = oracle.run(spec)
    rep = o["report"]
    S0 = np.sum(spec.d0) / 1000 * A
    S1 = (np.sum(o["d_t"]) + np.sum(o["I_t"])) / 1000 * A
    assert S1 - S0 == pytest.approx(rep["cum_rain_m3"], rel=1e-10)
    assert rep["cum_infiltration_m3"] > 0
    # the unsaturated zone never holds more than zwt (theta_sat - theta_r)
    v = spec.infiltration
    zwt = np.maximum(np.nan_to_num(spec.elevation) - v.h0, 0.0)
    cap = zwt * (v.theta_sat - v.theta_r)[v.soil_class] * 1000.0
    assert (o["I_t"] <= cap * (1 + 1e-12)).all()


def test_groundwater_head_reaches_the_closure():
    # the module every step raises the head under a wet surface; a run that
    # reads the moving head infiltrates differently from one on the fixed h0
    spec = loop_cases.vg_plane(gw=True, every=1, T=2.0, impervious=False, outlets=False)
    o_gw = oracle.run(spec)
    spec_fixed = loop_cases.vg_plane(gw=False, T=2.0, impervious=False, outlets=False)
    o_fix = oracle.run(spec_fixed)
    assert o_gw["report"]["gw_updates"] > 0
    assert not np.array_equal(o_gw["I_t"], o_fix["I_t"])
    # closed budget with the aquifer
    A = spec.resolution ** 2
    ny, nx = spec.shape
    v = spec.infiltration
    S_gw0 = float(np.sum(A * 0.2 * (v.h0 - (np.nan_to_num(spec.elevation) - 2.0))))
    S0 = np.sum(spec.d0) / 1000 * A + S_gw0
    rep = o_gw["report"]
    S1 = (np.sum(o_gw["d_t"]) + np.sum(o_gw["I_t"])) / 1000 * A + rep["gw_storage_m3"]
    assert S1 - S0 == pytest.approx(rep["cum_rain_m3"], rel=1e-10)


def test_groundwater_needs_an_infiltration_state():
    spec = loop_cases.plane(wq=False, gw=True)
    spec.infiltration = None
    with pytest.raises(RuntimeError):
        oracle.run(spec, max_steps=1)
