"""Pins of the oracle's report-cadence volume error and of its rollback, both
derived here from the reference's text rather than from the oracle:

* volume_error (mass_balance_check.m:45-73): the storage change since the
  previous report, minus the fluxes of ONE step, whose rainfall is the
  delta_p_agg that update_spatial_BC has already set for the NEXT step
  (HydroPol2D_Main_While.m:721, 793-803).  On a flat closed basin under
  constant rain the depth stays uniform and equals the rain applied so far,
  so every term has a closed form in the accepted dt sequence.
* the catch block (HydroPol2D_Main_While.m:926-993): the state returns to the
  snapshot taken before the step, catch_index rises by one, the new dt is
  alfa_min dx / sqrt(g max(d_tot, d_p)) / catch_index / factor_reduction, and
  max_dt becomes min(1.5 dt, max_dt) for the rest of the run.
"""
import numpy as np
import pytest

from hydropol2d_amd import synthetic
from oracle import oracle


def flat_basin(ny, nx, rain_mmh, d0_mm, T_min, dx=2.0):
    spec = synthetic.tilted_plane_c1(ny, nx, dx=dx, routing_time_min=T_min)
    spec.elevation = np.zeros((ny, nx), order="F")
    spec.outlet_mask = np.zeros((ny, nx), np.uint8, order="F")
    spec.d0 = np.asfortranarray(np.full((ny, nx), float(d0_mm)))
    spec.infiltration = None
    # end-stamped series: the rate holds on (0, 1e4] min
    spec.rain_time_min = np.array([1e4, 2e4])
    spec.rain_intensity_mmh = np.array([rain_mmh, rain_mmh])
    return spec


def test_volume_error_closed_form_on_a_flat_closed_basin():
    r_mmh, T = 36.0, 10.0
    spec = flat_basin(1, 3, r_mmh, 0.0, T)
    res = oracle.run(spec)
    rep = res["report"]
    tcr = res["dt_hist"]            # tcr_k: the dt refreshed after accepted step k
    n = rep["steps"]
    mn = spec.min_time_step_s
    # rain of step j (mm): update_spatial_BC averages over (t_previous, t); the
    # initial call and the first two steps' calls all see an interval of
    # min_dt (t starts at min_dt, and k == 1 sets t_previous = tcr_1 while t
    # has only advanced by it), afterwards step j + 1 gets step j's dt
    dts = np.concatenate([[mn], tcr[:-1]])          # dt used by step j = 1..n
    rain = r_mmh / 3600.0 * np.concatenate([[mn] * 3, dts[2:n - 1]])
    depth = np.cumsum(rain)                        # uniform depth after step j
    np.testing.assert_allclose(res["d_t"].ravel(), depth[-1], rtol=1e-13)
    # report steps: after step j, t has advanced to the end of step j + 1;
    # a report is due once t / T reaches the next 5 % threshold
    t_end = (mn + np.cumsum(tcr)) / 60.0
    thresholds, nxt = [], spec.report_every_percent
    for j in range(n):
        if t_end[j] / T * 100.0 >= nxt or t_end[j] / T * 100.0 >= 100.0:
            thresholds.append(j)
            nxt += spec.report_every_percent
    assert rep["n_reports"] == len(thresholds) >= 2
    A = spec.resolution ** 2
    cells = 3
    S_prev = 0.0
    expect = []
    for j in thresholds:
        S = cells * A * depth[j] / 1000.0
        # the fluxes of the report: P = the next step's rain over the grid
        P_next = cells * A * (r_mmh / 3600.0 * dts[j]) / 1000.0 if j >= 2 else \
            cells * A * (r_mmh / 3600.0 * mn) / 1000.0
        expect.append((S - S_prev) - P_next)
        S_prev = S
    np.testing.assert_allclose(res["volume_error_hist"], expect, rtol=1e-9,
                               atol=1e-12 * cells * A)
    # the mix of a multi-step storage change with a one-step flux is not a
    # conservation error: here it is the rain of all but one step per report
    assert expect[1] > 0


def test_rollback_restores_the_snapshot_and_sets_the_dt():
    # a lake at rest (flat, uniform depth, no rain): nothing moves, so the dt
    # would stay at max_dt; a failure injected at step 5 rolls back
    spec = flat_basin(6, 7, 0.0, 50.0, 10.0)
    k_fail = 5
    before = oracle.run(spec, max_steps=k_fail - 1)
    rolled = oracle.run(spec, inject_fail_step=k_fail, stop_after_rollback=True)
    rep = rolled["report"]
    np.testing.assert_array_equal(rolled["d_t"], before["d_t"])
    assert rep["catch_index"] == 2 and rep["rejected_steps"] == 1
    # d_tot of the failed step = d_p (no rain, no infiltration), so the
    # celerity comes from the snapshot's deepest cell
    c = np.sqrt(9.81 * np.max(before["d_t"]) / 1000.0)
    nt = spec.alfa_min * spec.resolution / c / 2 / spec.factor_reduction
    assert nt < spec.max_time_step_s
    assert rolled["rollback_dt_s"] == pytest.approx(nt, rel=1e-14)
    assert rep["time_step_s"] == pytest.approx(nt, rel=1e-14)
    assert rolled["rollback_max_dt_s"] == pytest.approx(1.5 * nt, rel=1e-14)
    # the rest of the run keeps the lowered max_dt: every later refreshed dt is
    # 1.5 nt snapped down to the increments
    full = oracle.run(spec, inject_fail_step=k_fail, max_steps=40)
    inc = spec.time_step_increments_s
    after = full["dt_hist"][k_fail - 1:]
    np.testing.assert_allclose(after, np.floor(1.5 * nt / inc) * inc, rtol=1e-12)
    assert np.all(full["dt_hist"][: k_fail - 1] == spec.max_time_step_s)


def test_rollback_dt_is_capped_by_max_dt():
    # with max_dt below the celerity dt, the catch block's min(nt, max_dt) binds
    spec = flat_basin(4, 5, 0.0, 50.0, 2.0)
    spec.max_time_step_s = 0.3
    rolled = oracle.run(spec, inject_fail_step=3, stop_after_rollback=True)
    c = np.sqrt(9.81 * 50.0 / 1000.0)
    nt = spec.alfa_min * spec.resolution / c / 2 / spec.factor_reduction
    assert nt > 0.3
    assert rolled["rollback_dt_s"] == 0.3 and rolled["rollback_max_dt_s"] == 0.3


def test_rollback_celerity_uses_the_post_rain_depth():
    # under rain d_tot (after the step's rainfall) exceeds d_p, and the
    # celerity of the catch block uses max(d_tot, d_p) = d_tot
    r_mmh, k_fail = 3600.0, 5
    spec = flat_basin(3, 4, r_mmh, 20.0, 2.0)
    before = oracle.run(spec, max_steps=k_fail - 1)
    rolled = oracle.run(spec, inject_fail_step=k_fail, stop_after_rollback=True)
    np.testing.assert_array_equal(rolled["d_t"], before["d_t"])
    # the failed step's rain is step k_fail - 1's dt of rain (the one-step lag);
    # step k_fail - 1 ran with the dt refreshed after step k_fail - 2
    d_tot = np.max(before["d_t"]) + r_mmh / 3600.0 * before["dt_hist"][k_fail - 3]
    assert d_tot > np.max(before["d_t"]) + 1.0
    nt = spec.alfa_min * spec.resolution / np.sqrt(9.81 * d_tot / 1000.0) / 2 / \
        spec.factor_reduction
    assert rolled["rollback_dt_s"] == pytest.approx(nt, rel=1e-13)
