"""Pins of the oracle's map stacks (save_output_maps.m:24-27, 148-258, 832-848)
and Human_Instability_Module (mode 1, Human_Instability_Module.m:4-26):

* records: on a flat closed basin under constant rain the records are taken
  at the steps whose t_save passes a new multiple of record_time_maps (an
  independent walk over the accepted dt sequence), each holding the uniform
  depth the rain has built (its closed form), zero velocity and hazard;
* risk: at rest the risk is 0 where the person's weight exceeds the buoyancy
  and 1 where the water floats them (h > W / (w1 w2 rho)), NaN outside the
  domain; it rises with the velocity.
"""
import numpy as np
import pytest

from oracle import oracle
import test_oracle_ledger as ledger

HI = {"weight_person": 70.0, "width1_person": 0.35, "width2_person": 0.25, "mu": 0.5, "Cd": 1.1,
      "ro_water": 1000.0, "gravity": 9.81}


def test_map_records_on_a_flat_basin():
    r_mmh, T, rec = 36.0, 10.0, 2.0
    spec = ledger.flat_basin(2, 3, r_mmh, 0.0, T)
    spec.record_time_maps_min = rec
    spec.human_instability = HI
    res = oracle.run(spec, n_map_records=16)
    tcr = res["dt_hist"]
    n = res["report"]["steps"]
    mn = spec.min_time_step_s
    dts = np.concatenate([[mn], tcr[:-1]])
    depth = np.cumsum(r_mmh / 3600.0 * np.concatenate([[mn] * 3, dts[2:n - 1]]))
    # walk the steps: t (end of step k) and t_save = t + tcr_k; a record when the
    # number of record times below t_save grows (record times 0, rec, ..., T)
    t_end = (mn + np.concatenate([[0.0], np.cumsum(tcr)])) / 60.0
    times = np.arange(0.0, T + 1e-9, rec)
    last, steps, ts = 1, [], []
    for k in range(n):
        t_save = t_end[k] + tcr[k] / 60.0
        actual = int(np.sum(times < t_save))
        if actual > last:
            steps.append(k)
            ts.append(t_end[k])
        last = actual
    assert res["report"]["n_map_records"] == len(steps) >= 4
    np.testing.assert_allclose(res["map_record_t"], ts, rtol=1e-14)
    R = res["map_records"]
    for i, k in enumerate(steps):
        np.testing.assert_allclose(R["d"][:, :, i], depth[k], rtol=1e-13)
    assert np.all(R["velocity"] == 0.0) and np.all(R["hazard_dv"] == 0.0)
    assert np.all(R["risk"] == 0.0)              # shallow water at rest
    assert np.all(np.isnan(R["I_t"]))            # no infiltration in this run


def test_risk_at_rest_follows_buoyancy():
    # h* = W / (w1 w2 rho) = 70 / (0.35 * 0.25 * 1000) = 0.8 m: below it the
    # person stands (risk 0 at rest), above it the water floats them (risk 1)
    spec = ledger.flat_basin(3, 4, 0.0, 500.0, 1.0)
    d0 = np.full((3, 4), 500.0)
    d0[:, 2:] = 1200.0
    spec.d0 = np.asfortranarray(d0)
    z = np.zeros((3, 4))
    z[:, 2:] = -0.7            # the same water level: still water
    z[0, 0] = np.nan
    spec.elevation = np.asfortranarray(z)
    spec.d0[0, 0] = 0.0
    spec.human_instability = HI
    res = oracle.run(spec, max_steps=5)
    mr = res["max_risk"]
    assert np.isnan(mr[0, 0])
    np.testing.assert_array_equal(mr[1:, :2], 0.0)
    np.testing.assert_array_equal(mr[0, 1], 0.0)
    np.testing.assert_array_equal(mr[:, 2:], 1.0)


def test_risk_rises_with_velocity():
    # water running down a steep plane: the risk is positive where it moves,
    # and is larger on the faster, steeper of two planes
    out = []
    for slope in (0.02, 0.08):
        spec = ledger.flat_basin(6, 40, 200.0, 50.0, 2.0)
        spec.elevation = np.asfortranarray(np.tile(-slope * 2.0 * np.arange(40), (6, 1)))
        spec.outlet_mask = np.zeros((6, 40), np.uint8, order="F")
        spec.outlet_mask[:, -1] = 1
        spec.human_instability = HI
        res = oracle.run(spec)
        out.append(res["max_risk"])
    assert np.nanmax(out[0]) > 0
    assert np.nanmax(out[1]) > np.nanmax(out[0])
