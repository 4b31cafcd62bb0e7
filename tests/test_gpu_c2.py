"""The full-physics single-domain workload (sink-filled fractal DEM, 8 LULC
classes, 5 Green-Ampt soil classes, Gaussian-random-field rain on a
triangular hyetograph, critical-depth outlets on the last column, one
pollutant with build-up / wash-off, the groundwater module every 10 steps)
on the GPU:

* a 256 x 320 cut in FP64 against the oracle to 1e-9 (the tolerance and the
  dt-tie rule of the loop parity tests), from the workload's dry start and
  from a wet start, each over hundreds of steps with tens of groundwater calls;
* the FP32 error of every output field against the FP64 oracle on the wet cut;
* the 4096 x 4096 workload itself in FP32 after a spin-up, against the same
  run in FP64 on the device (depths within 1e-3 m, masses and heads within
  their bounds), with non-negative finite state and no water created.
"""
import numpy as np
import pytest

from hydropol2d_amd import synthetic
from oracle import oracle

pytestmark = pytest.mark.gpu


def _close(a, b, name, rtol=1e-9):
    a, b = np.asarray(a, float), np.asarray(b, float)
    np.testing.assert_array_equal(np.isnan(a), np.isnan(b), err_msg=name)
    scale = np.nanmax(np.abs(b)) if np.isfinite(b).any() else 1.0
    np.testing.assert_allclose(a, b, rtol=rtol, atol=rtol * (scale + 1e-300), equal_nan=True,
                               err_msg=name)


def c2_cut(precision="fp64", wet=False):
    spec = synthetic.full_physics_c2(256, 320, 2.0, precision=precision, routing_time_min=180.0)
    if wet:
        # a wet start: 5 to 60 mm of water over the cut, so that routing,
        # wash-off and the dt control are active from the first step
        rng = np.random.default_rng(7)
        spec.d0 = np.asfortranarray(rng.uniform(5.0, 60.0, spec.shape))
    return spec


def _run_pair(spec, steps):
    from hydropol2d_amd.simulation import Simulation
    with Simulation(spec) as s:
        s.run(steps)
        rep = s.report()
        st = s.state(infiltration=True, water_quality=True, groundwater=True)
        dts = s.dt_history(8192)
    return rep, st, dts


@pytest.mark.parametrize("wet,steps", [(False, 400), (True, 300)])
def test_c2_cut_fp64_matches_oracle(wet, steps):
    spec = c2_cut(wet=wet)
    print(f"c2 cut fp64 wet={wet} steps={steps}", flush=True)
    rep, st, dts = _run_pair(spec, steps)
    assert rep["steps"] == steps == len(dts)
    assert rep["gw_updates"] >= 2
    o = oracle.run(spec, max_steps=steps, dt_follow=dts)
    print(f"  {o['n_dt_ties']} dt ties, {rep['gw_updates']} groundwater calls, "
          f"t = {rep['t_min']:.4f} min", flush=True)
    assert o["n_dt_mismatch"] == 0
    orep = o["report"]
    for key in ("steps", "rejected_steps", "catch_index", "n_reports", "wq_substep_steps",
                "gw_updates"):
        assert rep[key] == orep[key], (key, rep[key], orep[key])
    _close(dts, o["dt_hist"][-len(dts):], "dt")
    for key in ("d_t", "outflow", "I_t", "dmax", "gw_h"):
        _close(st[key], o[key], key)
    for key in ("B_t", "P_conc", "Tot_Washed"):
        _close(st[key], o[key], key, 1e-8)
    for key in ("t_min", "time_step_s", "cum_rain_m3", "cum_infiltration_m3", "storage_m3",
                "outflow_volume_m3", "gw_storage_m3", "gw_total_recharge_m3"):
        assert rep[key] == pytest.approx(orep[key], rel=1e-9, abs=1e-9), key
    for key in ("wq_mass_kg", "wq_mass_lost_kg", "wq_tmin_s"):
        assert rep[key] == pytest.approx(orep[key], rel=1e-8, abs=1e-12), key
    scale = max(1.0, orep["storage_m3"] + orep["cum_infiltration_m3"] + orep["gw_storage_m3"])
    assert abs(rep["volume_error_m3"] - orep["volume_error_m3"]) <= 1e-9 * scale


def test_c2_cut_fp32_error_per_output_field():
    # the FP32 run against the FP64 oracle on the FP32 run's own time levels
    # (dt_force): the snap of the dt to 1 ms increments breaks ties of the
    # wash-off's tmin_wq differently in FP32, and a step taken 1 ms longer or
    # shorter is a different trajectory, not an arithmetic error
    steps = 300
    rep, st, dts = _run_pair(c2_cut("fp32", wet=True), steps)
    o = oracle.run(c2_cut(wet=True), max_steps=steps, dt_follow=dts, dt_force=True)
    free = oracle.run(c2_cut(wet=True), max_steps=steps)
    orep = o["report"]
    rows = {}

    def rel_field(a, b):
        return float(np.nanmax(np.abs(a - b)) / max(np.nanmax(np.abs(b)), 1e-300))

    rows["d_t [m]"] = (float(np.nanmax(np.abs(st["d_t"] - o["d_t"]))) / 1000.0, 1e-3)
    rows["dmax [m]"] = (float(np.nanmax(np.abs(st["dmax"] - o["dmax"]))) / 1000.0, 1e-3)
    rows["I_t [mm]"] = (float(np.nanmax(np.abs(st["I_t"] - o["I_t"]))), 0.05)
    rows["gw head [m]"] = (float(np.nanmax(np.abs(st["gw_h"] - o["gw_h"]))), 1e-3)
    rows["x-face outflow / max"] = (rel_field(st["outflow"][:, :, 0], o["outflow"][:, :, 0]), 5e-2)
    rows["y-face outflow / max"] = (rel_field(st["outflow"][:, :, 1], o["outflow"][:, :, 1]), 5e-2)
    rows["B_t / max"] = (rel_field(st["B_t"], o["B_t"]), 1e-2)
    for key, tol in (("cum_rain_m3", 1e-4), ("cum_infiltration_m3", 1e-3), ("storage_m3", 1e-3),
                     ("outflow_volume_m3", 2e-2), ("wq_mass_kg", 1e-3), ("gw_storage_m3", 1e-6)):
        rows[key + " (rel)"] = (abs(rep[key] - orep[key]) / max(abs(orep[key]), 1e-300), tol)
    rows["steps (rel)"] = (abs(rep["steps"] - orep["steps"]) / orep["steps"], 0.0)
    # not a bound of the north star: the same run against the oracle on its own dt
    rows["d_t free-running [m]"] = (float(np.nanmax(np.abs(st["d_t"] - free["d_t"]))) / 1000.0, 2e-2)
    print(f"\nFP32 - FP64 oracle, 256 x 320 full-physics cut, wet start, 300 steps "
          f"({o['n_dt_mismatch']} of the oracle's own dt differ from the FP32 run's)", flush=True)
    for k, (v, tol) in rows.items():
        print(f"  {k:32s} {v:10.3e}   (bound {tol:g})", flush=True)
    for k, (v, tol) in rows.items():
        assert v <= tol, (k, v, tol)


def test_c2_full_size_fp32_against_fp64_on_device():
    # the 4096^2 workload from a wet start (ponded water of 30 mm on average,
    # a smooth random field with a CV of 0.5 and 100 m correlation: a dry start
    # keeps the dt at its minimum for minutes of model time while the first
    # rain infiltrates), FP32 against FP64 on the device after the same steps
    from hydropol2d_amd.simulation import Simulation
    steps = 200
    d0 = np.asfortranarray(synthetic.gaussian_random_field(4096, 4096, 2.0, 30.0, 0.5, 100.0,
                                                           seed=11))
    out = {}
    for prec in ("fp32", "fp64"):
        spec = synthetic.full_physics_c2(4096, 4096, 2.0, precision=prec, routing_time_min=180.0)
        spec.d0 = d0
        print(f"c2 4096^2 {prec}, wet start: {steps} steps", flush=True)
        with Simulation(spec) as s:
            s.run_fixed(steps, graph_steps=50)
            rep = s.report()
            st = s.state(infiltration=True, water_quality=True, groundwater=True)
            dts = s.dt_history(steps)
        out[prec] = (rep, st, dts)
    A = spec.resolution ** 2
    z = spec.elevation
    wq, gw = spec.water_quality, spec.groundwater
    m0 = float(np.sum(wq.C1[wq.lulc_class] * (1 - np.exp(-wq.C2[wq.lulc_class] * wq.ADD_days)) *
                      A / 1e4))
    S_gw0 = float(np.sum(A * gw.Sy * (gw.h0 - (z - gw.soil_depth))))
    rep32, st32, dts32 = out["fp32"]
    rep64, st64, dts64 = out["fp64"]
    assert rep32["steps"] == rep64["steps"] == steps
    assert rep32["gw_updates"] == rep64["gw_updates"] >= steps // 10 - 1
    for k in ("d_t", "I_t", "B_t", "gw_h"):
        assert np.all(np.isfinite(st32[k])), k
    assert np.min(st32["d_t"]) >= 0.0 and np.min(st32["B_t"]) >= 0.0
    # no water is created: surface + soil + aquifer + outflow - initial <= rain
    # (the Green-Ampt drainage k_out and the recovery reset only remove water)
    n = 4096 * 4096
    surface = float(np.nansum(st32["d_t"], dtype=np.float64)) / 1000 * A
    soil = float(np.nansum(st32["I_t"], dtype=np.float64)) / 1000 * A
    aquifer = float(np.nansum(A * gw.Sy * (st32["gw_h"] - (z - gw.soil_depth)), dtype=np.float64))
    total = surface + soil + aquifer + rep32["cum_outflow_m3"]
    initial = float(np.sum(d0)) / 1000 * A + spec.infiltration.I_min_mm * n / 1000 * A + S_gw0
    print(f"water: rain {rep32['cum_rain_m3']:.6e}, change of surface + soil + aquifer + "
          f"outflow {total - initial:.6e} m3", flush=True)
    assert rep32["cum_rain_m3"] > 0
    assert total - initial <= rep32["cum_rain_m3"] + 1e-6 * initial
    assert float(np.nansum(st32["B_t"], dtype=np.float64)) <= m0 * (1 + 1e-5)
    # FP32 against FP64 on the same device path (the dt sequences may part at
    # snap ties; the comparison is after the same number of steps)
    assert rep32["t_min"] == pytest.approx(rep64["t_min"], rel=1e-2)
    err_d = float(np.max(np.abs(st32["d_t"] - st64["d_t"]))) / 1000.0
    err_h = float(np.max(np.abs(st32["gw_h"] - st64["gw_h"])))
    err_b = float(np.max(np.abs(st32["B_t"] - st64["B_t"]))) / float(np.max(st64["B_t"]))
    print(f"4096^2 fp32 vs fp64 after {steps} steps (t = {rep64['t_min']:.3f} min): "
          f"depth {err_d:.2e} m, head {err_h:.2e} m, B_t {err_b:.2e} of max", flush=True)
    assert err_d <= 1e-3 and err_h <= 1e-3 and err_b <= 1e-2
