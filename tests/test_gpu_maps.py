"""GPU parity of the device loop's map stacks (save_output_maps.m:24-27,
148-258) and Human_Instability_Module (mode 1) against the FP64 oracle: the
record times, every recorded field (depth, velocity, hazard, risk, I_t, the
groundwater depth and the pollutant maps) and the max_risk map, FP64 to 1e-9;
FP32 records within 1e-3 m of depth."""
import numpy as np
import pytest

import loop_cases
from hydropol2d_amd import synthetic
from oracle import oracle
from test_oracle_maps import HI

pytestmark = pytest.mark.gpu


def _close(a, b, name, rtol=1e-9):
    a, b = np.asarray(a, float), np.asarray(b, float)
    np.testing.assert_array_equal(np.isnan(a), np.isnan(b), err_msg=name)
    scale = np.nanmax(np.abs(b)) if np.isfinite(b).any() else 1.0
    np.testing.assert_allclose(a, b, rtol=rtol, atol=rtol * (scale + 1e-300), equal_nan=True,
                               err_msg=name)


def _plane(precision="fp64"):
    spec = synthetic.tilted_plane_c1(40, 52, routing_time_min=6.0, precision=precision)
    spec.rain_intensity_mmh = np.array([400.0, 0.0])  # enough water to move people
    spec.record_time_maps_min = 0.5
    spec.human_instability = HI
    return spec


def _wq_gw():
    spec = loop_cases.plane(C3=3e4, every=5, T=4.0)
    spec.record_time_maps_min = 0.5
    return spec


@pytest.mark.parametrize("case", ["plane_hi", "wq_gw"])
def test_map_records_match_oracle(case):
    from hydropol2d_amd.simulation import Simulation
    spec = _plane() if case == "plane_hi" else _wq_gw()
    print(f"map stacks {case}", flush=True)
    with Simulation(spec) as s:
        s.run()
        rec = s.map_records()
        rep, dts = s.report(), s.dt_history()
        mr = s.max_risk() if case == "plane_hi" else None
    o = oracle.run(spec, dt_follow=dts, n_map_records=64)
    assert o["n_dt_mismatch"] == 0
    assert rep["n_map_records"] == o["report"]["n_map_records"] == len(rec["t_min"]) >= 6
    np.testing.assert_allclose(rec["t_min"], o["map_record_t"], rtol=1e-12)
    for name, arr in o["map_records"].items():
        if np.all(np.isnan(arr)):
            assert name not in rec, name
            continue
        _close(rec[name], arr, name, 1e-8 if name.startswith("Pol") else 1e-9)
    if mr is not None:
        _close(mr, o["max_risk"], "max_risk")
        assert np.nanmax(mr) > 0.01
        assert np.nanmax(rec["velocity"]) > 0.1 and np.nanmax(rec["hazard_dv"]) > 0


def test_fp32_map_records_within_1mm():
    from hydropol2d_amd.simulation import Simulation
    with Simulation(_plane("fp32")) as s:
        s.run()
        rec, dts = s.map_records(), s.dt_history()
        mr = s.max_risk()
    o = oracle.run(_plane(), dt_follow=dts, dt_force=True, n_map_records=64)
    assert len(rec["t_min"]) == len(o["map_record_t"])
    err = float(np.nanmax(np.abs(rec["d"] - o["map_records"]["d"]))) / 1000.0
    print(f"fp32 map records: depth {err:.2e} m, max_risk "
          f"{np.nanmax(np.abs(mr - o['max_risk'])):.2e}", flush=True)
    assert err <= 1e-3
    assert np.nanmax(np.abs(mr - o["max_risk"])) <= 1e-3


def test_map_stacks_and_time_series_written_as_mat(tmp_path):
    # the device records written as save_map_hydro_N.mat and
    # current_time_series_outputs.mat, read back with SciPy's MAT-file reader
    from scipy.io import loadmat
    from hydropol2d_amd import io as hio
    from hydropol2d_amd.simulation import Simulation
    spec = _plane()
    with Simulation(spec) as s:
        s.run()
        rec = s.map_records()
        paths = hio.save_map_stacks(str(tmp_path), rec)
        ts = hio.write_time_series_outputs(str(tmp_path), s)
        hy = s.hydrograph()
    n = len(rec["t_min"])
    assert len(paths) == (n + 11) // 12
    H = loadmat(paths[0])["Maps"]["Hydro"][0, 0]
    np.testing.assert_array_equal(H["d"][0, 0], rec["d"][:, :, :12].astype(np.float32))
    np.testing.assert_array_equal(H["risk"][0, 0], rec["risk"][:, :, :12].astype(np.float32))
    T = loadmat(ts)["TempTS"][0, 0]          # the struct's (1, 1) element
    q = T["outlet"][0, 0]["hydrograph_m3_s"].ravel()
    np.testing.assert_array_equal(q, hy["hydro_q"][:hy["n_hydro"]])
