"""The benchmark configurations that are not bench lines, on the GPU.

* The 256 x 256 tilted plane (FP64, Green-Ampt, critical outlet, 1 h), the
  case the reference runs on a CPU as-is: the whole run against the FP64
  oracle at 1e-9.
* The 2048 x 2048 dam-break stress case (FP32, V valley, breach, dry bed): at
  full size, the size-independent properties (mirror symmetry of the valley,
  the volume ledger, non-negative depths, the dt collapse at the front), and
  the same setup cut to 320 x 128 against the FP64 oracle at 1e-9.
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


def test_tilted_plane_256_full_run_matches_oracle():
    from hydropol2d_amd.simulation import Simulation
    spec = synthetic.tilted_plane_c1(256, 256, routing_time_min=60.0)
    print("tilted plane 256 x 256, FP64, 1 h", flush=True)
    with Simulation(spec) as s:
        assert s.run()
        rep, st, hy, dts = s.report(), s.state(infiltration=True), s.hydrograph(), s.dt_history()
    o = oracle.run(spec)
    orep = o["report"]
    print(f"  {rep['steps']} steps", flush=True)
    assert rep["steps"] == orep["steps"] and rep["rejected_steps"] == orep["rejected_steps"]
    _close(dts, o["dt_hist"][-len(dts):], "dt")
    for key in ("d_t", "outflow", "I_t", "dmax"):
        _close(st[key], o[key], key)
    _close(hy["hydro_q"], o["hydro_q"], "hydrograph")
    for key in ("t_min", "outflow_volume_m3", "cum_rain_m3", "cum_infiltration_m3", "storage_m3",
                "outlet_runoff_volume_mm"):
        assert rep[key] == pytest.approx(orep[key], rel=1e-9, abs=1e-9), key


def test_dam_break_cut_matches_oracle():
    from hydropol2d_amd.simulation import Simulation
    spec = synthetic.dam_break_c3(320, 128, precision="fp64", routing_time_min=30.0)
    print("dam break 320 x 128, FP64, 300 steps", flush=True)
    with Simulation(spec) as s:
        s.run(300)
        rep, st, dts = s.report(), s.state(), s.dt_history()
    o = oracle.run(spec, max_steps=300)
    assert rep["steps"] == o["report"]["steps"] == 300
    _close(dts, o["dt_hist"][-len(dts):], "dt")
    _close(st["d_t"], o["d_t"], "d_t")
    _close(st["outflow"], o["outflow"], "outflow")


def test_dam_break_full_size_properties():
    from hydropol2d_amd.simulation import Simulation
    spec = synthetic.dam_break_c3(2048, 2048, precision="fp32", routing_time_min=30.0)
    A = spec.resolution ** 2
    v0 = float(np.sum(spec.d0, dtype=np.float64)) / 1000.0 * A
    print("dam break 2048 x 2048, FP32, 3000 steps", flush=True)
    with Simulation(spec) as s:
        s.run(3000)
        rep, st, dts = s.report(), s.state(), s.dt_history()
    d = st["d_t"]
    assert rep["steps"] == 3000 and rep["rejected_steps"] == 0
    assert np.isfinite(d).all() and d.min() >= 0
    # the valley, the breach and the dry bed are mirror images across the axis
    mirror = d[:, ::-1]
    wet = d.max()
    assert np.max(np.abs(d - mirror)) <= 1e-5 * wet, np.max(np.abs(d - mirror))
    # the water has left the reservoir through the breach
    assert d[400:, :].sum() > 0.01 * d.sum()
    # volume ledger: storage + outflow through the bottom row = the initial volume
    v1 = rep["storage_m3"] + rep["cum_outflow_m3"]
    assert abs(v1 - v0) <= 1e-5 * v0, (v1, v0)
    # the wet front drives the dt far below its cap
    assert dts.max() < 0.5 and dts.min() >= spec.min_time_step_s
