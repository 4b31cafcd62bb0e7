"""The launch-shape tuner times the run's next step with every candidate
shape.  Its trials must not move the run: a run tuned half-way (trials
executed) ends bit-identical to the same run that takes the tuned shape from
the tuner's cache (no trials) at the same point, in FP32 and FP64, for the
full-physics workload (Green-Ampt with soil classes, a rainfall pattern,
wash-off and groundwater), the surface step with lumped rain only, and a
Green-Ampt run without a soil-class map."""
import numpy as np
import pytest

from hydropol2d_amd import synthetic
from hydropol2d_amd.simulation import Simulation

pytestmark = pytest.mark.gpu


def _spec(kind, precision, ny, nx):
    spec = synthetic.full_physics_c2(ny, nx, 2.0, precision=precision, routing_time_min=180.0,
                                     water_quality=kind == "c2", groundwater=kind == "c2")
    rng = np.random.default_rng(11)
    spec.d0 = np.asfortranarray(rng.uniform(2.0, 40.0, spec.shape))
    if kind == "lumped":
        spec.rain_pattern = None
        spec.infiltration = None
    elif kind == "no_soil_map":
        spec.infiltration.soil_class = None
    return spec


def _run(spec, first, then):
    with Simulation(spec) as s:
        s.run_fixed(first)
        tune = s.autotune()
        s.run_fixed(then)
        st = s.state(infiltration=spec.infiltration is not None,
                     water_quality=spec.water_quality is not None,
                     groundwater=spec.groundwater is not None)
        rep = s.report()
    return tune, st, rep


# a grid size per case that no other test tunes, so that the first run's
# tuner call runs its trials and the second one reads its cache
@pytest.mark.parametrize("kind,precision,shape", [
    ("c2", "fp32", (197, 263)),
    ("c2", "fp64", (131, 149)),
    ("lumped", "fp32", (203, 211)),
    ("no_soil_map", "fp32", (139, 277)),
])
def test_tuning_leaves_the_run_unchanged(kind, precision, shape):
    print(f"autotune {kind} {precision} {shape}", flush=True)
    spec = _spec(kind, precision, *shape)
    t1, a, ra = _run(spec, 25, 35)
    t2, b, rb = _run(spec, 25, 35)
    assert t1["ms"] > 0, "the first tuner call must run its trials"
    assert t2["ms"] == -1.0, "the second tuner call must come from the cache"
    assert (t1["tj"], t1["waves_per_block"]) == (t2["tj"], t2["waves_per_block"])
    assert ra["steps"] == rb["steps"]
    for key in a:
        np.testing.assert_array_equal(np.asarray(a[key]), np.asarray(b[key]), err_msg=key)
    for key in ("t_min", "steps", "dt_min"):
        if key in ra:
            assert ra[key] == rb[key], key


def _with_maps_inflow_risk(spec):
    """Rainfall maps from gauges, two inflow hydrographs, the human-instability
    risk map and map records on top of a run (the step kernel's other optional
    arrays)."""
    from hydropol2d_amd import forcing
    from hydropol2d_amd.runspec import Inflow
    from test_oracle_maps import HI
    ny, nx = spec.shape
    dx = spec.resolution
    x_grid = 1000.0 + dx * np.arange(1, nx + 1)
    y_grid = 5000.0 - dx * np.arange(1, ny + 1)
    xg = np.array([1010.0, 1000.0 + dx * nx, 1100.0])
    yg = np.array([4990.0, 5000.0 - dx * ny, 4900.0])
    R = np.array([[80.0, 20.0, 50.0], [10.0, 60.0, 90.0]])
    spec.rain_maps_mmh = forcing.spatial_gauge_rain_maps(R, xg, yg, x_grid, y_grid)
    spec.rain_map_time_min = np.array([1.0, 2.0])
    spec.rain_pattern = None
    cells = np.zeros((ny, nx, 2), np.uint8)
    cells[ny // 3, nx // 2, 0] = 1
    cells[ny - 4:ny - 1, 1, 1] = 1
    spec.inflow = Inflow(time_min=np.array([1.0, 2.0, 1e4]),
                         rate_mmh=np.array([[900.0, 100.0], [0.0, 400.0], [300.0, 300.0]]),
                         cells=cells)
    spec.human_instability = HI
    spec.record_time_maps_min = 0.25
    return spec


# Each lane of the step kernel holds one or two grid rows; the arithmetic per
# cell is the same, so the two shapes must give bit-identical runs.  Even row
# counts, including 250 = 2 * 124 + 2 (a last slab of two rows).
@pytest.mark.parametrize("kind,precision,shape", [
    ("c2", "fp32", (198, 263)),
    ("c2", "fp64", (132, 149)),
    ("lumped", "fp32", (250, 211)),
    ("no_soil_map", "fp64", (250, 67)),
    ("maps", "fp32", (62, 90)),
    ("maps", "fp64", (40, 52)),
])
def test_rows_per_lane_bit_identical(kind, precision, shape):
    print(f"rows per lane {kind} {precision} {shape}", flush=True)
    spec = _spec("c2" if kind == "maps" else kind, precision, *shape)
    if kind == "maps":
        spec = _with_maps_inflow_risk(spec)
    out = []
    for rows in (1, 2):
        with Simulation(spec) as s:
            s.set_launch_shape(32, 2, rows)
            assert s.launch_shape() == {"tj": 32, "waves_per_block": 2, "rows": rows}
            s.run_fixed(60)
            st = s.state(infiltration=spec.infiltration is not None,
                         water_quality=spec.water_quality is not None,
                         groundwater=spec.groundwater is not None)
            extra = {}
            if kind == "maps":
                extra["max_risk"] = s.max_risk()
                extra.update({"rec_" + k: v for k, v in s.map_records().items()})
            out.append((st, extra, s.report()))
    (a, ea, ra), (b, eb, rb) = out
    assert ra["steps"] == rb["steps"] == 60
    for key in a:
        np.testing.assert_array_equal(np.asarray(a[key]), np.asarray(b[key]), err_msg=key)
    for key in ea:
        np.testing.assert_array_equal(ea[key], eb[key], err_msg=key)
    for key in ra:
        np.testing.assert_array_equal(np.asarray(ra[key]), np.asarray(rb[key]), err_msg=key)


def test_launch_shape_refuses_two_rows_on_an_odd_grid():
    spec = _spec("lumped", "fp32", 63, 40)
    with Simulation(spec) as s:
        assert s.launch_shape()["rows"] == 1
        with pytest.raises(Exception):
            s.set_launch_shape(32, 2, 2)
        with pytest.raises(Exception):
            s.set_launch_shape(8, 2, 1)
