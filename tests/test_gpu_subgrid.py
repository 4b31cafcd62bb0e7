"""GPU parity of the lookup-table subgrid model against the FP64 oracle,
through the C-ABI (the MATLAB-signature mirrors in hydropol2d_amd.subgrid):
the tables built on the GPU (every field, NaN pattern exact, values to
1e-12 relative), the routing step over several calls with outlets and
inactive cells (outputs to 1e-11 relative; the same number of internal
sub-steps), the rainfall volume change, and at 1024 x 1024 coarse cells the
closed-basin volume conservation of the step."""
import numpy as np
import pytest

from hydropol2d_amd import abi
from hydropol2d_amd import subgrid as sg
from oracle import oracle

pytestmark = pytest.mark.gpu

RES = (None,) * 12


class _Ras:
    def __init__(self, Z, cellsize):
        self.Z, self.cellsize = Z, cellsize


def _close(a, b, name, rtol=1e-12):
    a, b = np.asarray(a, float), np.asarray(b, float)
    np.testing.assert_array_equal(np.isnan(a), np.isnan(b), err_msg=name)
    fin = np.isfinite(b)
    np.testing.assert_array_equal(np.isinf(a), np.isinf(b), err_msg=name)
    if fin.any():
        scale = np.abs(b[fin]).max()
        np.testing.assert_allclose(a[fin], b[fin], rtol=rtol, atol=rtol * (scale + 1e-300),
                                   err_msg=name)


def _dem(fny, fnx, seed, tilt=3.0, holes=0.0):
    rng = np.random.default_rng(seed)
    z = rng.random((fny, fnx)) * 2.0 + np.linspace(0, tilt, fnx)[None, :]
    if holes:
        z[rng.random(z.shape) < holes] = np.nan
    return np.asfortranarray(z)


def _tables_equal(G, O):
    for k, a in O.fields.items():
        _close(G.fields[k], a, k)


@pytest.mark.parametrize("ny,nx,r,nl,fine_n,holes", [
    (3, 4, 2, 5, True, 0.0), (5, 6, 4, 20, False, 0.0), (4, 3, 6, 11, True, 0.05),
    (1, 7, 2, 6, True, 0.0), (6, 1, 4, 6, False, 0.0), (17, 23, 4, 20, True, 0.02)])
def test_tables_match_oracle(ny, nx, r, nl, fine_n, holes):
    print(f"subgrid tables {ny}x{nx} r={r} nl={nl}", flush=True)
    dx = 2.5
    dem = _dem(ny * r, nx * r, ny * nx + r, holes=holes)
    rough = (np.asfortranarray(0.02 + 0.05 * np.random.default_rng(2).random(dem.shape)) if fine_n
             else np.asfortranarray(0.02 + 0.05 * np.random.default_rng(3).random((ny, nx))))
    O = oracle.subgrid_build(dem, dx, rough, dx * r, ny, nx, nl)
    G, inv = sg.Subgrid_Properties_Lookup(_Ras(dem, dx), _Ras(rough, dx * r if not fine_n else dx),
                                          _Ras(np.zeros((ny, nx)), dx * r), dx * r, nr_levels=nl)
    _tables_equal(G, O)
    np.testing.assert_array_equal(np.isnan(inv), np.isnan(O.z_zmin))


def _setup(ny, nx, r, seed, nl=12, holes=0.0):
    dx = 2.0
    dem = _dem(ny * r, nx * r, seed, tilt=1.5, holes=holes)
    rough = np.asfortranarray(0.02 + 0.05 * np.random.default_rng(seed).random(dem.shape))
    O = oracle.subgrid_build(dem, dx, rough, dx * r, ny, nx, nl)
    G, _ = sg.Subgrid_Properties_Lookup(_Ras(dem, dx), _Ras(rough, dx), _Ras(np.zeros((ny, nx)), dx * r),
                                        dx * r, nr_levels=nl)
    return O, G, dx * r


def _gpu_step(G, d, ts, R, outlet_type=2, slope=0.0, rows=(), cols=(), prev=None, idx_nan=None):
    o = sg.Local_Inertial_Model_D4_Subgrid(1, *RES, 0, None, d, None, None, None, R * R, ts, R,
                                           None, outlet_type, slope, list(rows), list(cols), 1.0,
                                           prev, idx_nan, 0, None, None, None, None, None, None,
                                           None, None, None, None, 0, G)
    return {"outflow": o[7], "d_t": o[5], "I_tot_end_cell": o[6], "Hf": o[8], "C_a": o[13],
            "eta_t": o[14], "V_t": o[15], "qout_left": o[0], "qout_right": o[1], "qout_up": o[2],
            "qout_down": o[3], "outlet_flow": o[4]}


@pytest.mark.parametrize("case", ["closed", "critical_outlets", "normal_outlet_nan"])
def test_step_matches_oracle_over_calls(case):
    ny, nx, r = 9, 12, 4
    O, G, R = _setup(ny, nx, r, 5, holes=0.02 if case == "normal_outlet_nan" else 0.0)
    print(f"subgrid step {case}", flush=True)
    rng = np.random.default_rng(7)
    d = np.asfortranarray(rng.random((ny, nx)) * 1500.0)
    rows, cols, otype, slope, mask = (), (), 2, 0.0, None
    if case == "critical_outlets":
        rows, cols = (ny, ny, 4), (nx, nx - 1, nx)
    if case == "normal_outlet_nan":
        rows, cols, otype, slope = (ny, 1), (6, 5), 1, 0.02
        mask = np.zeros((ny, nx), np.uint8)
        mask[3, 3] = 1
    po = pg = None
    do = dg = d
    for _ in range(5):
        o = oracle.li_subgrid(O, do, 0.05, R, otype, slope, rows, cols, outflow_prev=po, idx_nan=mask)
        g = _gpu_step(G, dg, 0.05, R, otype, slope, rows, cols, pg, mask)
        for k in ("outflow", "d_t", "I_tot_end_cell", "Hf", "C_a", "eta_t", "V_t"):
            _close(g[k], o[k], k, 1e-11)
        do, po, dg, pg = o["d_t"], o["outflow"], g["d_t"], g["outflow"]
    assert np.abs(po[:, :, :2]).max() > 10.0
    if rows:
        assert po[:, :, 2].max() > 0
    # the reference's views of the outflow pages
    np.testing.assert_array_equal(g["qout_right"], g["outflow"][:, :, 0])
    np.testing.assert_array_equal(g["qout_left"][:, 1:], -g["outflow"][:, :-1, 0])
    np.testing.assert_array_equal(g["qout_down"][:-1, :], -g["outflow"][1:, :, 1])


def test_set_tables_uploads_host_tables():
    ny, nx, r = 5, 7, 2
    dem = _dem(ny * r, nx * r, 31)
    O = oracle.subgrid_build(dem, 1.0, np.full(dem.shape, 0.03), 2.0, ny, nx, 8)
    d = np.full((ny, nx), 800.0)
    g = _gpu_step(O, d, 0.02, 2.0)       # the oracle-built host tables, uploaded on first use
    o = oracle.li_subgrid(O, d, 0.02, 2.0)
    _close(g["d_t"], o["d_t"], "d_t", 1e-11)
    _close(g["outflow"], o["outflow"], "outflow", 1e-11)


def test_volume_change_matches_oracle():
    O, G, R = _setup(6, 8, 4, 11)
    rng = np.random.default_rng(4)
    d0 = rng.random((6, 8)) * 0.4
    delta = rng.normal(0, 30.0, (6, 8))
    go = sg.hp2d_subgrid_apply_volume_change(d0, delta, G, R * R)
    oo = oracle.subgrid_volume(O, d0, delta, R * R)
    for k, (a, b) in enumerate(zip(go, oo)):
        _close(a, b, f"out{k}", 1e-12)


def test_large_grid_closed_basin_conserves_volume():
    ny = nx = 1024
    r = 2
    print("subgrid 1024^2 build + 3 steps", flush=True)
    dem = _dem(ny * r, nx * r, 99, tilt=4.0)
    rough = np.full(dem.shape, 0.035)
    G, _ = sg.Subgrid_Properties_Lookup(_Ras(dem, 1.0), _Ras(rough, 1.0), _Ras(np.zeros((ny, nx)), 2.0),
                                        2.0, nr_levels=10)
    d = np.asfortranarray(np.random.default_rng(1).random((ny, nx)) * 800.0)
    _, V0, _, _, _ = sg.hp2d_subgrid_apply_volume_change(d / 1000.0, np.zeros((ny, nx)), G, 4.0)
    prev = None
    for _ in range(3):
        g = _gpu_step(G, d, 0.01, 2.0, prev=prev)
        d, prev = g["d_t"], g["outflow"]
    assert g["V_t"].sum() == pytest.approx(V0.sum(), rel=1e-11)
    assert np.abs(prev[:, :, :2]).max() > 10.0
