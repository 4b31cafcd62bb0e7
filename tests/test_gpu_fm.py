"""GPU parity of Full_Momentum_Model_D4 (ordinary grid) against the FP64
oracle through the C-ABI (the MATLAB-signature mirror): raw and
hydrostatically reconstructed HLL fluxes, cell- and face-based outlets,
inactive cells, momentum memory over repeated calls, all outputs to 1e-11
relative; at 2048 x 2048 the closed-basin volume conservation over calls."""
import numpy as np
import pytest

from hydropol2d_amd import local_inertial as li
from oracle import oracle

pytestmark = pytest.mark.gpu

RES = (None,) * 12


def _close(a, b, name, rtol=1e-11):
    a, b = np.asarray(a, float), np.asarray(b, float)
    np.testing.assert_array_equal(np.isnan(a), np.isnan(b), err_msg=name)
    scale = np.nanmax(np.abs(b)) if np.isfinite(b).any() else 1.0
    np.testing.assert_allclose(a, b, rtol=rtol, atol=rtol * (scale + 1e-300), equal_nan=True,
                               err_msg=name)


def _gpu(z, d, n, A, ts, R, outlet_type=2, slope=0.0, rows=(), cols=(), prev=None, idx_nan=None,
         d_tol=1.0, outlet_index=None, props=None):
    o = li.Full_Momentum_Model_D4(1, *RES, 0, z, d, None, n, None, A, ts, R, outlet_index,
                                  outlet_type, slope, list(rows), list(cols), d_tol, prev, idx_nan,
                                  0, 0, None, None, None, None, None, None, None, None, None, None,
                                  0, 0, None, props)
    return {"outflow": o[7], "d_t": o[5], "I_tot_end_cell": o[6], "Hf": o[8], "C_a": o[13],
            "qout_left": o[0], "qout_right": o[1], "qout_up": o[2], "qout_down": o[3],
            "outlet_flow": o[4]}


def _case(ny, nx, seed, nan_frac=0.0):
    rng = np.random.default_rng(seed)
    z = rng.random((ny, nx)) * 0.4 + np.linspace(1.0, 0.0, nx)[None, :] + np.linspace(0, 0.3, ny)[:, None]
    if nan_frac:
        z[rng.random((ny, nx)) < nan_frac] = np.nan
    d = rng.random((ny, nx)) * 400.0
    d[rng.random((ny, nx)) < 0.15] = 0.0
    d[np.isnan(z)] = np.nan
    n = np.asfortranarray(0.02 + 0.06 * rng.random((ny, nx)))
    return np.asfortranarray(z), np.asfortranarray(d), n


@pytest.mark.parametrize("shape", [(1, 19), (17, 1), (2, 3), (64, 80), (131, 77)])
@pytest.mark.parametrize("hr,outlet", [(0, "cell2"), (1, "cell1"), (0, "faces"), (1, "faces")])
def test_matches_oracle_over_calls(shape, hr, outlet):
    ny, nx = shape
    print(f"full momentum {shape} HR={hr} outlet={outlet}", flush=True)
    z, d, n = _case(ny, nx, ny * 7 + nx, nan_frac=0.04 if ny * nx > 50 else 0.0)
    R, ts = 2.0, 0.002
    rows, cols, otype, props, okw = (), (), 2, {"flag_HR_full_momentum": hr}, {}
    nan = np.zeros((ny, nx), np.uint8)
    if ny * nx > 50:
        nan[1, 1] = 1
    oi = np.zeros((ny, nx), np.uint8)
    if outlet.startswith("cell"):
        otype = int(outlet[-1])
        rows, cols = [ny, ny, 1], [nx, nx, nx]            # a repeat and a second cell
        oi[ny - 1, nx - 1] = oi[0, nx - 1] = 1
    else:
        right = np.zeros((ny, nx), np.uint8)
        right[:, -1] = 1
        down = np.zeros((ny, nx), np.uint8)
        down[-1, :] = 1
        props.update(face_right=right, face_left=np.zeros((ny, nx), np.uint8),
                     face_up=np.zeros((ny, nx), np.uint8), face_down=down)
        okw["face_masks"] = (right, None, None, down)
    prev_o = prev_g = None
    do = dg = d
    for k in range(6):
        o = oracle.full_momentum(z, do, n, 4.0, ts, R, otype, 0.01, rows, cols, prev_o, nan, 1.0,
                                 None, oi, hr, **okw)
        g = _gpu(z, dg, n, 4.0, ts, R, otype, 0.01, rows, cols, prev_g, nan, 1.0, oi, props)
        for key in ("outflow", "d_t", "I_tot_end_cell", "Hf", "C_a"):
            _close(g[key], o[key], f"{key} call {k}")
        do, prev_o, dg, prev_g = o["d_t"], o["outflow"], g["d_t"], g["outflow"]
    assert np.nanmax(np.abs(prev_o[:, :, 3:5])) > 1e-4
    np.testing.assert_array_equal(g["qout_right"], g["outflow"][:, :, 0])
    np.testing.assert_array_equal(g["qout_up"], g["outflow"][:, :, 1])
    np.testing.assert_array_equal(g["outlet_flow"], g["outflow"][:, :, 2])


def test_three_page_memory_and_grid_cell_area():
    # the local-inertial outflow (3 pages) as the momentum memory, per-cell areas
    ny, nx = 20, 24
    z, d, n = _case(ny, nx, 5)
    prev = np.asfortranarray(np.random.default_rng(2).normal(0, 500.0, (ny, nx, 3)))
    A = np.asfortranarray(np.full((ny, nx), 4.0) * (1 + 0.1 * np.random.default_rng(3).random((ny, nx))))
    o = oracle.full_momentum(z, d, n, A, 0.003, 2.0, 2, 0.0, [ny], [nx], prev)
    g = _gpu(z, d, n, A, 0.003, 2.0, 2, 0.0, [ny], [nx], prev)
    for key in ("outflow", "d_t", "I_tot_end_cell", "Hf", "C_a"):
        _close(g[key], o[key], key)


def test_large_closed_basin_conserves_volume():
    ny = nx = 2048
    print("full momentum 2048^2", flush=True)
    z, d, n = _case(ny, nx, 11)
    v0 = d.sum()
    prev = None
    for _ in range(4):
        g = _gpu(z, d, n, 4.0, 0.002, 2.0, prev=prev)
        d, prev = g["d_t"], g["outflow"]
    assert d.sum() == pytest.approx(v0, rel=1e-11)
    assert np.abs(prev[:, :, 3]).max() > 1e-4
