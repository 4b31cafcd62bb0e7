"""Pins of the FP64 CPU oracle's Full_Momentum_Model_D4 (ordinary grid,
HLL finite volumes with the implicit Manning predictor) on closed forms and
invariants:

* the HLL flux of a dam-break pair at rest, recovered from the depth change
  (mass) and from the momentum of both cells (momentum flux + reflective wall
  pressure), along x and along y (the y faces' S -> N sign convention);
* the exact implicit Manning friction: a uniform flow on a flat bed keeps its
  depth and its discharge becomes the positive root of a*Q^2 + Q = Q*;
* a flat lake at rest stays at rest; the hydrostatic reconstruction reduces to
  the raw fluxes on a flat bed;
* volume conservation in a closed basin, and an outlet that removes exactly
  what its page reports (cell- and face-based outlets);
* the 10 m/s velocity cap.
"""
import math

import numpy as np
import pytest

from oracle import oracle

G = 9.81


def _fm(z, d, dt_min, R=2.0, **kw):
    ny, nx = np.shape(z)
    return oracle.full_momentum(z, d, kw.pop("n", 0.03), R * R, dt_min, R, **kw)


@pytest.mark.parametrize("axis", ["x", "y"])
def test_hll_dam_break_pair_mass_and_momentum(axis):
    hL, hR, R, dt_min = 0.8, 0.3, 2.0, 1e-5
    dt, A = dt_min * 60, R * R
    if axis == "x":
        z = np.zeros((1, 2))
        d = np.array([[hL, hR]]) * 1000
    else:  # row 0 is north; flow from the deep southern cell (row 1) to the north
        z = np.zeros((2, 1))
        d = np.array([[hR], [hL]]) * 1000
    o = _fm(z, d, dt_min, R, n=0.0)
    cL = math.sqrt(G * hL)
    Fh = cL * (hL - hR) / 2                        # HLL, sL = -cL, sR = +cL
    Fqn = 0.25 * G * (hL ** 2 + hR ** 2)
    deep = (0, 0) if axis == "x" else (1, 0)
    shallow = (0, 1) if axis == "x" else (0, 0)
    np.testing.assert_allclose(d[deep] - o["d_t"][deep], 1000 * dt * R * Fh / A, rtol=1e-9)
    np.testing.assert_allclose(o["d_t"][shallow] - d[shallow], 1000 * dt * R * Fh / A, rtol=1e-9)
    # momentum: flux in/out plus the wall pressure 0.5 g h^2 of the outer faces
    mom_deep = dt * R / A * (0.5 * G * hL ** 2 - Fqn) * min(1.0, o["d_t"][deep] / 1000 / hL)
    mom_shallow = dt * R / A * (Fqn - 0.5 * G * hR ** 2)
    # both cells gain momentum towards the shallow side: +hu is east, +hv is north
    page = 3 if axis == "x" else 4
    assert mom_deep > 0 and mom_shallow > 0
    np.testing.assert_allclose(o["outflow"][deep + (page,)], mom_deep, rtol=1e-9)
    np.testing.assert_allclose(o["outflow"][shallow + (page,)], mom_shallow, rtol=1e-9)
    other = 4 if axis == "x" else 3
    assert np.all(o["outflow"][:, :, other] == 0)


@pytest.mark.parametrize("n,h,q0", [(0.03, 0.5, 0.4), (0.1, 0.2, 0.05), (0.015, 1.2, 2.0)])
def test_exact_implicit_friction_on_uniform_flow(n, h, q0):
    ny, nx, R, dt_min = 7, 9, 5.0, 0.02
    dt = dt_min * 60
    z = np.zeros((ny, nx))
    d = np.full((ny, nx), h * 1000)
    prev = np.zeros((ny, nx, 5))
    prev[:, :, 3] = q0                                 # hu memory page
    o = _fm(z, d, dt_min, R, n=n, outflow_prev=prev)
    a = dt * G * n ** 2 / h ** (7 / 3)
    q = 2 * q0 / (1 + math.sqrt(1 + 4 * a * q0))        # a q^2 + q = q0
    inner = (slice(2, -2), slice(2, -2))
    np.testing.assert_allclose(o["outflow"][inner + (3,)], q, rtol=1e-12)
    np.testing.assert_allclose(o["d_t"][inner], h * 1000, rtol=1e-12)
    assert np.all(o["outflow"][inner + (4,)] == 0)


def test_flat_lake_at_rest():
    o = _fm(np.full((6, 8), 3.0), np.full((6, 8), 450.0), 0.05)
    assert np.all(o["outflow"][:, :, [0, 1, 3, 4]] == 0)
    np.testing.assert_array_equal(o["d_t"], np.full((6, 8), 450.0))


def test_hydrostatic_reconstruction_reduces_to_raw_fluxes_on_a_flat_bed():
    # on a flat bed the reconstructed face depths are the cell depths and the
    # pressure corrections vanish; on a stepped bed they are not
    rng = np.random.default_rng(3)
    d = rng.random((7, 9)) * 500 + 20
    prev = np.zeros((7, 9, 5))
    prev[:, :, 3:] = rng.normal(0, 0.05, (7, 9, 2))
    flat = np.full((7, 9), 2.0)
    a = _fm(flat, d, 0.01, flag_HR=1, outflow_prev=prev)
    b = _fm(flat, d, 0.01, flag_HR=0, outflow_prev=prev)
    for k in ("d_t", "outflow", "Hf"):   # (z + h) - z and h * (hu / h) round differently
        np.testing.assert_allclose(a[k], b[k], rtol=1e-12, atol=1e-12)
    step = flat + (np.arange(9)[None, :] % 2) * 0.1
    a = _fm(step, d, 0.01, flag_HR=1, outflow_prev=prev)
    b = _fm(step, d, 0.01, flag_HR=0, outflow_prev=prev)
    assert np.abs(a["d_t"] - b["d_t"]).max() > 1e-3


@pytest.mark.parametrize("flag_HR", [0, 1])
def test_closed_basin_conserves_volume(flag_HR):
    ny, nx, R = 12, 15, 2.0
    rng = np.random.default_rng(flag_HR + 5)
    z = rng.random((ny, nx)) * 0.3 + np.linspace(0, 1.0, nx)[None, :]
    z[3, 4] = np.nan
    d = rng.random((ny, nx)) * 400
    d[np.isnan(z)] = np.nan
    total = np.nansum(d)
    prev = None
    for _ in range(20):
        o = _fm(z, d, 0.002, R, flag_HR=flag_HR, outflow_prev=prev)
        assert np.nansum(o["d_t"]) == pytest.approx(total, rel=1e-12)
        assert np.isnan(o["d_t"][3, 4]) and np.isnan(o["I_tot_end_cell"][3, 4])
        d, prev = o["d_t"], o["outflow"]
    assert np.abs(prev[:, :, 3]).max() > 1e-3


@pytest.mark.parametrize("outlet_type", [1, 2])
@pytest.mark.parametrize("faces", [False, True])
def test_outlet_removes_exactly_its_page(outlet_type, faces):
    ny, nx, R, dt_min = 8, 10, 2.0, 0.003
    rng = np.random.default_rng(9)
    z = rng.random((ny, nx)) * 0.2 + np.linspace(1.0, 0.0, nx)[None, :]
    d = rng.random((ny, nx)) * 300 + 50
    kw = dict(outlet_type=outlet_type, slope_outlet=0.01)
    if faces:
        right = np.zeros((ny, nx), bool)
        right[:, -1] = True
        down = np.zeros((ny, nx), bool)
        down[-1, -2:] = True
        kw["face_masks"] = (right, None, None, down)
    else:
        kw.update(row_outlet=[3, 4, 4, 5], col_outlet=[nx, nx, nx, nx])   # one repeat
    o = _fm(z, d, dt_min, R, **kw)
    out_mm = o["outflow"][:, :, 2].sum() * dt_min / 60
    assert out_mm > 0
    assert d.sum() - np.nansum(o["d_t"]) == pytest.approx(out_mm, rel=1e-10)
    nz = o["outflow"][:, :, 2] > 0
    assert np.all(o["Hf"][:, :, 2][nz] > 0)
    if faces:  # the corner cell drains through its right and its bottom face
        assert o["outflow"][-1, -1, 2] > 0 and o["outflow"][-1, -2, 2] > 0


def test_velocity_cap():
    ny, nx = 5, 5
    d = np.full((ny, nx), 100.0)
    prev = np.zeros((ny, nx, 5))
    prev[:, :, 3] = 3.0                         # 30 m/s at 0.1 m depth
    o = _fm(np.zeros((ny, nx)), d, 1e-6, 2.0, n=0.0, outflow_prev=prev)
    speed = np.hypot(o["outflow"][:, :, 3], o["outflow"][:, :, 4]) / (o["d_t"] / 1000)
    assert speed.max() <= 10.0 + 1e-9
    assert speed[2, 2] == pytest.approx(10.0, rel=1e-9)
