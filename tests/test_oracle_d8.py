"""Pins of the FP64 CPU oracle's D8 routines (Bates_Inertial_8D,
CA_Routing_8D, build_up_wash_off_8D and the caller's eight-direction inflow
gather) on closed forms and invariants.

* Bates_Inertial_8D: Manning normal flow on a plane tilted in x and y is a
  fixed point of the update on each downhill page, with its own slope (a wrong
  diagonal length, friction exponent, page order or unit conversion breaks
  it); the critical cap h*sqrt(g*h); a lake at rest; quarter-turn
  equivariance of every page (a wrong neighbour offset breaks it).
* CA_Routing_8D: a lone wet cell on a flat bed shares its outflow equally
  among its eight neighbours and keeps the Vol_min share (closed form); a lake
  at rest; volume conservation of routing + gather in a closed basin; no cell
  sends more than it holds; quarter-turn equivariance.
* build_up_wash_off_8D: pollutant mass conservation (the eight inflows gather
  exactly what the neighbours send) with and without internal sub-steps; the
  outlet mass leaves as Out_Conc says; Tot_Washed = max(B_begin - B_t, 0);
  tmin_wq = 10 s when no cell gains mass; quarter-turn equivariance.
"""
import math

import numpy as np
import pytest

import d8_cases as dc
from oracle import oracle

G = 9.81


def _plane(ny, nx, R, sx, sy, h):
    """Bed falling by sx per metre along +j (east) and sy along +i (south);
    depth h everywhere."""
    j = np.arange(nx)[None, :] * R
    i = np.arange(ny)[:, None] * R
    z = np.asfortranarray(100.0 - sx * j - sy * i + 0 * j * i)
    d = np.full((ny, nx), h * 1000.0, order="F")
    return z, d


# ------------------------------------------------------- Bates_Inertial_8D ----

@pytest.mark.parametrize("dt_min", [0.005, 0.05, 0.5])
def test_bates8d_normal_flow_is_fixed_point_on_every_downhill_page(dt_min):
    ny, nx, R, n, h = 9, 11, 4.0, 0.04, 0.35
    sx, sy = 0.02, 0.008
    z, d = _plane(ny, nx, R, sx, sy, h)
    slope = {"right": sx, "down": sy, "se": (sx + sy) / math.sqrt(2), "ne": (sx - sy) / math.sqrt(2)}
    qprev = np.zeros((ny, nx, 9), order="F")
    for name, s in slope.items():
        q_m2s = h ** (5 / 3) * math.sqrt(s) / n           # Manning unit discharge
        qprev[:, :, dc.PAGES.index(name)] = q_m2s / R * 1000 * 3600
    o = oracle.bates_8d(z, d, d, n, R * R, dt_min, R, 1, 0.0, [], [], qprev)
    inner = (slice(1, -1), slice(1, -1))
    for name in slope:
        p = dc.PAGES.index(name)
        np.testing.assert_allclose(o["outflow"][inner + (p,)], qprev[inner + (p,)], rtol=1e-12,
                                   err_msg=name)
        np.testing.assert_allclose(o["Hf"][inner + (p,)], h, rtol=1e-12)
    for name in ("left", "up", "sw", "nw", "outlet"):
        assert np.all(o["outflow"][inner + (dc.PAGES.index(name),)] == 0.0), name
    # mass leaving each cell = the pages' volume over the step
    vol = qprev[inner].sum(axis=2) / 1000 / 3600 * R * R * dt_min * 60
    np.testing.assert_allclose(o["I_tot_end_cell"][inner], vol, rtol=1e-12)
    np.testing.assert_allclose(o["d_t"][inner], d[inner] - vol / (R * R) * 1000, rtol=1e-12)


def test_bates8d_outlet_normal_slope_is_fixed_point():
    ny, nx, R, n, h, S0 = 5, 6, 2.0, 0.03, 0.2, 0.015
    z = np.full((ny, nx), 10.0, order="F")
    d = np.full((ny, nx), h * 1000, order="F")
    qn = h ** (5 / 3) * math.sqrt(S0) / n / R * 3.6e6
    qprev = np.zeros((ny, nx, 9), order="F")
    qprev[2, 3, 4] = qn
    o = oracle.bates_8d(z, d, d, n, R * R, 0.1, R, 1, S0, [3], [4], qprev)
    assert o["outflow"][2, 3, 4] == pytest.approx(qn, rel=1e-12)
    out = o["outflow"].copy()
    out[2, 3, 4] = 0
    assert np.all(out == 0)


def test_bates8d_critical_cap():
    ny, nx, R, h = 5, 5, 1.0, 0.05
    z, d = _plane(ny, nx, R, 0.3, 0.0, h)
    o = oracle.bates_8d(z, d, d, 1e-4, R * R, 0.5, R, 1, 0.0, [], [], None, flag_critical=1)
    cap = h * math.sqrt(G * h) / R * 3.6e6
    np.testing.assert_allclose(o["outflow"][1:-1, 1:-1, 1], cap, rtol=1e-12)
    free = oracle.bates_8d(z, d, d, 1e-4, R * R, 0.5, R, 1, 0.0, [], [], None)
    assert np.all(free["outflow"][1:-1, 1:-1, 1] > 2 * cap)


def test_bates8d_lake_at_rest():
    rng = np.random.default_rng(3)
    z = np.asfortranarray(rng.integers(0, 40, (7, 8)) / 8.0)  # binary fractions: y = 6 exactly
    d = np.asfortranarray((6.0 - z) * 1000.0)
    o = oracle.bates_8d(z, d, d, 0.03, 4.0, 0.2, 2.0, 2, 0.0, [], [], None)
    assert np.all(o["outflow"] == 0) and np.all(o["I_tot_end_cell"] == 0)
    np.testing.assert_array_equal(o["d_t"], d)


def test_bates8d_dry_and_nan_cells():
    z = dc.rough_dem(8, 9, 5, nan_frac=0.1)
    d = dc.depths_mm(z, 5, wet_frac=0.5)
    o = oracle.bates_8d(z, d, d, 0.03, 4.0, 0.05, 2.0, 1, 0.01, [8], [9], None)
    nan = np.isnan(z)
    assert np.all(np.isnan(o["outflow"][nan])) and np.all(np.isnan(o["d_t"][nan]))
    dry = (d == 0) & ~nan
    assert np.all(o["outflow"][dry] == 0)
    assert np.all(o["outflow"][~nan] >= 0)


@pytest.mark.parametrize("critical", [0, 1])
def test_bates8d_quarter_turn_equivariance(critical):
    ny, nx, R = 7, 10, 3.0
    z = dc.rough_dem(ny, nx, 11)
    d = dc.depths_mm(z, 11)
    rows, cols = [7, 4], [10, 1]
    a = oracle.bates_8d(z, d, d, 0.05, R * R, 0.02, R, 2, 0.0, rows, cols, None, critical)
    rr, cc = dc.rot90_outlets(rows, cols, ny, nx)
    b = oracle.bates_8d(np.rot90(z), np.rot90(d), np.rot90(d), 0.05, R * R, 0.02, R, 2, 0.0, rr, cc,
                        None, critical)
    assert np.count_nonzero(a["outflow"]) > 40
    np.testing.assert_allclose(b["outflow"], dc.rot90_pages(a["outflow"]), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(b["d_t"], np.rot90(a["d_t"]), rtol=1e-12)


# ----------------------------------------------------------- CA_Routing_8D ----

@pytest.mark.parametrize("depth_mm", [5.0, 80.0, 900.0])
def test_ca8d_lone_wet_cell_shares_equally(depth_mm):
    ny, nx, R, n, dt_min = 5, 5, 2.0, 0.03, 0.1
    A = R * R
    z = np.zeros((ny, nx), order="F")
    d = np.zeros((ny, nx), order="F")
    d[2, 2] = depth_mm
    o = oracle.ca_8d(z, d, n, A, dt_min, None, R, np.zeros((ny, nx)), 1, 0.0, [], [], 0, 1.0)
    h = depth_mm / 1000
    v_m = (1 / n) * h ** (2 / 3) * math.sqrt(h / R)   # steepest: the cardinal faces
    # weights 1/9 each (eight faces of volume A*h, plus the Vol_min share A*h)
    I = min(h * A, v_m * h * 9 * R * dt_min * 60, 0.9 * A * h + A * h)
    page = I / 9 / (dt_min * 60) / A * 1000 * 3600
    got = o["outflow"][2, 2, :]
    np.testing.assert_allclose(np.delete(got, 4), page, rtol=1e-12)
    assert got[4] == 0
    np.testing.assert_allclose(o["d_t"][2, 2], depth_mm - 8 / 9 * I / A * 1000, rtol=1e-12)
    assert np.all(np.delete(o["outflow"].reshape(-1, 9, order="F"), 12, axis=0) == 0)


def test_ca8d_lake_at_rest():
    rng = np.random.default_rng(4)
    z = np.asfortranarray(rng.integers(0, 40, (6, 9)) / 8.0)
    d = np.asfortranarray((6.0 - z) * 1000.0)
    o = oracle.ca_8d(z, d, 0.03, 4.0, 0.2, None, 2.0, np.ones((6, 9)), 1, 0.0, [], [], 0, 1.0)
    assert np.all(o["outflow"] == 0)
    np.testing.assert_array_equal(o["d_t"], d)


@pytest.mark.parametrize("critical", [0, 1])
def test_ca8d_closed_basin_conserves_volume(critical):
    ny, nx, R, dt_min = 12, 15, 2.0, 0.05
    z = dc.rough_dem(ny, nx, 21, nan_frac=0.05)
    d = dc.depths_mm(z, 21)
    Iprev = np.zeros((ny, nx))
    total0 = np.nansum(d)
    for _ in range(5):
        o = oracle.ca_8d(z, d, 0.04, R * R, dt_min, None, R, Iprev, 1, 0.0, [], [], critical, 1.0)
        assert np.nanmin(o["d_t"]) >= -1e-9           # a cell never sends more than it holds
        assert np.count_nonzero(o["outflow"]) > 50
        qin, d_new = oracle.gather(o["outflow"], dt_min, o["d_t"])
        assert np.nansum(d_new) == pytest.approx(total0, rel=1e-12)
        d, Iprev = d_new, o["I_tot_end_cell"]


def test_ca8d_outlet_removes_exactly_its_page():
    ny, nx, R, dt_min = 8, 10, 2.0, 0.05
    z = dc.rough_dem(ny, nx, 8)
    d = dc.depths_mm(z, 8, wet_frac=1.0)
    o = oracle.ca_8d(z, d, 0.04, R * R, dt_min, None, R, np.zeros((ny, nx)), 1, 0.02,
                     [8, 8], [9, 10], 0, 1.0)
    qin, d_new = oracle.gather(o["outflow"], dt_min, o["d_t"])
    lost = o["outflow"][:, :, 4].sum() * dt_min / 60   # no page crosses the grid edge
    assert o["outflow"][:, :, 4].sum() > 0
    assert d.sum() - d_new.sum() == pytest.approx(lost, rel=1e-10)


@pytest.mark.parametrize("critical", [0, 1])
def test_ca8d_quarter_turn_equivariance(critical):
    ny, nx, R = 9, 6, 2.5
    z = dc.rough_dem(ny, nx, 13, nan_frac=0.05)
    d = dc.depths_mm(z, 13)
    Ip = np.random.default_rng(1).random((ny, nx))
    rows, cols = [9], [6]
    a = oracle.ca_8d(z, d, 0.05, R * R, 0.03, None, R, Ip, 2, 0.0, rows, cols, critical, 2.0)
    rr, cc = dc.rot90_outlets(rows, cols, ny, nx)
    b = oracle.ca_8d(np.rot90(z), np.rot90(d), 0.05, R * R, 0.03, None, R, np.rot90(Ip), 2, 0.0,
                     rr, cc, critical, 2.0)
    assert np.count_nonzero(a["outflow"]) > 30
    np.testing.assert_allclose(b["outflow"], dc.rot90_pages(a["outflow"]), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(b["d_t"], np.rot90(a["d_t"]), rtol=1e-12)


def test_gather_d4_and_d8_route_each_page_to_its_neighbour():
    rng = np.random.default_rng(2)
    ny, nx = 5, 6
    for k in (5, 9):
        q = np.zeros((ny, nx, k), order="F")
        for p in range(k):
            if p == 4:
                continue
            q[:, :, :] = 0
            q[2, 3, p] = 1.0 + p
            qin, _ = oracle.gather(q)
            di, dj = {0: (0, -1), 1: (0, 1), 2: (-1, 0), 3: (1, 0), 5: (-1, 1), 6: (1, 1),
                      7: (1, -1), 8: (-1, -1)}[p]
            want = np.zeros((ny, nx))
            want[2 + di, 3 + dj] = 1.0 + p
            np.testing.assert_array_equal(qin, want, err_msg=f"page {p}")
    q = rng.random((ny, nx, 9))
    qin, _ = oracle.gather(q)
    assert np.isfinite(qin).all()


# ----------------------------------------------------- build_up_wash_off_8D ----

@pytest.mark.parametrize("C3,split", [(2e3, 0), (1e6, 1)])
def test_washoff8d_closed_basin_conserves_mass(C3, split):
    C3a, C4a, q, B, _ = dc.wq_case(10, 12, 3, outlet=False, C3=C3)
    o = oracle.washoff_8d(C3a, C4a, q, B, 0.5, 4.0, min_Bt=0.0)
    assert o["split"] == split
    assert o["mass_lost"] == 0.0
    assert o["B_t"].sum() == pytest.approx(B.sum(), rel=1e-12)
    assert np.abs(o["B_t"] - B).max() > 1e-6


def test_washoff8d_outlet_mass_and_tot_washed():
    C3a, C4a, q, B, _ = dc.wq_case(9, 11, 4, outlet=True, C3=2e3)
    tw0 = np.full(B.shape, 0.25)
    ts = 0.5
    o = oracle.washoff_8d(C3a, C4a, q, B, ts, 4.0, Tot_Washed=tw0, min_Bt=0.0)
    assert o["split"] == 0
    out_kg = o["Out_Conc"] * (q[:, :, 4].sum() / 1000 * 4.0) / 1000 * ts / 60
    assert B.sum() - o["B_t"].sum() == pytest.approx(out_kg, rel=1e-10)
    np.testing.assert_allclose(o["Tot_Washed"], np.maximum(B - o["B_t"], 0) + tw0, rtol=1e-14)


def test_washoff8d_no_gain_gives_ten_second_tmin():
    C3a, C4a, q, B, _ = dc.wq_case(6, 7, 5, C3=0.01)
    q[:, :, [0, 1, 2, 3, 5, 6, 7, 8]] = 0.0       # only the outlet page: every cell loses
    q[:, :, 4] = 3.0
    o = oracle.washoff_8d(C3a, C4a, q, B, 0.2, 4.0, min_Bt=0.0)
    assert o["tmin_wq"] == 10.0


def test_washoff8d_quarter_turn_equivariance():
    C3a, C4a, q, B, mask = dc.wq_case(7, 10, 6, C3=2e5, nan_frac=0.05)
    a = oracle.washoff_8d(C3a, C4a, q, B, 0.3, 4.0, nan_mask=mask, min_Bt=0.01)
    b = oracle.washoff_8d(np.rot90(C3a), np.rot90(C4a), dc.rot90_pages(q), np.rot90(B), 0.3, 4.0,
                          nan_mask=np.rot90(mask), min_Bt=0.01)
    np.testing.assert_allclose(b["B_t"], np.rot90(a["B_t"]), rtol=1e-12, equal_nan=True)
    np.testing.assert_allclose(b["P_conc"], np.rot90(a["P_conc"]), rtol=1e-12, equal_nan=True)
    assert b["tmin_wq"] == pytest.approx(a["tmin_wq"], rel=1e-12)
