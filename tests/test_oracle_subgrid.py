"""Pins of the FP64 CPU oracle's lookup-table subgrid model
(Subgrid_Properties_Lookup, the SFINCS table queries,
hp2d_subgrid_apply_volume_change and Local_Inertial_Model_D4_Subgrid).

* Tables against brute force on the fine DEM, written here from the SFINCS
  definitions with numpy slicing of the patches: each cell's level volumes
  are the fine-cell volumes below the level (so a wrong patch, bisection or
  area fails), and each face's tables are the pooled half patches either side
  of the face (a wrong half or direction fails).
* Queries: volume -> level -> volume round trip; the closed form above the
  table top.
* Step: Manning normal flow is a fixed point on flat-bottomed subgrid cells
  (face depth = upstream level - mean bed of the two half patches, n = the
  fine n); lake at rest; closed-basin volume conservation; the outlet removes
  exactly what its page reports; wet-cell depths never go negative.
"""
import math

import numpy as np
import pytest

from oracle import oracle

G = 9.81


def _fine_dem(fny, fnx, seed=0, tilt=3.0):
    rng = np.random.default_rng(seed)
    return np.asfortranarray(rng.random((fny, fnx)) * 2.0 + np.linspace(0, tilt, fnx)[None, :])


def _patch(dem, i, j, r):
    return dem[i * r:(i + 1) * r, j * r:(j + 1) * r]


@pytest.mark.parametrize("r,nl", [(2, 5), (4, 10), (6, 20)])
def test_cell_tables_match_brute_force(r, nl):
    ny, nx, dx = 3, 4, 5.0
    dem = _fine_dem(ny * r, nx * r, seed=r)
    dem[0, 0] = np.nan                     # a hole: the cell keeps its finite cells only
    T = oracle.subgrid_build(dem, dx, np.full(dem.shape, 0.03), dx * r, ny, nx, nl)
    fa = dx * dx
    for i in range(ny):
        for j in range(nx):
            z = _patch(dem, i, j, r).ravel(order="F")
            z = z[np.isfinite(z)]
            assert T.z_zmin[i, j] == z.min() and T.z_zmax[i, j] == z.max()
            volmax = np.maximum(z.max() - z, 0).sum() * fa
            assert T.z_volmax[i, j] == pytest.approx(volmax, rel=1e-12)
            for k in range(nl):
                zl = T.z_level[i, j, k]
                assert T.z_volume[i, j, k] == pytest.approx(volmax * k / (nl - 1), rel=1e-12, abs=1e-12)
                assert np.maximum(zl - z, 0).sum() * fa == pytest.approx(T.z_volume[i, j, k],
                                                                         rel=1e-9, abs=1e-9)
                assert T.z_area[i, j, k] == (zl > z).sum() * fa


def _face_brute(z, n, zmin, zmax, nl):
    levels = np.linspace(zmin, zmax, nl) if zmax > zmin else np.full(nl, zmin)
    hav, nre, pwe = [], [], []
    for zu in levels:
        h = np.maximum(zu - z, 0)
        hav.append(h.mean())
        pwe.append((h > 0).mean())
        den = (np.maximum(zu - np.maximum(z, zmin), 0) ** (5 / 3) / n).mean()
        nre.append(h.mean() ** (5 / 3) / den if h.mean() > 0 and den > 0 else np.inf)
    return np.array(hav), np.array(nre), np.array(pwe)


@pytest.mark.parametrize("r", [2, 4])
def test_face_tables_pool_the_half_patches(r):
    ny, nx, dx, nl = 3, 3, 2.0, 7
    dem = _fine_dem(ny * r, nx * r, seed=10 + r)
    rough = np.asfortranarray(0.02 + 0.05 * np.random.default_rng(3).random(dem.shape))
    T = oracle.subgrid_build(dem, dx, rough, dx * r, ny, nx, nl)
    h = r // 2
    for i in range(ny):
        for j in range(nx - 1):  # u: east half of (i, j), west half of (i, j+1)
            A, B = _patch(dem, i, j, r)[:, -h:], _patch(dem, i, j + 1, r)[:, :h]
            nA, nB = _patch(rough, i, j, r)[:, -h:], _patch(rough, i, j + 1, r)[:, :h]
            z = np.concatenate([A.ravel(order="F"), B.ravel(order="F")])
            n = np.concatenate([nA.ravel(order="F"), nB.ravel(order="F")])
            zmin, zmax = max(A.min(), B.min()), max(A.max(), B.max())
            hav, nre, pwe = _face_brute(z, n, zmin, zmax, nl)
            assert (T.u_zmin[i, j], T.u_zmax[i, j]) == (zmin, zmax)
            np.testing.assert_allclose(T.u_havg[i, j], hav, rtol=1e-12, atol=1e-15)
            np.testing.assert_allclose(T.u_nrep[i, j], nre, rtol=1e-12)
            np.testing.assert_allclose(T.u_pwet[i, j], pwe, rtol=1e-15)
            assert T.u_navg[i, j] == pytest.approx(n.mean(), rel=1e-14)
    for i in range(ny - 1):
        for j in range(nx):  # v: south half of (i, j), north half of (i+1, j)
            A, B = _patch(dem, i, j, r)[-h:, :], _patch(dem, i + 1, j, r)[:h, :]
            z = np.concatenate([A.ravel(order="F"), B.ravel(order="F")])
            n = np.concatenate([_patch(rough, i, j, r)[-h:, :].ravel(order="F"),
                                _patch(rough, i + 1, j, r)[:h, :].ravel(order="F")])
            zmin, zmax = max(A.min(), B.min()), max(A.max(), B.max())
            hav, nre, pwe = _face_brute(z, n, zmin, zmax, nl)
            np.testing.assert_allclose(T.v_havg[i, j], hav, rtol=1e-12, atol=1e-15)
            np.testing.assert_allclose(T.v_nrep[i, j], nre, rtol=1e-12)
    for side, sl in (("north", (slice(0, h), slice(None))), ("south", (slice(r - h, r), slice(None))),
                     ("west", (slice(None), slice(0, h))), ("east", (slice(None), slice(r - h, r)))):
        p = _patch(dem, 1, 2, r)[sl]
        q = _patch(rough, 1, 2, r)[sl]
        hav, nre, pwe = _face_brute(p.ravel(order="F"), q.ravel(order="F"), p.min(), p.max(), nl)
        np.testing.assert_allclose(getattr(T, f"{side}_havg")[1, 2], hav, rtol=1e-12, atol=1e-15)
        np.testing.assert_allclose(getattr(T, f"{side}_nrep")[1, 2], nre, rtol=1e-12)


def test_odd_refinement_ratio_is_refused():
    dem = _fine_dem(9, 9)
    with pytest.raises(ValueError):
        oracle.subgrid_build(dem, 1.0, np.full(dem.shape, 0.03), 3.0, 3, 3, 5)


def test_volume_level_round_trip_and_closed_form_above_top():
    r, ny, nx, dx = 4, 4, 5, 2.5
    dem = _fine_dem(ny * r, nx * r, seed=7)
    T = oracle.subgrid_build(dem, dx, np.full(dem.shape, 0.03), dx * r, ny, nx, 12)
    rng = np.random.default_rng(1)
    zs = T.z_zmin + rng.random((ny, nx)) * (T.z_zmax - T.z_zmin)
    V, zb, A = oracle.subgrid_queries(T, zs)
    np.testing.assert_allclose(zb, zs, rtol=0, atol=1e-11)
    assert np.all(A > 0) and np.all(A <= T.cell_area)
    zs_hi = T.z_zmax + 0.37
    V_hi, zb_hi, A_hi = oracle.subgrid_queries(T, zs_hi)
    np.testing.assert_allclose(V_hi, T.z_volmax + 0.37 * T.cell_area, rtol=1e-13)
    np.testing.assert_allclose(zb_hi, zs_hi, rtol=1e-13)
    assert np.all(A_hi == T.cell_area)
    V_lo, _, A_lo = oracle.subgrid_queries(T, T.z_zmin - 0.1)
    assert np.all(V_lo == 0) and np.all(A_lo == 0)


def test_volume_change_adds_exactly_the_rain_volume():
    r, ny, nx, dx = 2, 3, 3, 4.0
    dem = _fine_dem(ny * r, nx * r, seed=2)
    T = oracle.subgrid_build(dem, dx, np.full(dem.shape, 0.03), dx * r, ny, nx, 9)
    d0 = np.random.default_rng(5).random((ny, nx)) * 0.5
    d1, Vo, Vn, A, res = oracle.subgrid_volume(T, d0, 7.5, T.cell_area)
    np.testing.assert_allclose(Vn - Vo, 7.5 / 1000 * T.cell_area, rtol=1e-12)
    np.testing.assert_allclose(res, 0.0, atol=1e-12)
    d2, *_ = oracle.subgrid_volume(T, d1, -7.5, T.cell_area)
    np.testing.assert_allclose(d2, d0, atol=1e-11)
    # more taken than held: the cell empties and the residual is what was clipped
    d3, Vo3, Vn3, _, res3 = oracle.subgrid_volume(T, d0, -1e6, T.cell_area)
    assert np.all(d3 == 0) and np.all(Vn3 == 0) and np.all(res3 == 0)


def _flat_cells(ny, nx, r, dx, zc, n=0.03):
    """A fine DEM whose every coarse cell is flat at zc[i, j]."""
    dem = np.kron(zc, np.ones((r, r)))
    return np.asfortranarray(dem), np.full(dem.shape, n)


@pytest.mark.parametrize("dt_min", [0.01, 0.1])
def test_manning_normal_flow_is_a_fixed_point(dt_min):
    ny, nx, r, dx, n, S, D = 4, 7, 2, 5.0, 0.035, 0.004, 0.6
    R = dx * r
    zc = 50.0 - S * R * np.arange(nx)[None, :] + np.zeros((ny, 1))
    dem, rough = _flat_cells(ny, nx, r, dx, zc, n)
    T = oracle.subgrid_build(dem, dx, rough, R, ny, nx, 6)
    HG = D + S * R / 2                       # upstream level minus the mean bed of the halves
    q = HG ** (5 / 3) * math.sqrt(S) / n     # Manning unit discharge [m2/s]
    prev = np.zeros((ny, nx, 3), order="F")
    prev[:, :-1, 0] = q / R * 1000 * 3600    # x faces (the last column has none)
    o = oracle.li_subgrid(T, np.full((ny, nx), D * 1000), dt_min, R, 2, 0.0, outflow_prev=prev)
    assert o["n_substeps"] == 1
    np.testing.assert_allclose(o["outflow"][:, :-1, 0], prev[:, :-1, 0], rtol=1e-12)
    np.testing.assert_allclose(o["Hf"][:, :-1, 0], HG, rtol=1e-12)
    assert np.all(o["outflow"][:, :, 1] == 0)
    # the volumes moved downstream by exactly the face fluxes
    V0, _, _ = oracle.subgrid_queries(T, T.z_zmin + D)
    dV = o["V_t"] - V0
    np.testing.assert_allclose(dV[:, 0], -q * R * dt_min * 60, rtol=1e-10)
    np.testing.assert_allclose(dV[:, -1], q * R * dt_min * 60, rtol=1e-10)
    np.testing.assert_allclose(dV[:, 1:-1], 0.0, atol=1e-9)


def test_lake_at_rest():
    r, ny, nx, dx = 4, 5, 6, 2.0
    dem = _fine_dem(ny * r, nx * r, seed=8)
    T = oracle.subgrid_build(dem, dx, np.full(dem.shape, 0.04), dx * r, ny, nx, 10)
    level = np.nanmax(dem) + 0.25
    d = (level - T.z_zmin) * 1000
    o = oracle.li_subgrid(T, d, 0.2, dx * r)
    assert np.all(o["outflow"] == 0)
    np.testing.assert_allclose(o["eta_t"], level, rtol=1e-13)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_closed_basin_conserves_volume_and_outlet_removes_its_page(seed):
    r, ny, nx, dx = 4, 6, 8, 2.5
    R = dx * r
    dem = _fine_dem(ny * r, nx * r, seed=20 + seed, tilt=1.0)
    rough = np.asfortranarray(0.02 + 0.05 * np.random.default_rng(seed).random(dem.shape))
    T = oracle.subgrid_build(dem, dx, rough, R, ny, nx, 15)
    d = np.random.default_rng(seed).random((ny, nx)) * 1500.0
    V0, _, _ = oracle.subgrid_queries(T, T.z_zmin + d / 1000)
    prev = None
    for _ in range(6):
        o = oracle.li_subgrid(T, d, 0.1, R, outflow_prev=prev)
        assert o["V_t"].sum() == pytest.approx(V0.sum(), rel=1e-12)
        assert np.all(o["d_t"] >= 0)
        d, prev = o["d_t"], o["outflow"]
    assert np.abs(prev[:, :, :2]).max() > 100.0
    rows, cols = [ny, ny, 3], [nx, nx - 1, nx]
    o = oracle.li_subgrid(T, d, 0.1, R, 2, 0.0, rows, cols, outflow_prev=prev)
    V1, _, _ = oracle.subgrid_queries(T, T.z_zmin + d / 1000)
    out_m3 = o["outflow"][:, :, 2].sum() / 1000 / 3600 * R * R * 0.1 * 60
    assert out_m3 > 0
    assert V1.sum() - o["V_t"].sum() == pytest.approx(out_m3, rel=1e-10)
    assert np.all(o["Hf"][:, :, 2][o["outflow"][:, :, 2] == 0] == 0)
