"""Pins of the FP64 CPU oracle's mainline infiltration step
(Infiltration_Module.m:101-583) on closed forms and the storage ledger."""
import numpy as np
import pytest

from oracle import oracle

SOIL = dict(theta_sat=0.45, theta_r=0.05, n_vg=1.6, alpha_vg=3.0)


def _case(d_mm, I_mm, ks=20.0, zwt=2.0, **kw):
    shape = np.shape(d_mm) if np.ndim(d_mm) else (1, 1)
    d = np.broadcast_to(np.asarray(d_mm, float), shape)
    # water table zwt below the surface: h_t = elevation - soil_depth + (soil_depth - zwt)
    elev, depth = 100.0, 3.0
    h_t = elev - depth + (depth - zwt)
    args = dict(time_step_min=1.0, resolution=5.0, d_t=d, I_t=np.broadcast_to(I_mm, shape),
                elevation=elev, h_t=h_t, soil_depth=depth, ksat_mmh=ks, **SOIL)
    args.update(kw)
    return oracle.infiltration(**args)


def test_supply_limited_infiltrates_all_ponding():
    # capacity and storage far above the supply: f = d/dt, all ponded water enters
    r = _case([[3.0]], 0.0, ks=1000.0)
    np.testing.assert_allclose(r["f"], 3.0 * 60.0, rtol=1e-14)
    assert r["d_t"][0, 0] == 0.0
    np.testing.assert_allclose(r["I_t"], 3.0, rtol=1e-14)


def test_storage_limited_fills_the_bucket():
    # 0.1 m of unsaturated zone (zwt) holds 0.1*(0.45-0.05) = 40 mm; 35 mm held
    r = _case([[500.0]], 35.0, ks=1e5, zwt=0.1)
    np.testing.assert_allclose(r["f"], 5.0 * 60.0, rtol=1e-12)
    np.testing.assert_allclose(r["I_t"], 40.0, rtol=1e-12)
    np.testing.assert_allclose(r["d_t"], 495.0, rtol=1e-12)


@pytest.mark.parametrize("wK", [0.0, 0.5, 1.0])
def test_capacity_with_full_head_cap_and_wetted_top(wK):
    # 2 m of ponding: the driving head is capped at dh_max = 1 m; the wetted top
    # layer reaches saturation (Mualem Kr(1) = 1), so C = Ks (1/Ltop + 1)
    r = _case([[2000.0]], 0.0, ks=5.0, zwt=2.5, Ksurf_weight=wK)
    np.testing.assert_allclose(r["f"], 5.0 * (1.0 / 0.05 + 1.0), rtol=1e-12)


def test_impervious_and_no_unsaturated_zone():
    r = _case([[10.0, 10.0]], 0.0, idx_imp=np.array([[True, False]]), zwt=1.0)
    assert r["f"][0, 0] == 0.0 and r["f"][0, 1] > 0
    r = _case([[10.0]], 0.0, zwt=0.0)   # water table at the surface
    assert r["f"][0, 0] == 0.0 and r["d_t"][0, 0] == 10.0


def test_storage_ledger_closes():
    # nothing clamps: the water that leaves the surface enters the soil;
    # errors(4) (the storage change) is zero to round-off, inf_volume is the
    # infiltrated volume, and the cumulative map advances by f dt / 1000
    rng = np.random.default_rng(0)
    d = rng.uniform(0.0, 20.0, (6, 7))
    I = rng.uniform(0.0, 50.0, (6, 7))
    cum = np.ones((6, 7))
    r = _case(d, I, ks=rng.uniform(1, 30, (6, 7)), zwt=1.5, cum_inf=cum)
    A = 25.0
    assert abs(r["dS_inf"]) < 1e-12 * (np.sum(d + I) * A / 1000)
    np.testing.assert_allclose(r["inf_volume"], np.sum(r["f"] * (1 / 60) * A) / 1000, rtol=1e-12)
    np.testing.assert_allclose(r["cum_inf"], 1.0 + r["f"] * (1 / 60) / 1000, rtol=1e-15)
    np.testing.assert_array_equal(r["I_p"], I)
    assert (r["f"] >= 0).all() and (r["f"] <= d * 60 + 1e-9).all()
