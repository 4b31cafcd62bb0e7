"""Pins of the FP64 CPU oracle's groundwater module step (Groundwater_Module.m
with its scheduler, recharge bucket and Boussinesq / local update) on the
scheduler's timing, telescoping recharge sums and water conservation."""
import numpy as np
import pytest

from oracle import oracle


def _setup(ny=10, nx=12, seed=0, ksat_gw=36.0, **kw):
    rng = np.random.default_rng(seed)
    x = np.arange(nx)[None, :] * 5.0
    elev = np.asfortranarray(100.0 - 0.01 * x + rng.uniform(0, 0.05, (ny, nx)))
    depth = 2.0
    z_bed = elev - depth
    h = np.asfortranarray(z_bed + 0.5 + rng.uniform(0, 0.1, (ny, nx)))
    uz = (elev - h) * 0.4 * 1000.0
    I = np.asfortranarray(uz * rng.uniform(0.3, 0.9, (ny, nx)))
    d = np.asfortranarray(rng.uniform(0, 5.0, (ny, nx)))
    gm = oracle.GroundwaterModule(elev, depth, 0.45, 0.05, 1.6, 3.0, 20.0, ksat_gw, 0.2, 5.0, **kw)
    state = dict(h=h, I=I, d=d, cum=np.zeros((ny, nx), order="F"),
                 maxd=np.full((ny, nx), -np.inf, order="F"))
    return gm, state, elev, z_bed


def _call(gm, s, t, dt=1.0, T=1e6):
    return gm.step(dt, t, T, s["h"], s["I"], s["d"], s["cum"], s["maxd"])


def test_async_accumulates_the_drained_storage():
    # nothing due yet: the pending exchange telescopes to the unsaturated-zone
    # storage drained so far, and the recharge total to its volume
    gm, s, elev, _ = _setup()
    I0 = s["I"].copy()
    for k in range(5):
        q, err, ran, _ = _call(gm, s, k + 1.0)
        assert not ran
    np.testing.assert_allclose(gm.pending["net"], (I0 - s["I"]) / 1000.0, rtol=1e-12, atol=1e-15)
    assert gm.s.elapsed_s == 300.0
    drained = np.sum(np.maximum(I0 - s["I"], 0)) / 1000.0 * 25.0
    np.testing.assert_allclose(gm.s.total_recharge_m3, drained, rtol=1e-9)


def test_update_fires_at_the_stable_timestep():
    # dt_stable = C dx^2 / (4 max(K H / Sy)): K = 3750 mm/h, H up to 0.6 m,
    # Sy = 0.2 gives about 550 s, so the update runs on the 10th one-minute call
    gm, s, elev, z_bed = _setup(ksat_gw=3750.0)
    K = 3750.0 / 1000 / 3600
    H = np.max(s["h"] - z_bed)
    dt_stable = 0.25 * 25.0 / (K * H / 0.2) / 4.0
    expected = int(np.ceil(dt_stable / 60.0))
    for k in range(1, 40):
        _, _, ran, dtg = _call(gm, s, float(k))
        if ran:
            break
    assert k == expected, (k, expected, dt_stable)
    assert dtg == 60.0 * k


def test_synchronous_local_update_closes_the_water_balance():
    # flag_baseflow = 0, flag_groundwater_async = 0: every call moves the
    # exchange into the head (dh = R dt / Sy) and the excess above the ground
    # onto the surface
    gm, s, elev, _ = _setup(flag_baseflow=0, flag_async=0)
    for k in range(3):
        h0, d0, I0 = s["h"].copy(), s["d"].copy(), s["I"].copy()
        _, _, ran, _ = _call(gm, s, k + 1.0)
        assert ran
        storage_in = (I0 - s["I"]) / 1000.0           # left the unsaturated zone
        gained = 0.2 * (s["h"] - h0) + (s["d"] - d0) / 1000.0
        np.testing.assert_allclose(gained, storage_in, rtol=1e-9, atol=1e-12)
        assert np.all(gm.pending["net"] == 0.0)


def test_boussinesq_path_conserves_total_water():
    # closed domain, synchronous Boussinesq updates: surface + unsaturated +
    # saturated storage + pending exchange is constant
    gm, s, elev, z_bed = _setup(flag_async=0, ksat_gw=360.0)

    def total():
        return (np.sum(s["d"]) / 1000 + np.sum(s["I"]) / 1000 + 0.2 * np.sum(s["h"] - z_bed)
                + np.nansum(gm.pending["net"])) * 25.0

    w0 = total()
    for k in range(4):
        _, _, ran, _ = _call(gm, s, k + 1.0)
        assert ran
    np.testing.assert_allclose(total(), w0, rtol=1e-10)
    assert np.all(s["maxd"] >= s["h"] - z_bed - 1e-12)


def test_final_step_forces_the_update():
    gm, s, _, _ = _setup()
    _, _, ran, _ = _call(gm, s, 1.0, T=5.0)
    assert not ran
    _, _, ran, _ = _call(gm, s, 5.0, T=5.0)
    assert ran
