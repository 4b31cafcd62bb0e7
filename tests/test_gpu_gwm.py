"""GPU parity of the groundwater module step (Groundwater_Module.m:37-404, no
layered soil) against the FP64 oracle, call by call.

Tolerance: the scheduler's decisions (whether an update runs, its dt, the
update count) must agree exactly.  The recharge bucket evaluates van
Genuchten / Mualem powers with the device's `pow`, which may differ from the
host library's in the last place, so every field is compared to a relative
1e-12 (absolute 1e-12 of its own scale); the Boussinesq update behind it is
bit-identical to the oracle for identical inputs (test_gpu_gw.py).
"""
import numpy as np
import pytest

from oracle import oracle

pytestmark = pytest.mark.gpu

RTOL = 1e-12


def _setup(ny, nx, seed, ksat_gw=36.0, nan_corner=False, **kw):
    from hydropol2d_amd.groundwater import Groundwater_Module
    rng = np.random.default_rng(seed)
    x = np.arange(nx)[None, :] * 5.0
    elev = np.asfortranarray(100.0 - 0.01 * x + rng.uniform(0, 0.05, (ny, nx)))
    domain = None
    if nan_corner:
        elev[: ny // 3, : nx // 3] = np.nan
        domain = ~np.isnan(elev)
    depth = 2.0
    z_bed = elev - depth
    h = np.asfortranarray(z_bed + 0.5 + rng.uniform(0, 0.1, (ny, nx)))
    I = np.asfortranarray((elev - h) * 0.4 * 1000.0 * rng.uniform(0.3, 0.9, (ny, nx)))
    d = np.asfortranarray(rng.uniform(0, 5.0, (ny, nx)))
    theta_s = rng.uniform(0.4, 0.5, (ny, nx))
    args = (elev, depth, theta_s, 0.05, 1.6, 3.0, 20.0, ksat_gw, 0.2, 5.0)
    gm_o = oracle.GroundwaterModule(*args, domain=domain, **kw)
    gm_g = Groundwater_Module(*args, domain=domain, **kw)
    state = dict(h=h, I=I, d=d, cum=np.zeros((ny, nx), order="F"),
                 maxd=np.full((ny, nx), -np.inf, order="F"))
    return gm_o, gm_g, state


def _close(a, b, what):
    scale = np.nanmax(np.abs(b)) if np.isfinite(b).any() else 0.0
    np.testing.assert_allclose(a, b, rtol=RTOL, atol=RTOL * max(scale, 1e-300), equal_nan=True,
                               err_msg=what)


def _run(gm_o, gm_g, state, calls, dt=1.0, T=1e6):
    so = {k: v.copy(order="F") for k, v in state.items()}
    sg = {k: v.copy(order="F") for k, v in state.items()}
    n_ran = 0
    for k in range(1, calls + 1):
        t = k * dt
        q_o, e_o, ran_o, dtg = gm_o.step(dt, t, T, so["h"], so["I"], so["d"], so["cum"], so["maxd"])
        q_g, e_g, ran_g = gm_g.step(dt, t, T, sg["h"], sg["I"], sg["d"], sg["cum"], sg["maxd"])
        assert bool(ran_o) == ran_g, f"call {k}: update decision"
        n_ran += ran_g
        for key in so:
            _close(sg[key], so[key], f"call {k}: {key}")
        _close(q_g, q_o, f"call {k}: q_exf")
        for key in ("net", "recharge", "capillary"):
            _close(gm_g.pending[key], gm_o.pending[key], f"call {k}: pending {key}")
        st = gm_g.state
        assert st["n_updates"] == gm_o.s.n_updates
        assert st["last_dt_s"] == gm_o.s.last_dt_s == (dtg if ran_o else 0.0)
        assert st["elapsed_s"] == gm_o.s.elapsed_s
        for key in ("total_recharge_m3", "total_capillary_m3", "total_exfiltration_m3"):
            np.testing.assert_allclose(st[key], getattr(gm_o.s, key), rtol=RTOL, atol=1e-300,
                                       err_msg=f"call {k}: {key}")
        err_scale = 0.2 * np.nansum(np.abs(so["h"])) * 25.0
        assert abs(e_g - e_o) <= 1e-12 * err_scale, (k, e_g, e_o)
    return n_ran


def test_async_scheduler_and_boussinesq_update():
    # the stable timestep of this aquifer is about 9 one-minute calls
    gm_o, gm_g, s = _setup(20, 24, seed=1, ksat_gw=3750.0)
    assert _run(gm_o, gm_g, s, 24) >= 2


def test_synchronous_local_update():
    gm_o, gm_g, s = _setup(16, 20, seed=2, flag_baseflow=0, flag_async=0)
    assert _run(gm_o, gm_g, s, 4) == 4


def test_synchronous_boussinesq_with_cells_outside():
    gm_o, gm_g, s = _setup(33, 47, seed=3, ksat_gw=360.0, nan_corner=True, flag_async=0)
    assert _run(gm_o, gm_g, s, 4) == 4


def test_head_change_trigger_and_final_step():
    # a small head-change cap fires the update before the stable timestep;
    # the last call of the run forces one
    gm_o, gm_g, s = _setup(12, 14, seed=4, max_head_change_m=2e-4)
    n = _run(gm_o, gm_g, s, 6, T=6.0)
    assert n >= 2


def test_large_grid_water_balance():
    # 384 x 512 cells, synchronous Boussinesq on a closed domain: surface +
    # unsaturated + saturated storage + pending exchange stays constant
    from hydropol2d_amd.groundwater import Groundwater_Module
    ny, nx = 384, 512
    rng = np.random.default_rng(5)
    elev = np.asfortranarray(100.0 - 0.002 * np.arange(nx)[None, :] * 5 + rng.uniform(0, 0.05, (ny, nx)))
    z_bed = elev - 2.0
    h = np.asfortranarray(z_bed + 0.5 + rng.uniform(0, 0.1, (ny, nx)))
    I = np.asfortranarray((elev - h) * 0.4 * 1000.0 * 0.5)
    d = np.asfortranarray(rng.uniform(0, 5.0, (ny, nx)))
    gm = Groundwater_Module(elev, 2.0, 0.45, 0.05, 1.6, 3.0, 20.0, 360.0, 0.2, 5.0, flag_async=0)

    def total():
        return (np.sum(d) / 1000 + np.sum(I) / 1000 + 0.2 * np.sum(h - z_bed)
                + np.nansum(gm.pending["net"])) * 25.0

    w0 = total()
    for k in range(3):
        _, _, ran = gm.step(1.0, k + 1.0, 1e6, h, I, d)
        assert ran
    np.testing.assert_allclose(total(), w0, rtol=1e-10)


def test_rejects_wrong_layout():
    from hydropol2d_amd.groundwater import Groundwater_Module
    gm = Groundwater_Module(np.full((4, 5), 10.0), 2.0, 0.45, 0.05, 1.6, 3.0, 20.0, 36.0, 0.2, 5.0)
    h = np.full((4, 5), 8.5)  # C order
    with pytest.raises(ValueError):
        gm.step(1.0, 1.0, 10.0, h, np.asfortranarray(h), np.asfortranarray(h))
