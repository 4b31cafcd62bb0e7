"""Pins of the oracle's legacy explicit Green-Ampt closure
(HydroPol2D_Main_While_Ensemble.m:243-298) and of its derived parameters
(HydroPol2D_preprocessing.m:2249-2253):

* against the reference's ponded Green-Ampt fixture (P1-INFIL-GA-001: the
  implicit solution F - psi*dtheta*ln(1 + F/(psi*dtheta)) = Ks*t), with the
  explicit scheme's error falling at first order in dt;
* against hand-computed single-cell steps for every branch: the ponded and the
  supply-limited closure, the Lu cap, the recovery clock (reset, decrement, the
  I_min reset once it runs out), the k_out drainage and the I_min floor;
* and that the step loop applies exactly this closure, cell by cell.
"""
import os

import numpy as np
import pytest

from hydropol2d_amd import synthetic
from hydropol2d_amd.runspec import GreenAmpt
from oracle import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "P1-INFIL-GA-001_reference.csv")


def _fixture():
    data = np.genfromtxt(GOLDEN, delimiter=",", names=True)
    return data["time_s"], data["cumulative_infiltration_m"], data["Ks_m_s"][0], \
        data["psi_delta_theta_m"][0]


def test_fixture_is_the_ponded_green_ampt_root():
    # the fixture's own closed form: F - S ln(1 + F/S) = Ks t, S = psi*dtheta
    t, F, Ks, S = _fixture()
    np.testing.assert_allclose(F - S * np.log1p(F / S), Ks * t, rtol=1e-9, atol=1e-15)


def _ponded_run(dt_s, t, F, Ks, S, d_pond=1.0, dtheta=0.5):
    """Explicit closure held ponded at d_pond mm (the surface is refilled to it
    every step, so the supply never limits), psi chosen so that
    (d_pond + psi) * dtheta = S: the capacity is then Ks (1 + S/I), the
    ponded Green-Ampt rate.  Starts from the fixture's first non-zero point."""
    ks_mmh = Ks * 1000.0 * 3600.0
    p = oracle.ga_params(ks_mmh, S * 1000.0 / dtheta - d_pond, dtheta, 0.0, lu_m=1e3)
    assert p["k_out"] == 0.0 and p["Lu_mm"] == 1e6
    dt_h = dt_s / 3600.0
    I = np.array([F[1] * 1000.0])
    T = np.array([p["Tr"]])
    out = [I[0]]
    tt = t[1]
    for t_next in t[2:]:
        n = int(round((t_next - tt) / dt_s))
        for _ in range(n):
            d, I_new, T, f = oracle.ga_step(p, False, dt_h, np.array([d_pond]), I, T)
            assert f[0] < d_pond / dt_h          # ponded: the capacity governs
            assert d[0] == pytest.approx(d_pond - f[0] * dt_h, rel=1e-15)
            I = I_new
        tt = t_next
        out.append(I[0])
    return np.array(out) / 1000.0


def test_ponded_closure_converges_to_the_fixture_at_first_order():
    t, F, Ks, S = _fixture()
    err = {}
    for dt in (120.0, 60.0, 30.0):
        Fm = _ponded_run(dt, t, F, Ks, S)
        err[dt] = np.max(np.abs(Fm - F[1:]))
    # forward Euler on dF/dt = Ks (1 + S/F): the error halves with the step
    assert 1.8 < err[120.0] / err[60.0] < 2.2
    assert 1.8 < err[60.0] / err[30.0] < 2.2
    # and is small against the 24 h infiltration of the fixture
    assert err[30.0] < 2e-4 * F[-1]


# hand-computed single-cell steps; the class has ksat = 25.4 mm/h (so
# sqrt(ksat/25.4) = 1), psi = 100 mm, dtheta = 0.2, theta_s - theta_r = 0.4 and
# Lu = 75 mm: kr = 1/75 1/h, Tr = 4.5 h, k_out = 0.4 * (1/75) * 0.075 * 1000 = 0.4 mm/h
def _cls():
    return oracle.ga_params(25.4, 100.0, 0.2, 0.4, lu_m=0.075)


def test_derived_parameters_by_hand():
    p = _cls()
    assert p["Tr"] == pytest.approx(4.5, rel=1e-15)
    assert p["Lu_mm"] == pytest.approx(75.0, rel=1e-15)
    assert p["k_out"] == pytest.approx(0.4, rel=1e-14)
    # Lu derived from ksat: 4 sqrt(ksat/25.4) inches; ksat = 6.35 mm/h -> 2 in
    q = oracle.ga_params(6.35, 100.0, 0.2, 0.3)
    assert q["Lu_mm"] == pytest.approx(50.8, rel=1e-14)
    assert q["Tr"] == pytest.approx(9.0, rel=1e-15)
    assert q["k_out"] == pytest.approx(0.3 * (1 / 150) * 0.0508 * 1000, rel=1e-14)


@pytest.mark.parametrize("case", [
    # (name, first, d [mm], I_p [mm], T_in [h], dt_h, d_after, I_t, T_out)
    # ponded: i_a = 100 > C = 25.4 (1 + 110*0.2/20) = 53.34 -> f = C, clock reset
    ("ponded", False, 10.0, 20.0, 1.0, 0.1, 10.0 - 5.334, 20.0 + 5.334, 4.5),
    # supply-limited: i_a = 50 <= C = 25.4 * 2.05 = 52.07 -> f = 50 and the
    # k_out drainage 0.4 * 0.1 leaves the store
    ("supply", False, 5.0, 20.0, 1.0, 0.1, 0.0, 20.0 + 5.0 - 0.04, 0.9),
    # Lu cap: I_p = 80 > 75 -> C = ksat = 25.4 < i_a = 100
    ("lu_cap", False, 10.0, 80.0, 1.0, 0.1, 10.0 - 2.54, 80.0 + 2.54, 4.5),
    # dry surface, clock running: T 0.3 -> 0.2, drainage only
    ("drain", False, 0.0, 20.0, 0.3, 0.1, 0.0, 20.0 - 0.04, 0.2),
    # clock runs out: T 0.05 -> -0.05 < 0 -> I_t = I_min
    ("recovered", False, 0.0, 20.0, 0.05, 0.1, 0.0, 1e-3, -0.05),
    # I_min floor of the drainage: 0.01 - 0.04 < I_min
    ("floor", False, 0.0, 0.01, 1.0, 0.1, 0.0, 1e-3, 0.9),
    # first step: no Lu cap and no drainage, clock set to Tr:
    # C = 25.4 (1 + 110*0.2/80) = 32.385 < i_a = 100
    ("first", True, 10.0, 80.0, 0.0, 0.1, 10.0 - 3.2385, 80.0 + 3.2385, 4.5),
])
def test_single_cell_steps_by_hand(case):
    _, first, d, Ip, Tin, dt_h, d_after, I_t, T_out = case
    got_d, got_I, got_T, _ = oracle.ga_step(_cls(), first, dt_h, np.array([d]), np.array([Ip]),
                                            np.array([Tin]), I_min_mm=1e-3)
    assert got_d[0] == pytest.approx(d_after, rel=1e-12, abs=1e-12)
    assert got_I[0] == pytest.approx(I_t, rel=1e-12)
    assert got_T[0] == pytest.approx(T_out, rel=1e-12, abs=1e-12)


def test_the_loop_applies_the_closure():
    # cells isolated by NaN cells on all four sides (no face carries flow), so
    # each step is rain, the Green-Ampt closure and the 1e-6 mm clean-up only;
    # the loop's depths and infiltrated depths equal the closure iterated on the
    # loop's own dt, for cells that start at different depths and soil moisture
    ny, nx = 5, 7
    spec = synthetic.tilted_plane_c1(ny, nx, routing_time_min=5.0)
    ii, jj = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    valid = (ii % 2 == 1) & (jj % 2 == 1)
    spec.elevation = np.asfortranarray(np.where(valid, 0.0, np.nan))
    spec.outlet_mask = np.zeros((ny, nx), np.uint8, order="F")
    spec.d0 = np.asfortranarray(np.where(valid, 0.5 + ii * jj, 0.0))
    # an end-stamped series: 20 mm/h on (0, 1e4] and (1e4, 2e4] min
    spec.rain_time_min = np.array([1e4, 2e4])
    spec.rain_intensity_mmh = np.array([20.0, 20.0])
    I0 = np.asfortranarray(np.where(valid, 0.5 + 5.0 * jj + ii, 1.0))
    spec.infiltration = GreenAmpt(ksat_mmh=np.array([25.4]), psi_mm=np.array([100.0]),
                                  dtheta=np.array([0.2]), theta_sat_minus_r=np.array([0.4]),
                                  lu_m=np.array([0.075]), I0_mm=I0, I_min_mm=1e-3)
    r = oracle.run(spec)
    n_steps = r["report"]["steps"]
    assert n_steps > 20 and r["report"]["rejected_steps"] == 0
    dts = np.concatenate([[spec.min_time_step_s], r["dt_hist"][:-1]])
    # rain of each step (update_spatial_BC lags one step; the first three steps
    # get the initial interval's rain, see HydroPol2D_Main_While.m:112, 734-740)
    rain = 20.0 / 3600.0 * np.concatenate([[spec.min_time_step_s] * 3, dts[2:-1]])
    p = _cls()
    v = valid.ravel(order="F")
    d = spec.d0.ravel(order="F")[v].copy()
    I = I0.ravel(order="F")[v].copy()
    T = np.zeros_like(d)
    for k in range(n_steps):
        d, I, T, _ = oracle.ga_step(p, k == 0, dts[k] / 3600.0, d + rain[k], I, T)
        d = np.where(d < 1e-6, 0.0, d)
    # (to the rounding of the rain depth, which the loop forms as avg * dt / 60)
    np.testing.assert_allclose(r["I_t"].ravel(order="F")[v], I, rtol=1e-13)
    np.testing.assert_allclose(r["d_t"].ravel(order="F")[v], d, rtol=1e-13, atol=1e-13)
    assert np.all(np.isnan(r["I_t"].ravel(order="F")[~v]))
