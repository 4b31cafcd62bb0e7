"""GPU checks of the pinned closed forms and of the Inertial_Solver drop-in:

* hp2d_inertial_solver (Inertial_Solver.m:19-26) against the oracle on
  several pages, with Hf = 0 faces (Inf/NaN in the reference, zeroed only by
  its caller) and the n^2 map broadcast over the pages, plus its closed forms
  (frictionless and zero-depth faces);
* the device loop's volume_error on a flat closed basin under constant rain
  against the closed form of mass_balance_check.m:45-73 in the accepted dt
  sequence;
* the device loop's rollback on a lake at rest: the dt the catch block sets
  and the lowered max_dt that holds afterwards (HydroPol2D_Main_While.m:926-993).
"""
import numpy as np
import pytest

from oracle import oracle
import test_oracle_ledger as ledger

pytestmark = pytest.mark.gpu


def test_inertial_solver_dropin_matches_oracle():
    from hydropol2d_amd import local_inertial
    rng = np.random.default_rng(11)
    ny, nx = 37, 53
    for pages in (1, 2, 3):
        q_p = np.asfortranarray(rng.normal(0.0, 0.5, (ny, nx, pages)))
        Hf = np.asfortranarray(rng.uniform(0.0, 2.0, (ny, nx, pages)))
        Hf[rng.random(Hf.shape) < 0.2] = 0.0          # dry faces
        S = np.asfortranarray(rng.normal(0.0, 0.01, (ny, nx, pages)))
        n_sq = np.asfortranarray(rng.uniform(0.01, 0.1, (ny, nx)) ** 2)
        print(f"inertial_solver pages={pages}", flush=True)
        got = local_inertial.Inertial_Solver(1, q_p, 0.7, Hf, S, n_sq)
        ref = oracle.inertial_solver(q_p, 0.7, Hf, S, n_sq)
        fin = np.isfinite(ref)
        np.testing.assert_array_equal(np.isfinite(got), fin)
        np.testing.assert_allclose(got[fin], ref[fin], rtol=1e-12, atol=1e-15)
        # Hf = 0: 0/0 or x/Inf as in MATLAB: never a finite non-zero discharge
        dry = Hf == 0.0
        assert np.all((got[dry] == 0.0) | ~np.isfinite(got[dry]))


def test_inertial_solver_closed_forms():
    from hydropol2d_amd import local_inertial
    ny, nx = 8, 9
    q_p = np.asfortranarray(np.linspace(-1.0, 1.0, ny * nx * 2).reshape(ny, nx, 2))
    Hf = np.asfortranarray(np.full((ny, nx, 2), 0.8))
    S = np.asfortranarray(np.linspace(-0.02, 0.02, ny * nx * 2).reshape(ny, nx, 2))
    dt = 0.5
    # n = 0: q = q_p - g Hf dt S exactly
    got = local_inertial.Inertial_Solver(1, q_p, dt, Hf, S, np.zeros((ny, nx)))
    np.testing.assert_allclose(got, q_p - 9.81 * Hf * dt * S, rtol=1e-15, atol=1e-15)
    # q_p = 0 and S = 0: no flow
    got0 = local_inertial.Inertial_Solver(1, np.zeros_like(q_p), dt, Hf, np.zeros_like(S),
                                          np.full((ny, nx), 0.03 ** 2))
    assert np.all(got0 == 0.0)


def test_loop_volume_error_closed_form():
    from hydropol2d_amd.simulation import Simulation
    r_mmh, T = 36.0, 10.0
    spec = ledger.flat_basin(1, 3, r_mmh, 0.0, T)
    print("loop volume_error flat 1x3", flush=True)
    with Simulation(spec) as sim:
        sim.run()
        rep, dts_acc = sim.report(), sim.dt_history()
        d = sim.state()["d_t"]
    o = oracle.run(spec)
    np.testing.assert_array_equal(dts_acc, o["dt_hist"][-len(dts_acc):])
    n = rep["steps"]
    tcr = o["dt_hist"]
    mn = spec.min_time_step_s
    dts = np.concatenate([[mn], tcr[:-1]])
    rain = r_mmh / 3600.0 * np.concatenate([[mn] * 3, dts[2:n - 1]])
    depth = np.cumsum(rain)
    np.testing.assert_allclose(d.ravel(), depth[-1], rtol=1e-13)
    # the last report's volume error: storage since the previous report minus
    # the next step's rain over the grid
    t_end = (mn + np.cumsum(tcr)) / 60.0
    reps, nxt = [], spec.report_every_percent
    for j in range(n):
        if t_end[j] / T * 100.0 >= nxt or t_end[j] / T * 100.0 >= 100.0:
            reps.append(j)
            nxt += spec.report_every_percent
    assert rep["n_reports"] == len(reps)
    A = spec.resolution ** 2
    j1, j0 = reps[-1], reps[-2]
    expect = 3 * A * (depth[j1] - depth[j0]) / 1000.0 - 3 * A * (r_mmh / 3600.0 * dts[j1]) / 1000.0
    assert rep["volume_error_m3"] == pytest.approx(expect, rel=1e-9, abs=1e-12)


def test_loop_rollback_dt_closed_form():
    from hydropol2d_amd.simulation import Simulation
    spec = ledger.flat_basin(6, 7, 0.0, 50.0, 10.0)
    k_fail = 5
    print("loop rollback lake at rest", flush=True)
    with Simulation(spec) as sim:
        sim.inject_failure(k_fail)
        sim.run(40)
        rep, dts = sim.report(), sim.dt_history()
    assert rep["catch_index"] == 2 and rep["rejected_steps"] == 1 and rep["steps"] == 40
    nt = spec.alfa_min * spec.resolution / np.sqrt(9.81 * 0.05) / 2 / spec.factor_reduction
    inc = spec.time_step_increments_s
    assert np.all(dts[: k_fail - 1] == spec.max_time_step_s)
    np.testing.assert_allclose(dts[k_fail - 1:], np.floor(1.5 * nt / inc) * inc, rtol=1e-12)
