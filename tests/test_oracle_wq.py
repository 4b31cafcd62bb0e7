"""Pins of the FP64 CPU oracle's build-up / wash-off (A15) on the reference's
validation closed forms, its golden fixture and invariants.

Each check is chosen so that a plausible restatement mistake (a wrong
neighbour page, a lost 3600, a mis-scheduled sub-step, the mass model's B
factor dropped) fails it.
"""
import csv
import math
import os

import numpy as np
import pytest

import wq_cases
from oracle import oracle

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _rows(name):
    with open(os.path.join(GOLD, name)) as f:
        return list(csv.DictReader(f))


def test_dry_buildup_formula():
    # run_water_quality_washoff_validation.m:78-103: B0 = C1(1-e^{-C2 ADD}) A/1e4
    from hydropol2d_amd.water_quality import initial_buildup
    P = wq_cases.base_parameters()
    B = initial_buildup(P["max_buildup_kg_ha"], P["buildup_rate_per_day"],
                        P["antecedent_dry_days"], P["area_m2"])
    assert abs(B - P["initial_buildup_kg"]) < 1e-12
    assert initial_buildup([56.04, 10.0], [1.0, 0.5], 5.0, 400.0, [1, 0])[1] == 0.0


def test_single_cell_exponential_washoff():
    # :105-180 thresholds: RMSE < 1e-6 kg, exported-load error < 0.01 %,
    # mass residual < 1e-6 kg, no negative mass
    t, m, ref, conc, res = wq_cases.single_cell(wq_cases.oracle_washoff)
    rmse = float(np.sqrt(np.mean((m - ref) ** 2)))
    assert rmse < 1e-6, rmse
    P = wq_cases.base_parameters()
    exp_err = np.max(np.abs((P["initial_buildup_kg"] - m) - (P["initial_buildup_kg"] - ref)))
    assert exp_err / (P["initial_buildup_kg"] - ref[-1]) * 100 < 0.01
    assert np.max(np.abs(res)) < 1e-6
    assert (m >= -1e-12).all()
    # outlet concentration follows C0 e^{-lambda t}
    c_ref = P["target_initial_concentration_mg_L"] * np.exp(-P["lambda_h"] * t / 60)
    np.testing.assert_allclose(conc[1:], c_ref[1:], rtol=2e-4)


def test_two_cell_transfer():
    # :182-252: b1 = B0 e^{-lt}, b2 = e^{-lt}(B0 + l B0 t); RMSE of the total < 1e-6 kg
    t, b1, b2, r1, r2, exp, rexp, res = wq_cases.two_cell(wq_cases.oracle_washoff)
    rmse = float(np.sqrt(np.mean(((b1 + b2) - (r1 + r2)) ** 2)))
    assert rmse < 1e-6, rmse
    # each cell separately: a swapped inflow page would send cell 1's mass the wrong way
    assert np.max(np.abs(b1 - r1)) < 2e-6
    assert np.max(np.abs(b2 - r2)) < 2e-6
    assert np.max(np.abs(exp - rexp)) / rexp[-1] * 100 < 0.01
    assert np.max(np.abs(res)) < 1e-6
    assert (b1 >= -1e-12).all() and (b2 >= -1e-12).all()


def test_golden_buildup_washoff_fixture():
    rows = _rows("P1-WQ-001_reference.csv")
    model = wq_cases.fixture_buildup_washoff(wq_cases.oracle_washoff, rows)
    ref = np.array([float(r["surface_mass_kg"]) for r in rows])
    np.testing.assert_allclose(model, ref, rtol=1e-9, atol=1e-12)
    # the fixture's wash-off rate column is k*Q*B
    k, Q = float(rows[0]["washoff_coefficient"]), float(rows[0]["runoff_reference"])
    np.testing.assert_allclose(k * Q * model, [float(r["washoff_rate_kg_s"]) for r in rows],
                               rtol=1e-9)


def test_substep_single_cell_quirk():
    # a cell that only loses mass has dW > 0, so tmin_wq is the whole step; a
    # step long enough to overdraw it gives floor(1)+1 = 2 sub-steps of T/2 and 0
    A, q, lam, T = 400.0, 30.0, 3.0, 3000.0
    Qm3 = q / 1000 / 3600 * A
    r = oracle.washoff(lam / Qm3, 1.0, 0, 0, 0, 0, q, [[2.0]], T / 60, A)
    B_exp, steps, out = wq_cases.substep_chain_closed_form([2.0], [lam], T)
    assert steps == 2 and B_exp[0] < 0
    assert r["B_t"][0, 0] == 0.0
    np.testing.assert_allclose(r["mass_lost"], B_exp[0], rtol=1e-12)
    np.testing.assert_allclose(r["Tot_Washed"][0, 0], out[0], rtol=1e-12)


@pytest.mark.parametrize("T", [40.0, 95.0, 600.0])
def test_substep_schedule_two_cells(T):
    # upstream cell feeds a nearly empty downstream cell: dW < 0 there sets
    # tmin_wq = 3600 B2/|dW2| and the call sub-steps
    A, q = 400.0, 30.0
    Qm3 = q / 1000 / 3600 * A
    lam = np.array([6.0, 2.0])
    B0 = np.array([3.0, 0.05])
    C3 = (lam / Qm3).reshape(1, 2)
    qr = np.array([[q, 0.0]])
    qo = np.array([[0.0, q]])
    r = oracle.washoff(C3, 1.0, 0, qr, 0, 0, qo, B0.reshape(1, 2), T / 60, A)
    B_exp, steps, out = wq_cases.substep_chain_closed_form(B0, lam, T)
    assert steps >= 2
    np.testing.assert_allclose(r["B_t"][0], np.maximum(B_exp, 0), rtol=1e-11, atol=1e-14)
    np.testing.assert_allclose(r["Tot_Washed"][0], out, rtol=1e-11)
    # Out_Conc is the outlet cell's mean wash-off over the step per outlet volume
    np.testing.assert_allclose(r["Out_Conc"], 1000 * (out[1] / (T / 3600)) / (q / 1000 * A),
                               rtol=1e-11)


def test_no_substep_single_call_is_one_euler_step():
    A, q, lam = 400.0, 30.0, 0.5
    Qm3 = q / 1000 / 3600 * A
    r = oracle.washoff(lam / Qm3, 1.0, 0, 0, 0, 0, q, [[1.5]], 6.0, A)
    np.testing.assert_allclose(r["B_t"][0, 0], 1.5 * (1 - lam * 0.1), rtol=1e-14)
    # a cell that only loses mass (dW > 0) leaves tmin_wq at the whole step
    assert r["tmin_wq"] == 6.0 * 60
    # P_conc = 1e6 W / (max(q,10) A)
    np.testing.assert_allclose(r["P_conc"][0, 0], 1e6 * lam * 1.5 / (q * A), rtol=1e-14)


def _random_case(seed, ny=9, nx=7, signed=False):
    rng = np.random.default_rng(seed)
    lo = -50.0 if signed else 0.0
    qs = [rng.uniform(lo, 80.0, (ny, nx)) for _ in range(4)]
    qo = np.zeros((ny, nx))
    qo[ny - 1, rng.integers(0, nx)] = 60.0
    B = rng.uniform(0.0, 3.0, (ny, nx))
    C3 = rng.uniform(5.0, 40.0, (ny, nx))
    C4 = rng.uniform(0.8, 1.4, (ny, nx))
    return C3, C4, qs, qo, B


@pytest.mark.parametrize("ts", [0.01, 2.0])
def test_mass_conservation_closed_grid(ts):
    # no outlet and no flow across the domain edge: total mass is conserved
    # whatever the sub-stepping (ts = 2 min forces it)
    C3, C4, qs, qo, B = _random_case(3)
    ql, qr, qu, qd = qs
    ql[:, 0] = 0
    qr[:, -1] = 0
    qu[0, :] = 0
    qd[-1, :] = 0
    qo[:] = 0
    r = oracle.washoff(C3, C4, ql, qr, qu, qd, qo, B, ts, 25.0)
    lost = r["mass_lost"]
    np.testing.assert_allclose(r["B_t"].sum() - lost, B.sum(), rtol=1e-12)
    assert (r["B_t"] >= 0).all()


def test_mass_budget_with_outlet():
    # everything that leaves the grid leaves through the outlet page or an edge face
    C3, C4, qs, qo, B = _random_case(4)
    ql, qr, qu, qd = qs
    ts, A = 0.05, 25.0
    r = oracle.washoff(C3, C4, ql, qr, qu, qd, qo, B, ts, A)
    W = lambda q: C3 * (np.maximum(q, 0) / 1000 / 3600 * A) ** C4 * B
    left_edge = W(ql)[:, 0].sum() + W(qr)[:, -1].sum() + W(qu)[0, :].sum() + W(qd)[-1, :].sum()
    out_mass = (left_edge + W(qo).sum()) * ts / 60
    np.testing.assert_allclose(B.sum() - r["B_t"].sum() + r["mass_lost"], out_mass, rtol=1e-10)


def test_rating_curve_model_and_thresholds():
    # flag_wq_model = 1: W = C3 Q^C4 (1 + max(B_extra - Br, 0)) with B_extra
    # = 0 at or below Bmin and Bm at or above Bmax; no wash-off below min_Bt
    A = 100.0
    Bmin, Bmax, min_Bt = 20.0, 40.0, 15.0     # g/m2
    B = np.array([[1.0, 3.0, 5.0]])           # kg: 10, 30, 50 g/m2
    q = np.full((1, 3), 36.0)
    C3 = np.full((1, 3), 2.0)
    r = oracle.washoff(C3, 1.0, 0, 0, 0, 0, q, B, 0.001, A, flag_wq_model=1, Bmin=Bmin, Bmax=Bmax,
                       min_Bt=min_Bt)
    Q = 36.0 / 1000 / 3600 * A
    Br, Bm = Bmin / 1000 * A, Bmax / 1000 * A
    W = np.array([0.0, 2.0 * Q * (1 + (3.0 - Br)), 2.0 * Q * (1 + (Bm - Br))])
    np.testing.assert_allclose(r["tot_W_out"][0], W, rtol=1e-14)
    np.testing.assert_allclose(r["B_t"][0], B[0] - W * 0.001 / 60, rtol=1e-14)


def test_nan_mask_first_pass():
    C3, C4, qs, qo, B = _random_case(5)
    mask = np.zeros(B.shape, dtype=bool)
    mask[2:4, 1:3] = True
    r = oracle.washoff(C3, C4, *qs, qo, B, 0.01, 25.0, nan_mask=mask)
    assert (r["tot_W_out"][mask] == 0).all()
    # masked cells keep their mass minus nothing, plus what neighbours send in
    r0 = oracle.washoff(C3, C4, *qs, qo, B, 0.01, 25.0)
    assert not np.array_equal(r["B_t"], r0["B_t"])
    assert (r["B_t"][mask] >= B[mask] - 1e-15).all()


def test_negative_face_rates_carry_nothing():
    # signed inputs: a negative face rate is an inflow face and washes nothing off
    C3, C4, qs, qo, B = _random_case(6, signed=True)
    r = oracle.washoff(C3, C4, *qs, qo, B, 0.01, 25.0)
    r_clamped = oracle.washoff(C3, C4, *[np.maximum(q, 0) for q in qs], qo, B, 0.01, 25.0)
    np.testing.assert_array_equal(r["B_t"], r_clamped["B_t"])
    assert np.isfinite(r["B_t"]).all()
