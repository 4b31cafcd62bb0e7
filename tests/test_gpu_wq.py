"""GPU parity of the build-up / wash-off kernels against the FP64 oracle.

Tolerance: 1e-12 relative on every output (FP64 throughout; the kernels
follow the oracle's operation order, the remaining difference is the device
pow against the host libm pow, about 1 ulp per call).
"""
import numpy as np
import pytest

import wq_cases
from oracle import oracle

pytestmark = pytest.mark.gpu

RTOL = 1e-12
KEYS = ("B_t", "P_conc", "Out_Conc", "tmin_wq", "tot_W_out", "mass_lost", "Tot_Washed")


def gpu_washoff(C3, C4, ql, qr, qu, qd, qo, B_t, time_step_min, cell_area, nan_mask=None,
                flag_wq_model=0, mass_lost=0.0, Tot_Washed=None, Bmin=0.0, Bmax=0.0, min_Bt=0.0):
    """The oracle's call shape on the MATLAB-signature mirror."""
    from hydropol2d_amd.water_quality import build_up_wash_off
    B = np.asarray(B_t, dtype=np.float64)
    B = B.reshape(1, -1) if B.ndim == 1 else B
    ny, nx = B.shape
    nan5 = None if nan_mask is None else np.repeat(np.asarray(nan_mask)[:, :, None], 5, axis=2)
    out = build_up_wash_off(C3, C4, ql, qr, qu, qd, qo, B, time_step_min, nx, ny, cell_area,
                            None, nan5, flag_wq_model, mass_lost,
                            np.zeros((ny, nx)) if Tot_Washed is None else Tot_Washed,
                            Bmin, Bmax, min_Bt)
    return dict(zip(KEYS, out))


def _random_case(seed, ny, nx, signed=False):
    rng = np.random.default_rng(seed)
    lo = -50.0 if signed else 0.0
    qs = [rng.uniform(lo, 80.0, (ny, nx)) * (rng.uniform(size=(ny, nx)) > 0.2) for _ in range(4)]
    qo = np.zeros((ny, nx))
    qo[ny - 1, rng.integers(0, nx, size=3)] = rng.uniform(10, 90, 3)
    # no zero mass: a zero-mass cell that receives wash-off gives tmin_wq = 0
    # (an endless sub-step loop in the reference) unless min_Bt > 0
    B = rng.uniform(0.01, 3.0, (ny, nx))
    C3 = rng.uniform(5.0, 40.0, (ny, nx))
    C4 = rng.uniform(0.8, 1.4, (ny, nx))
    return C3, C4, qs, qo, B


def _compare(a, b):
    for k in KEYS:
        np.testing.assert_allclose(np.asarray(a[k]), np.asarray(b[k]), rtol=RTOL, atol=1e-300,
                                   err_msg=k)


@pytest.mark.parametrize("seed,ny,nx,ts,area,signed,model,min_bt,masked", [
    (0, 17, 13, 0.01, 4.0, False, 0, 0.0, False),
    (1, 64, 50, 0.5, 4.0, True, 0, 0.0, False),      # signed local-inertial inputs
    (2, 33, 71, 5.0, 400.0, False, 0, 0.0, False),   # long step: ~4 sub-steps
    (3, 40, 40, 0.02, 4.0, False, 1, 5.0, False),    # rating curve + min_Bt
    (4, 25, 30, 0.02, 4.0, False, 0, 0.0, True),     # idx_nan mask
    (5, 300, 257, 1.0, 400.0, True, 1, 2.0, True),   # everything, ragged size, sub-steps
    (6, 1, 1, 0.2, 4.0, False, 0, 0.0, False),       # single cell
    (9, 20, 20, 3.0, 400.0, False, 0, 50.0, False),  # most cells below min_Bt
])
def test_dropin_matches_oracle(seed, ny, nx, ts, area, signed, model, min_bt, masked):
    C3, C4, qs, qo, B = _random_case(seed, ny, nx, signed)
    mask = None
    if masked:
        mask = np.zeros((ny, nx), dtype=bool)
        mask[: max(ny // 4, 1), : max(nx // 5, 1)] = True
    kw = dict(nan_mask=mask, flag_wq_model=model, mass_lost=-0.25,
              Tot_Washed=np.full((ny, nx), 0.5), Bmin=3.0, Bmax=20.0, min_Bt=min_bt)
    ref = oracle.washoff(C3, C4, *qs, qo, B, ts, area, **kw)
    got = gpu_washoff(C3, C4, *qs, qo, B, ts, area, **kw)
    _compare(got, ref)


def test_repeated_calls_reuse_workspace():
    # the handle's workspace is reused across calls of different step lengths
    # min_Bt > 0 keeps cells emptied by one call out of the next call's tmin_wq
    C3, C4, qs, qo, B = _random_case(7, 48, 48)
    for ts in (5.0, 0.01, 2.0):
        ref = oracle.washoff(C3, C4, *qs, qo, B, ts, 400.0, min_Bt=0.1)
        got = gpu_washoff(C3, C4, *qs, qo, B, ts, 400.0, min_Bt=0.1)
        _compare(got, ref)
        B = ref["B_t"]


def test_validation_single_cell_on_gpu():
    # run_water_quality_washoff_validation.m:105-180 thresholds on the GPU path,
    # and the GPU series equal to the oracle's
    t, m, ref, conc, res = wq_cases.single_cell(gpu_washoff, duration_min=5.0)
    t2, m2, _, conc2, _ = wq_cases.single_cell(wq_cases.oracle_washoff, duration_min=5.0)
    np.testing.assert_allclose(m, m2, rtol=RTOL)
    np.testing.assert_allclose(conc, conc2, rtol=RTOL)
    assert float(np.sqrt(np.mean((m - ref) ** 2))) < 1e-6
    assert np.max(np.abs(res)) < 1e-6


def test_validation_two_cell_on_gpu():
    out = wq_cases.two_cell(gpu_washoff, duration_min=5.0)
    ref = wq_cases.two_cell(wq_cases.oracle_washoff, duration_min=5.0)
    np.testing.assert_allclose(out[1], ref[1], rtol=RTOL)
    np.testing.assert_allclose(out[2], ref[2], rtol=RTOL)
    t, b1, b2, r1, r2, exp, rexp, res = out
    assert float(np.sqrt(np.mean(((b1 + b2) - (r1 + r2)) ** 2))) < 1e-6


def test_golden_fixture_on_gpu():
    import csv
    import os
    with open(os.path.join(os.path.dirname(__file__), "golden", "P1-WQ-001_reference.csv")) as f:
        rows = list(csv.DictReader(f))[:4]
    model = wq_cases.fixture_buildup_washoff(gpu_washoff, rows)
    np.testing.assert_allclose(model, [float(r["surface_mass_kg"]) for r in rows], rtol=1e-9)


def test_zero_mass_cells_below_min_bt():
    C3, C4, qs, qo, B = _random_case(10, 30, 30)
    B[::3, ::2] = 0.0
    ref = oracle.washoff(C3, C4, *qs, qo, B, 2.0, 400.0, min_Bt=0.01)
    got = gpu_washoff(C3, C4, *qs, qo, B, 2.0, 400.0, min_Bt=0.01)
    _compare(got, ref)


def test_endless_substepping_is_refused():
    # tmin_wq = 0 (a zero-mass cell receiving wash-off, min_Bt = 0): the
    # library returns an error instead of looping
    from hydropol2d_amd import _lib
    C3, C4, qs, qo, B = _random_case(11, 8, 8)
    B[4, 4] = 0.0
    with pytest.raises(_lib.HP2DError):
        gpu_washoff(C3, C4, *qs, qo, B, 2.0, 4.0)


def test_large_grid_mass_conservation():
    # 2048^2 closed grid (no outlet, no flow across the edge), sub-stepping
    # step: total mass - mass_lost is conserved (size-independent property)
    ny = nx = 2048
    C3, C4, qs, qo, B = _random_case(8, ny, nx)
    ql, qr, qu, qd = qs
    ql[:, 0] = 0
    qr[:, -1] = 0
    qu[0, :] = 0
    qd[-1, :] = 0
    qo[:] = 0
    got = gpu_washoff(C3, C4, ql, qr, qu, qd, qo, B, 3.0, 400.0)
    np.testing.assert_allclose(got["B_t"].sum() - got["mass_lost"], B.sum(), rtol=1e-11)
    assert (got["B_t"] >= 0).all()
    ref = oracle.washoff(C3, C4, ql, qr, qu, qd, qo, B, 3.0, 400.0)
    _compare(got, ref)
