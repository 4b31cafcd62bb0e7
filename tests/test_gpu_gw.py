"""GPU parity of the Boussinesq groundwater kernels against the FP64 oracle.

Tolerance: the kernels perform the oracle's IEEE operations in the same
order (no transcendental functions, contraction off), so heads, fluxes,
velocities and exfiltration are required to match bit for bit; the scalar
error is compared to 1e-12 of the storage it is a difference of.
"""
import numpy as np
import pytest

import gw_cases
from oracle import oracle

pytestmark = pytest.mark.gpu


def gpu_solver(dt, dx, dy, h, z0, Sy, R, K, courant, h_soil, catchment_mask=None,
               dirichlet_mask=None, h_dirichlet=None):
    """The oracle's call shape on the MATLAB-signature mirror."""
    from hydropol2d_amd.groundwater import Boussinesq_2D_explicit as B
    h_t, u_x, u_y, q_exf, q_river, err = B(dt, dx, dy, h, z0, Sy, R, K, None, None, None, None,
                                           courant, h_soil, catchment_mask, dirichlet_mask,
                                           h_dirichlet, None)
    assert not q_river.any()
    return dict(h_t=h_t, u_x=u_x, u_y=u_y, q_exf=q_exf, error=err, n_sub=B.last_substeps)


def _case(seed, ny, nx):
    rng = np.random.default_rng(seed)
    z0 = rng.uniform(0, 1.0, (ny, nx)) + np.linspace(0, 2, nx)[None, :]
    h = z0 + rng.uniform(0.2, 3.0, (ny, nx))
    K = rng.uniform(1e-5, 5e-4, (ny, nx))
    Sy = rng.uniform(0.1, 0.3, (ny, nx))
    R = rng.uniform(0, 3e-6, (ny, nx))
    hs = rng.uniform(1.0, 4.0, (ny, nx))
    return z0, h, K, Sy, R, hs


def _sloped(seed, ny, nx, g=0.3, dx=5.0):
    # thin aquifer on a sloping bed: the Courant limit of the reference binds
    # (several sub-steps) while the donor cap keeps the scheme stable
    rng = np.random.default_rng(seed)
    x = np.arange(nx) * dx
    z0 = np.tile(100 - g * x, (ny, 1)) + rng.uniform(0, 0.01, (ny, nx))
    h = z0 + 0.05 + rng.uniform(0, 0.02, (ny, nx))
    K = rng.uniform(3e-3, 1e-2, (ny, nx))
    Sy = rng.uniform(0.15, 0.25, (ny, nx))
    return z0, h, K, Sy


def _same(got, ref):
    for k in ("h_t", "u_x", "u_y", "q_exf"):
        np.testing.assert_array_equal(got[k], ref[k], err_msg=k)
    assert got["n_sub"] == ref["n_sub"]
    scale = np.nansum(np.abs(ref["h_t"])) + 1.0
    assert abs(got["error"] - ref["error"]) <= 1e-12 * scale * 1e3


@pytest.mark.parametrize("seed,ny,nx,T,masks,kmul", [
    (0, 17, 23, 60.0, False, 1.0),
    (1, 64, 40, 600.0, True, 100.0),     # several sub-steps, catchment + Dirichlet
    (2, 2, 2, 10.0, False, 1.0),         # smallest grid the reference accepts
    (3, 130, 257, 1800.0, True, 100.0),  # ragged, many sub-steps, clamps firing
    (6, 40, 50, 1800.0, True, 1.0),      # exfiltration (thin soil), one sub-step
])
def test_dropin_matches_oracle(seed, ny, nx, T, masks, kmul):
    z0, h, K, Sy, R, hs = _case(seed, ny, nx)
    K = K * kmul
    cm = dm = hd = None
    if masks:
        cm = np.ones((ny, nx), dtype=bool)
        cm[: ny // 5, : nx // 4] = False
        h[~cm] = np.nan
        dm = np.zeros((ny, nx), dtype=bool)
        dm[:, -1] = True
        hd = z0 + 0.5
    ref = oracle.boussinesq_2d(T, 5.0, 4.0, h, z0, Sy, R, K, 0.25, hs, cm, dm, hd)
    got = gpu_solver(T, 5.0, 4.0, h, z0, Sy, R, K, 0.25, hs, cm, dm, hd)
    _same(got, ref)


def test_sloped_multistep_matches_oracle():
    z0, h, K, Sy = _sloped(7, 96, 160)
    R = np.full(h.shape, 1e-6)
    ref = oracle.boussinesq_2d(3600.0, 5.0, 5.0, h, z0, Sy, R, K, 0.25, 0.06)
    got = gpu_solver(3600.0, 5.0, 5.0, h, z0, Sy, R, K, 0.25, 0.06)
    assert ref["n_sub"] >= 5
    assert np.nanmax(ref["q_exf"]) > 0
    _same(got, ref)


def test_repeated_calls():
    z0, h, K, Sy, R, hs = _case(4, 33, 31)
    for T in (600.0, 5.0, 120.0):
        ref = oracle.boussinesq_2d(T, 5.0, 5.0, h, z0, Sy, R, K, 0.25, hs)
        got = gpu_solver(T, 5.0, 5.0, h, z0, Sy, R, K, 0.25, hs)
        _same(got, ref)
        h = ref["h_t"]


@pytest.mark.parametrize("ratio,expected", [(0.5, 1), (2.4, 5), (2.6, 4), (5.5, 6)])
def test_substep_schedule_on_gpu(ratio, expected):
    ny, nx = 4, 30
    x = np.arange(nx) * 10.0
    h = np.tile(50.0 - 0.01 * x, (ny, 1))
    K = np.full((ny, nx), 1e-3)
    ux = -(0.5 * (K[:, :-1] + K[:, 1:])) * ((h[:, 1:] - h[:, :-1]) / 10.0)
    dtg = np.min(0.25 * 10.0 / np.abs(ux))
    dm = np.ones((ny, nx), dtype=bool)
    got = gpu_solver(ratio * dtg, 10.0, 10.0, h, 0.0, 0.2, 0.0, K, 0.25, 100.0, None, dm, h)
    assert got["n_sub"] == expected


def test_hillslope_steady_on_gpu():
    res, ratio, passed = gw_cases.steady_dupuit(gpu_solver)
    ref, _, _ = gw_cases.steady_dupuit(gw_cases.oracle_solver)
    for a, b in zip(res, ref):
        np.testing.assert_array_equal(a["H_model"], b["H_model"])
    assert passed


def test_hillslope_transient_on_gpu():
    s, passed = gw_cases.transient_linearized(gpu_solver, duration_s=3 * 86400)
    r, _ = gw_cases.transient_linearized(gw_cases.oracle_solver, duration_s=3 * 86400)
    np.testing.assert_array_equal(s["q_model"], r["q_model"])
    assert s["rmse"].max() < 0.02 and s["max_err"].max() < 0.05


def test_large_grid_closed_box():
    # 2048^2, several sub-steps: storage conserved (size-independent) and
    # equal to the oracle bit for bit
    ny = nx = 2048
    z0, h, K, Sy = _sloped(1, ny, nx)
    got = gpu_solver(1800.0, 5.0, 5.0, h, z0, Sy, 0.0, K, 0.25, 100.0)
    assert got["n_sub"] >= 4
    s0 = np.sum(Sy * (h - z0))
    np.testing.assert_allclose(np.sum(Sy * np.maximum(got["h_t"] - z0, 0)), s0, rtol=1e-11)
    ref = oracle.boussinesq_2d(1800.0, 5.0, 5.0, h, z0, Sy, 0.0, K, 0.25, 100.0)
    _same(got, ref)


# The FP32 runs' form of the sub-step: two passes that recompute the face
# fluxes (FP32 face arithmetic and divergence on FP64 heads, thicknesses and
# head differences, reciprocal multiplies), the masked heads in two
# alternating buffers.  It must follow the oracle's FP64 head change to FP32
# precision over many sub-steps, with the same sub-step count, on catchment
# masks, Dirichlet cells, clamps and exfiltration, and with an even and an odd
# sub-step count.
RTOL32 = 1e-5


def _fp32_form(T, dx, dy, h, z0, Sy, R, K, hs, cm=None, dm=None, hd=None):
    from hydropol2d_amd.groundwater import Boussinesq_2D_explicit as B
    ref = oracle.boussinesq_2d(T, dx, dy, h, z0, Sy, R, K, 0.25, hs, cm, dm, hd)
    h_t, u_x, u_y, q_exf, _, err = B(T, dx, dy, h, z0, Sy, R, K, None, None, None, None, 0.25,
                                     hs, cm, dm, hd, None, precision="fp32")
    assert B.last_substeps == ref["n_sub"]
    np.testing.assert_array_equal(np.isnan(h_t), np.isnan(ref["h_t"]))
    dh_ref = ref["h_t"] - h
    scale = np.nanmax(np.abs(dh_ref)) + 1e-300
    assert np.nanmax(np.abs((h_t - h) - dh_ref)) <= RTOL32 * scale
    assert np.nanmax(np.abs(q_exf - ref["q_exf"])) <= RTOL32 * (np.nanmax(np.abs(ref["q_exf"])) + 1e-300)
    for got, r in ((u_x, ref["u_x"]), (u_y, ref["u_y"])):
        np.testing.assert_array_equal(np.isnan(got), np.isnan(r))
        assert np.nanmax(np.abs(got - r)) <= RTOL32 * (np.nanmax(np.abs(r)) + 1e-300)
    dS = np.nansum(np.abs(np.broadcast_to(Sy, h.shape) * dh_ref)) * dx * dy
    assert abs(err - ref["error"]) <= RTOL32 * dS + 1e-9
    return ref["n_sub"]


@pytest.mark.parametrize("seed,ny,nx,T,masks,kmul", [
    (0, 17, 23, 60.0, False, 1.0),
    (1, 64, 40, 600.0, True, 100.0),
    (3, 130, 257, 1800.0, True, 100.0),
    (6, 40, 50, 1800.0, True, 1.0),
])
def test_fp32_form_follows_oracle(seed, ny, nx, T, masks, kmul):
    z0, h, K, Sy, R, hs = _case(seed, ny, nx)
    K = K * kmul
    cm = dm = hd = None
    if masks:
        cm = np.ones((ny, nx), dtype=bool)
        cm[: ny // 5, : nx // 4] = False
        h[~cm] = np.nan
        dm = np.zeros((ny, nx), dtype=bool)
        dm[:, -1] = True
        hd = z0 + 0.5
    n_sub = _fp32_form(T, 5.0, 4.0, h, z0, Sy, R, K, hs, cm, dm, hd)
    print(f"fp32-form Boussinesq seed {seed}: {n_sub} sub-steps", flush=True)


def test_fp32_form_even_and_odd_substep_counts():
    # the final head sits in one of two buffers by the parity of the count
    z0, h, K, Sy = _sloped(7, 96, 160)
    R = np.full(h.shape, 1e-6)
    counts = set()
    for T in (600.0, 900.0, 1200.0, 3600.0):
        n_sub = _fp32_form(T, 5.0, 5.0, h, z0, Sy, R, K, 0.06)
        assert n_sub >= 2
        counts.add(n_sub % 2)
    assert counts == {0, 1}
