"""GPU parity of the D8 drop-ins (Bates_Inertial_8D, CA_Routing_8D,
build_up_wash_off_8D and the eight-direction inflow gather) against the FP64
oracle, through the C-ABI (the MATLAB-signature mirrors).

Tolerances: the kernels evaluate the reference arithmetic in the same order
as the oracle (no FMA contraction), so the pages agree to 1e-12 relative
(pow and sqrt of the device library may differ from glibc in the last bit);
the gather adds in the same order and is bit-exact.  At 2048 x 2048 the
routing + gather is checked through volume conservation in a closed basin."""
import numpy as np
import pytest

import d8_cases as dc
from hydropol2d_amd import local_inertial as li
from hydropol2d_amd import water_quality as wq
from oracle import oracle

pytestmark = pytest.mark.gpu

RES = (None,) * 12  # the twelve reservoir arguments (flag_reservoir = 0)
SHAPES = [(1, 17), (23, 1), (2, 2), (3, 5), (64, 64), (131, 77)]


def _close(a, b, name, rtol=1e-12):
    a, b = np.asarray(a, float), np.asarray(b, float)
    np.testing.assert_array_equal(np.isnan(a), np.isnan(b), err_msg=name)
    scale = np.nanmax(np.abs(b)) if np.isfinite(b).any() else 1.0
    np.testing.assert_allclose(a, b, rtol=rtol, atol=rtol * (scale + 1e-300), equal_nan=True,
                               err_msg=name)


def _outlets(ny, nx):
    return [ny, max(1, ny - 1)], [nx, nx]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("outlet_type,critical", [(1, 0), (2, 1), (2, 0)])
def test_bates8d_matches_oracle(shape, outlet_type, critical):
    ny, nx = shape
    print(f"bates8d {shape} outlet {outlet_type} critical {critical}", flush=True)
    z = dc.rough_dem(ny, nx, 31 + ny, nan_frac=0.05 if ny * nx > 8 else 0.0)
    d_tot = dc.depths_mm(z, 3)
    d_p = dc.depths_mm(z, 4)
    rows, cols = _outlets(ny, nx)
    rough = np.asfortranarray(0.02 + 0.05 * np.random.default_rng(5).random((ny, nx)))
    qprev = np.asfortranarray(np.random.default_rng(6).random((ny, nx, 9)) * 300.0)
    qprev[np.isnan(z)] = np.nan
    o = oracle.bates_8d(z, d_tot, d_p, rough, 4.0, 0.02, 2.0, outlet_type, 0.01, rows, cols,
                        qprev, critical)
    g = li.Bates_Inertial_8D(*RES, 0, z, d_tot, d_p, rough, 4.0, 0.02, 2.0, None, outlet_type,
                             0.01, rows, cols, 1.0, qprev, None, critical)
    for k, name in enumerate(oracle.D8_PAGES):
        _close(g[k], o["outflow"][:, :, k], name)
    _close(g[9], o["d_t"], "d_t")
    _close(g[10], o["I_tot_end_cell"], "I_tot_end_cell")
    _close(g[11], o["outflow"], "outflow")
    _close(g[12], o["Hf"], "Hf")


def test_bates8d_time_loop_matches_oracle():
    # ten calls feeding each call's outflow into the next (d_t becomes d_tot
    # after the caller's gather, d_p the previous depth)
    ny, nx = 40, 57
    z = dc.rough_dem(ny, nx, 9, relief=0.3)
    d = dc.depths_mm(z, 9, wet_frac=1.0, scale=200.0)
    rows, cols = [ny], [nx]
    qg = qo = np.zeros((ny, nx, 9))
    dg = do = d
    dpg = dpo = d
    for _ in range(10):
        o = oracle.bates_8d(z, do, dpo, 0.04, 4.0, 0.005, 2.0, 1, 0.02, rows, cols, qo)
        dpo = do
        _, do = oracle.gather(o["outflow"], 0.005, o["d_t"])
        qo = o["outflow"]
        g = li.Bates_Inertial_8D(*RES, 0, z, dg, dpg, 0.04, 4.0, 0.005, 2.0, None, 1, 0.02,
                                 rows, cols, 1.0, qg, None, 0)
        dpg = dg
        _, dg = li.outflows_to_inflows(g[11], 0.005, g[9])
        qg = g[11]
    _close(qg, qo, "outflow", 1e-10)
    _close(dg, do, "d", 1e-10)
    assert np.nanmax(qo) > 10.0


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("outlet_type,critical,reset", [(1, 0, True), (2, 1, False),
                                                        (1, 1, False)])
def test_ca8d_matches_oracle(shape, outlet_type, critical, reset):
    ny, nx = shape
    print(f"ca8d {shape} outlet {outlet_type} critical {critical} reset {reset}", flush=True)
    z = dc.rough_dem(ny, nx, 41 + nx, nan_frac=0.05 if ny * nx > 8 else 0.0)
    d = dc.depths_mm(z, 12)
    rows, cols = _outlets(ny, nx)
    rough = np.asfortranarray(0.02 + 0.05 * np.random.default_rng(7).random((ny, nx)))
    h0 = np.asfortranarray(np.random.default_rng(8).random((ny, nx)) * 5.0)
    Ip = np.zeros((ny, nx)) if reset else np.random.default_rng(9).random((ny, nx)) * 0.05
    o = oracle.ca_8d(z, d, rough, 4.0, 0.02, h0, 2.0, Ip, outlet_type, 0.01, rows, cols,
                     critical, 1.5)
    g = li.CA_Routing_8D(*RES, 0, z, d, rough, 4.0, 0.02, h0, 2.0, Ip, None, outlet_type, 0.01,
                         rows, cols, None, critical, 1.5)
    for k, name in enumerate(oracle.D8_PAGES):
        _close(g[k], o["outflow"][:, :, k], name)
    _close(g[9], o["d_t"], "d_t")
    _close(g[10], o["I_tot_end_cell"], "I_tot_end_cell")


@pytest.mark.parametrize("shape", [(1, 9), (7, 1), (33, 65)])
@pytest.mark.parametrize("pages", [5, 9])
def test_gather_is_bit_exact(shape, pages):
    ny, nx = shape
    rng = np.random.default_rng(ny * nx + pages)
    q = np.asfortranarray(rng.random((ny, nx, pages)) * 100.0)
    d = np.asfortranarray(rng.random((ny, nx)) * 50.0)
    qo, do = oracle.gather(q, 0.07, d)
    qg, dg = li.outflows_to_inflows(q, 0.07, d)
    np.testing.assert_array_equal(qg, qo)
    np.testing.assert_array_equal(dg, do)


def test_ca8d_large_closed_basin_conserves_volume():
    # 2048 x 2048: routing + gather move water without creating or losing it
    import ctypes as C

    from hydropol2d_amd import _lib
    ny = nx = 2048
    print("ca8d 2048^2 closed basin", flush=True)
    z = dc.rough_dem(ny, nx, 77, relief=0.5)
    d = dc.depths_mm(z, 77, wet_frac=0.9, scale=100.0)
    Ip = np.zeros((ny, nx))
    v0 = d.sum()
    for _ in range(3):
        g = li.CA_Routing_8D(*RES, 0, z, d, 0.03, 4.0, 0.01, None, 2.0, Ip, None, 1, 0.0, [], [],
                             None, 0, 1.0)
        assert g[9].min() >= -1e-9
        q = np.stack(g[:9], axis=2)
        _, d = li.outflows_to_inflows(q, 0.01, g[9])
        Ip = g[10]
    assert d.sum() == pytest.approx(v0, rel=1e-11)
    L = _lib.load()
    h = li._Handle.get(ny, nx, 0)
    t = C.c_double()
    _lib.check(L.hp2d_last_kernel_time(h, C.byref(t)), handle=h)
    assert 0.0 < t.value < 50.0


@pytest.mark.parametrize("C3,model,nan_frac", [(2e3, 0, 0.0), (1e6, 0, 0.0), (2e5, 1, 0.05),
                                               (5e4, 0, 0.05)])
@pytest.mark.parametrize("shape", [(1, 13), (9, 11), (70, 45)])
def test_washoff8d_matches_oracle(C3, model, nan_frac, shape):
    ny, nx = shape
    print(f"washoff8d {shape} C3 {C3} model {model}", flush=True)
    C3a, C4a, q, B, mask = dc.wq_case(ny, nx, ny + nx, outlet=True, C3=C3, nan_frac=nan_frac,
                                      edge_out=True)
    kw = dict(Bmin=50.0, Bmax=400.0, min_Bt=0.05) if model == 1 else dict(min_Bt=0.01)
    tw0 = np.full((ny, nx), 0.1)
    o = oracle.washoff_8d(C3a, C4a, q, B, 0.4, 4.0, nan_mask=mask, flag_wq_model=model,
                          mass_lost=-0.5, Tot_Washed=tw0, **kw)
    g = wq.build_up_wash_off_8D(C3a, C4a, *[q[:, :, k] for k in range(9)], B, 0.4, nx, ny, 4.0,
                                None, mask, model, -0.5, tw0, kw.get("Bmin", 0.0),
                                kw.get("Bmax", 0.0), kw["min_Bt"])
    _close(g[0], o["B_t"], "B_t", 1e-11)
    _close(g[1], o["P_conc"], "P_conc", 1e-11)
    assert g[2] == pytest.approx(o["Out_Conc"], rel=1e-11)
    assert g[3] == pytest.approx(o["tmin_wq"], rel=1e-11)
    _close(g[4], o["tot_W_out"], "tot_W_out", 1e-11)
    assert g[5] == pytest.approx(o["mass_lost"], rel=1e-11, abs=1e-14)
    _close(g[6], o["Tot_Washed"], "Tot_Washed", 1e-11)
