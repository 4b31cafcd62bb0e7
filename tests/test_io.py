"""The on-disk formats (host code, no GPU): the GeoTIFF reader against Pillow's
decoder on the reference's own rasters and on files of every supported
encoding, and the Level 5 MAT-file writer against SciPy's reader, including
the Maps / TempTS layouts of save_output_maps.m."""
import os
import struct
import zlib

import numpy as np
import pytest
from PIL import Image
from scipy.io import loadmat

from hydropol2d_amd import io as hio

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("name", ["Static_DEM.tif", "Static_LULC.tif"])
def test_reference_rasters_match_pillow(name):
    path = os.path.join(GOLDEN, name)
    g = hio.GRIDobj(path)
    ref = np.array(Image.open(path), dtype=np.float64)
    np.testing.assert_array_equal(np.nan_to_num(g.Z, nan=-1e300), np.nan_to_num(ref, nan=-1e300))
    assert g.Z.shape == (100, 50)
    # ModelPixelScale (20, 20) and the tie point (0, 0) -> (0, 2000)
    assert (g.cellsize, g.dy, g.x_ul, g.y_ul) == (20.0, 20.0, 0.0, 2000.0)
    np.testing.assert_allclose(g.x()[:2], [10.0, 30.0])
    np.testing.assert_allclose(g.y()[:2], [1990.0, 1970.0])


@pytest.mark.parametrize("comp", ["raw", "packbits", "tiff_lzw", "tiff_deflate"])
def test_pillow_encodings(tmp_path, comp):
    rng = np.random.default_rng(1)
    a = rng.normal(100.0, 30.0, (37, 53)).astype(np.float32)
    a[3, 4] = np.nan
    p = str(tmp_path / f"x_{comp}.tif")
    Image.fromarray(a, mode="F").save(p, compression=comp)
    g = hio.GRIDobj(p)
    np.testing.assert_array_equal(np.isnan(g.Z), np.isnan(a))
    np.testing.assert_array_equal(g.Z[~np.isnan(a)], a[~np.isnan(a)].astype(np.float64))


def _write_tiff(path, a, *, tile=None, predictor=1, deflate=False, big_endian=False,
                nodata=None, tie=(0, 0, 0, 500000.0, 4100000.0, 0), scale=(30.0, 30.0, 0.0),
                epsg=32611):
    """A minimal single-band GeoTIFF writer (test infrastructure)."""
    bo = ">" if big_endian else "<"
    a = np.ascontiguousarray(a)
    dt = a.dtype.newbyteorder(bo)
    fmt = {"f": 3, "i": 2, "u": 1}[a.dtype.kind]
    h, w = a.shape
    nb = a.dtype.itemsize

    def encode(block):
        bh, bw = block.shape
        raw = bytearray(block.astype(dt).tobytes())
        if predictor == 2:
            s = np.frombuffer(bytes(raw), dtype=dt).reshape(bh, bw).astype(np.int64)
            d = s.copy()
            d[:, 1:] = s[:, 1:] - s[:, :-1]
            raw = bytearray(d.astype(dt).tobytes())
        elif predictor == 3:
            out = bytearray()
            be_bytes = np.frombuffer(block.astype(a.dtype.newbyteorder(">")).tobytes(),
                                     dtype=np.uint8).reshape(bh, bw, nb)
            for r in range(bh):
                row = be_bytes[r].T.reshape(-1)          # byte planes, most significant first
                diff = np.diff(row.astype(np.int16), prepend=0).astype(np.uint8)
                out += diff.tobytes()
            raw = out
        return zlib.compress(bytes(raw)) if deflate else bytes(raw)

    blocks, (bw, bh) = [], (tile if tile else (w, h))
    if tile:
        for by in range(0, h, bh):
            for bx in range(0, w, bw):
                blk = np.zeros((bh, bw), dtype=a.dtype)
                sub = a[by:by + bh, bx:bx + bw]
                blk[:sub.shape[0], :sub.shape[1]] = sub
                blocks.append(encode(blk))
    else:
        blocks.append(encode(a))
    data = b"".join(blocks)
    offs, o = [], 8
    for b in blocks:
        offs.append(o)
        o += len(b)
    entries = []
    extra = bytearray()
    base = 8 + len(data)

    def add(tag, typ, vals):
        size = {3: 2, 4: 4, 12: 8, 2: 1}[typ]
        code = {3: "H", 4: "I", 12: "d"}.get(typ)
        payload = vals if typ == 2 else struct.pack(bo + code * len(vals), *vals)
        cnt = len(vals)
        entries.append((tag, typ, cnt, payload, size * cnt))

    add(256, 4, [w]); add(257, 4, [h]); add(258, 3, [8 * nb]); add(259, 3, [8 if deflate else 1])
    add(262, 3, [1]); add(277, 3, [1]); add(284, 3, [1]); add(317, 3, [predictor]); add(339, 3, [fmt])
    if tile:
        add(322, 4, [bw]); add(323, 4, [bh]); add(324, 4, offs); add(325, 4, [len(b) for b in blocks])
    else:
        add(273, 4, offs); add(278, 4, [h]); add(279, 4, [len(data)])
    add(33550, 12, list(scale)); add(33922, 12, list(tie))
    add(34735, 3, [1, 1, 0, 1, 3072, 0, 1, epsg])
    if nodata is not None:
        add(42113, 2, str(nodata).encode() + b"\0")
    entries.sort(key=lambda e: e[0])
    ifd_at = base
    n_ent = len(entries)
    extra_at = ifd_at + 2 + 12 * n_ent + 4
    ifd = struct.pack(bo + "H", n_ent)
    for tag, typ, cnt, payload, size in entries:
        if size <= 4:
            ifd += struct.pack(bo + "HHI", tag, typ, cnt) + payload.ljust(4, b"\0")
        else:
            ifd += struct.pack(bo + "HHII", tag, typ, cnt, extra_at + len(extra))
            extra += payload
            if len(extra) % 2:
                extra += b"\0"
    ifd += struct.pack(bo + "I", 0)
    head = (b"MM" if big_endian else b"II") + struct.pack(bo + "HI", 42, ifd_at)
    with open(path, "wb") as f:
        f.write(head + data + ifd + extra)


@pytest.mark.parametrize("kw", [
    dict(tile=(16, 16), nodata=-32768),                                   # tiled int16
    dict(tile=(16, 8), predictor=2, deflate=True, nodata=-32768),         # + horizontal predictor
    dict(predictor=2, deflate=True, big_endian=True, nodata=-32768),      # big-endian strips
])
def test_int16_tiles_predictor_nodata(tmp_path, kw):
    rng = np.random.default_rng(2)
    a = rng.integers(-500, 3000, (45, 61)).astype(np.int16)
    a[0, :5] = -32768
    p = str(tmp_path / "i16.tif")
    _write_tiff(p, a, **kw)
    g = hio.GRIDobj(p)
    want = a.astype(np.float64)
    want[a == -32768] = np.nan
    np.testing.assert_array_equal(np.isnan(g.Z), np.isnan(want))
    np.testing.assert_array_equal(g.Z[~np.isnan(want)], want[~np.isnan(want)])
    assert g.epsg == 32611 and g.cellsize == 30.0
    assert (g.x_ul, g.y_ul) == (500000.0, 4100000.0)


@pytest.mark.parametrize("big_endian", [False, True])
def test_float_predictor(tmp_path, big_endian):
    rng = np.random.default_rng(3)
    a = rng.normal(0, 1e3, (20, 33)).astype(np.float32)
    p = str(tmp_path / "fp.tif")
    _write_tiff(p, a, predictor=3, deflate=True, big_endian=big_endian)
    np.testing.assert_array_equal(hio.GRIDobj(p).Z, a.astype(np.float64))


def test_crop_trims_the_nan_border():
    Z = np.full((6, 7), np.nan)
    Z[2:4, 1:5] = 1.0
    g = hio.GRIDobj(Z=Z, cellsize=10.0, x_ul=100.0, y_ul=900.0)
    c = g.crop()
    assert c.Z.shape == (2, 4)
    assert (c.x_ul, c.y_ul) == (110.0, 880.0)


def test_unsupported_tiff_is_refused(tmp_path):
    from hydropol2d_amd._lib import HP2DError
    p = tmp_path / "bad.tif"
    p.write_bytes(b"II*\0" + b"\0" * 20)
    with pytest.raises(HP2DError):
        hio.GRIDobj(str(p))


def test_mat_writer_round_trip(tmp_path):
    rng = np.random.default_rng(4)
    d = rng.normal(size=(5, 7, 3))
    r = rng.random((5, 7, 3)).astype(np.float32)
    p = str(tmp_path / "m.mat")
    hio.write_mat(p, {"Maps.Hydro.d": d, "Maps.Hydro.risk": r, "Maps.WQ_States.Pol_Mass_Map": d[:, :, :1],
                      "k": 17.0, "label": "HydroPol2D", "mask": np.eye(3, dtype=bool),
                      "idx": np.arange(4, dtype=np.int32)})
    m = loadmat(p, squeeze_me=False)
    H = m["Maps"]["Hydro"][0, 0]
    np.testing.assert_array_equal(H["d"][0, 0], d)
    got_r = H["risk"][0, 0]
    assert got_r.dtype == np.float32
    np.testing.assert_array_equal(got_r, r)
    np.testing.assert_array_equal(m["Maps"]["WQ_States"][0, 0]["Pol_Mass_Map"][0, 0], d[:, :, :1])
    assert m["k"][0, 0] == 17.0
    assert str(m["label"][0]) == "HydroPol2D"
    np.testing.assert_array_equal(m["mask"], np.eye(3, dtype=np.uint8))
    np.testing.assert_array_equal(m["idx"].ravel(), np.arange(4))


def test_map_stacks_in_chunks_of_twelve(tmp_path):
    rng = np.random.default_rng(5)
    n = 15
    rec = {"t_min": np.arange(n, dtype=float), "d": rng.random((4, 6, n)),
           "velocity": rng.random((4, 6, n)), "hazard_dv": rng.random((4, 6, n))}
    paths = hio.save_map_stacks(str(tmp_path), rec)
    assert [os.path.basename(p) for p in paths] == ["save_map_hydro_1.mat", "save_map_hydro_2.mat"]
    m1, m2 = (loadmat(p)["Maps"]["Hydro"][0, 0] for p in paths)
    assert m1["d"][0, 0].shape == (4, 6, 12) and m2["d"][0, 0].shape == (4, 6, 3)
    np.testing.assert_allclose(m2["velocity"][0, 0], rec["velocity"][:, :, 12:].astype(np.float32))
