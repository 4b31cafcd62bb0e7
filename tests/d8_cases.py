"""Inputs shared by the D8 oracle pins and the D8 GPU parity tests
(Bates_Inertial_8D, CA_Routing_8D, build_up_wash_off_8D)."""
import numpy as np

PAGES = ("left", "right", "up", "down", "outlet", "ne", "se", "sw", "nw")

# np.rot90 (a quarter turn anticlockwise): the flow page each page becomes
# (east -> north, north -> west, west -> south, south -> east, NE -> NW, ...)
ROT90_PAGE = {"left": "down", "right": "up", "up": "left", "down": "right", "outlet": "outlet",
              "ne": "nw", "nw": "sw", "sw": "se", "se": "ne"}


def rot90_pages(a):
    """The 9-page stack of the rotated grid, from the pages of the original."""
    out = np.zeros((a.shape[1], a.shape[0], 9), order="F")
    for p, name in enumerate(PAGES):
        out[:, :, PAGES.index(ROT90_PAGE[name])] = np.rot90(a[:, :, p])
    return out


def rot90_outlets(rows, cols, ny, nx):
    """1-based outlet rows/cols after np.rot90 of an ny x nx grid."""
    r0, c0 = np.asarray(rows) - 1, np.asarray(cols) - 1
    return (nx - 1 - c0) + 1, r0 + 1


def rough_dem(ny, nx, seed, relief=2.0, nan_frac=0.0):
    rng = np.random.default_rng(seed)
    z = rng.random((ny, nx)) * relief + np.linspace(0, 1.0, nx)[None, :]
    if nan_frac:
        z[rng.random((ny, nx)) < nan_frac] = np.nan
    return np.asfortranarray(z)


def depths_mm(z, seed, wet_frac=0.8, scale=300.0):
    rng = np.random.default_rng(seed + 7)
    d = rng.random(z.shape) * scale
    d[rng.random(z.shape) > wet_frac] = 0.0
    d[np.isnan(z)] = np.nan
    return np.asfortranarray(d)


def wq_case(ny, nx, seed, outlet=True, edge_out=False, C3=5.0, C4=1.2, nan_frac=0.0):
    """Nine non-negative flow pages [mm/h] and a pollutant mass [kg]; the
    pages leaving the grid are zero unless edge_out (so the mass stays on
    the grid apart from what the outlet page carries)."""
    rng = np.random.default_rng(seed)
    q = rng.random((ny, nx, 9)) * 40.0
    q[rng.random((ny, nx, 9)) < 0.3] = 0.0
    if not edge_out:
        q[:, 0, 0] = 0.0
        q[:, -1, 1] = 0.0
        q[0, :, 2] = 0.0
        q[-1, :, 3] = 0.0
        for p, (di, dj) in zip((5, 6, 7, 8), ((-1, 1), (1, 1), (1, -1), (-1, -1))):
            if di < 0:
                q[0, :, p] = 0.0
            else:
                q[-1, :, p] = 0.0
            if dj < 0:
                q[:, 0, p] = 0.0
            else:
                q[:, -1, p] = 0.0
    if not outlet:
        q[:, :, 4] = 0.0
    else:
        keep = np.zeros((ny, nx), bool)
        keep[-1, -3:] = True
        q[:, :, 4] = np.where(keep, q[:, :, 4] + 5.0, 0.0)
    B = rng.random((ny, nx)) * 2.0 + 0.05
    mask = np.zeros((ny, nx), np.uint8)
    if nan_frac:
        m = rng.random((ny, nx)) < nan_frac
        q[m] = np.nan
        B[m] = np.nan
        mask[m] = 1
    C3a = np.asfortranarray(np.full((ny, nx), C3) * (1 + 0.2 * rng.random((ny, nx))))
    C4a = np.asfortranarray(np.full((ny, nx), C4))
    return (C3a, C4a, np.asfortranarray(q), np.asfortranarray(B), np.asfortranarray(mask))
