"""CPU model of the stage kernel's launch geometry (sbp_capi.hip compute_geometry,
sbp_kernels.hpp stage2d_kernel / edge_point / interior_tile / dline).

Checks, for many grid shapes and all orders, that
* every global load of the interior tiles and of the table-driven edge rows
  stays inside [0,nx) x [0,ny);
* every grid point is written exactly once per stage (edge frame + interior
  tiles partition the grid);
* the XCD-aware block remap is a bijection onto the tiles.
"""
import itertools

import numpy as np
import pytest

BX = 256
NXCD = 8
CLS = {2: 1, 4: 4, 6: 6}
R = {2: 1, 4: 2, 6: 3}
TAPS = {2: 3, 4: 6, 6: 9}


def default_ty(nx, ny, cls):
    wi, hi = nx - 2 * cls, ny - 2 * cls
    if wi <= 0 or hi <= 0:
        return 16
    tx = (wi + BX - 1) // BX
    best = 4
    t = 4
    while t <= 64:
        if tx * ((hi + t - 1) // t) >= 960:
            best = t
        t *= 2
    return best


def geometry(nx, ny, order, ty=None):
    cls = CLS[order]
    ty = ty or default_ty(nx, ny, cls)
    wi, hi = nx - 2 * cls, ny - 2 * cls
    tiles_x = (wi + BX - 1) // BX if wi > 0 else 0
    tiles_y = (hi + ty - 1) // ty if hi > 0 else 0
    n_edge = 2 * cls * nx + 2 * cls * (ny - 2 * cls)
    eb = (n_edge + BX - 1) // BX
    neb = (eb + NXCD - 1) // NXCD * NXCD
    return dict(cls=cls, ty=ty, tiles_x=tiles_x, tiles_y=tiles_y, n_tiles=tiles_x * tiles_y,
                n_edge=n_edge, n_edge_blocks=neb, grid=neb + tiles_x * tiles_y)


def edge_decode(e, nx, ny, cls):
    nstrip = cls * nx
    nside = cls * (ny - 2 * cls)
    if e < nstrip:
        return e % nx, e // nx
    if e < 2 * nstrip:
        e -= nstrip
        return e % nx, ny - cls + e // nx
    if e < 2 * nstrip + nside:
        e -= 2 * nstrip
        return e % cls, cls + e // cls
    e -= 2 * nstrip + nside
    return nx - cls + e % cls, cls + e // cls


def dline_range(n, idx, order):
    cls, r, taps = CLS[order], R[order], TAPS[order]
    if idx < cls:
        start = 0
    elif idx >= n - cls:
        start = n - taps
    else:
        start = idx - r
    return start, start + taps - 1


def xcd_tile(bi, nt):
    xcd, local = bi % NXCD, bi // NXCD
    q, r = nt // NXCD, nt % NXCD
    return xcd * q + min(xcd, r) + local


def check(nx, ny, order, ty=None):
    g = geometry(nx, ny, order, ty)
    cls, r = g["cls"], R[order]
    count = np.zeros((ny, nx), dtype=np.int32)
    # edge blocks
    for b in range(g["n_edge_blocks"]):
        for t in range(BX):
            e = b * BX + t
            if e >= g["n_edge"]:
                continue
            i, j = edge_decode(e, nx, ny, cls)
            assert 0 <= i < nx and 0 <= j < ny
            count[j, i] += 1
            for n, idx in ((nx, i), (ny, j)):
                lo, hi = dline_range(n, idx, order)
                assert 0 <= lo and hi < n, (nx, ny, order, i, j)
    # interior tiles
    tiles = [xcd_tile(bi, g["n_tiles"]) for bi in range(g["n_tiles"])]
    assert sorted(tiles) == list(range(g["n_tiles"]))
    for tile in range(g["n_tiles"]):
        tx, tyb = tile % g["tiles_x"], tile // g["tiles_x"]
        x0 = cls + tx * BX
        j0 = cls + tyb * g["ty"]
        j1 = min(j0 + g["ty"], ny - cls)
        assert j1 > j0
        tid = np.arange(BX)
        i = x0 + tid
        active = i < nx - cls
        ic = np.minimum(i, nx - 1)
        hl, hr = tid < r, tid >= BX - r
        hc = np.where(hl, i - r, np.minimum(i + r, nx - 1))
        assert ic.min() >= 0 and ic.max() < nx
        assert hc[hl | hr].min() >= 0 and hc[hl | hr].max() < nx
        # rows touched: prologue j0-R..j0+R-1, prefetch min(jn+R, ny-1) and min(jn, ny-1)
        # for jn up to j1 + PF - 1 (PF <= 3 rows in flight)
        rows = list(range(j0 - r, j0 + r)) + [min(jn + r, ny - 1) for jn in range(j0, j1 + 3)]
        rows += [min(jn, ny - 1) for jn in range(j0, j1 + 3)]
        assert min(rows) >= 0 and max(rows) < ny
        # LDS neighbours read by active threads map to real columns
        for t in tid[active]:
            for s in range(2 * r + 1):
                col = x0 - r + t + s
                assert 0 <= col < nx
        for j in range(j0, j1):
            count[j, i[active]] += 1
    assert (count == 1).all(), (nx, ny, order, np.argwhere(count != 1)[:5])
    return g


SHAPES = [(3, 3), (8, 8), (12, 12), (13, 14), (37, 29), (256, 40), (257, 257), (300, 45),
          (520, 521), (531, 270), (1025, 65)]


@pytest.mark.parametrize("order", [2, 4, 6])
@pytest.mark.parametrize("shape", SHAPES)
def test_geometry_covers_grid_once(order, shape):
    nx, ny = shape
    m = {2: 3, 4: 8, 6: 12}[order]
    if nx < m or ny < m:
        pytest.skip("below the minimum grid of this order")
    check(nx, ny, order)


@pytest.mark.parametrize("ty", [1, 7, 16, 64, 300])
def test_geometry_tile_heights(ty):
    check(300, 300, 6, ty)
    check(64, 70, 4, ty)


def test_geometry_large_grid_parameters():
    g = geometry(4097, 4097, 4)
    assert g["tiles_x"] == 16 and g["n_edge_blocks"] % NXCD == 0
    assert g["n_tiles"] >= 1024 and g["ty"] == 64
    assert geometry(2049, 2049, 4)["ty"] == 16
    assert geometry(1025, 1025, 4)["ty"] == 4
    check(1025, 1025, 4)
    # remap is a bijection on the big grid too
    nt = g["n_tiles"]
    assert sorted(xcd_tile(b, nt) for b in range(nt)) == list(range(nt))
    for nt in range(1, 40):
        assert sorted(xcd_tile(b, nt) for b in range(nt)) == list(range(nt))
