"""CPU model of the two-stage launches (sbp_capi.hip compute_pair_geometry /
default_pair_rows, sbp_kernels.hpp pair_tile_rw / pair2d_kernel and the frame
patches frame_patch / frame_patch_pair of the D = CLS + R deep frame).

Checks, for many grid shapes and all orders, that
* every global load of a pair tile (operand rows with prefetch, halo
  columns, the mid-row and output-row pointwise loads) and of the frame
  patches' table rows stays inside [0,nx) x [0,ny);
* the pair tiles' output points and the frame patches partition the grid;
* every mid point an output point differentiates is computed by the same
  tile, at an unclamped column, from input columns the tile staged, and is an
  interior point of the operator (so the interior stencil is the right row);
* every point the second stage of a frame patch reads (zero-padded taps
  included: a stale non-finite value times zero would still poison the sum)
  lies in the patch grown by RW = TAPS - R - 1, where the patch computed the
  first stage, and the grown patch fits the LDS array;
* the register windows and FIFOs hold the rows each phase reads
  (pair_tile_rw);
* the ensemble kernels' persistent loops take every work item of every XCD
  chunk exactly once.
"""
import numpy as np
import pytest

from test_kernel_geometry import CLS, R, TAPS, NXCD, dline_range, xcd_tile

def pair_nt(esz, order):
    """Threads per two-stage tile."""
    return 256


def default_pair_rows(nx, ny, order, esz=8):
    NT = pair_nt(esz, order)
    r, d = R[order], CLS[order] + R[order]
    wi, hi = nx - 2 * d, ny - 2 * d
    if wi <= 0 or hi <= 0:
        return 16
    tx = (wi + NT - 2 * r - 1) // (NT - 2 * r)
    if esz == 8:
        t = 136 if order == 6 else min(40, max(8, (hi + 50) // 100 // 2 * 2))
    else:
        t = 48 if order == 6 else 16
    while t > 8 and tx * ((hi + t - 1) // t) < 400:
        t //= 2
    return t


def pair_geometry(nx, ny, order, rows=None, esz=8):
    NT = pair_nt(esz, order)
    r, d = R[order], CLS[order] + R[order]
    rows = rows or default_pair_rows(nx, ny, order, esz)
    wi, hi = nx - 2 * d, ny - 2 * d
    bxo = NT - 2 * r
    tiles_x = (wi + bxo - 1) // bxo if wi > 0 else 0
    n_tiles = tiles_x * ((hi + rows - 1) // rows) if (wi > 0 and hi > 0) else 0
    n_patches = frame_patch_count(nx, ny, d)
    neb = (n_patches + NXCD - 1) // NXCD * NXCD
    return dict(d=d, rows=rows, tiles_x=tiles_x, n_tiles=n_tiles, n_patches=n_patches, n_edge_blocks=neb, nt=NT)


PATCH_W = 64  # kPatchW


def frame_patch_count(nx, ny, d):
    return 2 * (-(-nx // PATCH_W)) + 2 * (-(-(ny - 2 * d) // PATCH_W))


def frame_patch(p, nx, ny, d, rw, taps):
    """(i0, i1, j0, j1) of the patch and (ei0, ei1, ej0, ej1): grown by rw,
    sideways at least to the taps-point lines of the closure rows."""
    nb = -(-nx // PATCH_W)
    if p < 2 * nb:
        top = p >= nb
        k = p - nb if top else p
        i0, i1 = k * PATCH_W, min(k * PATCH_W + PATCH_W, nx)
        j0, j1 = (ny - d, ny) if top else (0, d)
        ei0, ei1 = max(min(i0 - rw, nx - taps), 0), min(max(i1 + rw, taps), nx)
        ej0, ej1 = (max(ny - d - rw, 0), ny) if top else (0, min(d + rw, ny))
    else:
        nl = -(-(ny - 2 * d) // PATCH_W)
        right = p - 2 * nb >= nl
        k = p - 2 * nb - (nl if right else 0)
        j0, j1 = d + k * PATCH_W, min(d + k * PATCH_W + PATCH_W, ny - d)
        i0, i1 = (nx - d, nx) if right else (0, d)
        ej0, ej1 = max(j0 - rw, 0), min(j1 + rw, ny)
        ei0, ei1 = (max(nx - d - rw, 0), nx) if right else (0, min(d + rw, nx))
    return (i0, i1, j0, j1), (ei0, ei1, ej0, ej1)


def simulate_windows(order, nrows):
    """Replays the register windows and FIFOs of pair_tile_rw for a tile of
    nrows output rows (rows numbered from j0 = 0): every value a phase uses
    must come from the row the formula needs."""
    r = R[order]
    WN = 2 * r + 1
    win = [None] * WN          # operand rows, oldest first
    wmid = [None] * WN         # mid rows
    fdx = [None] * (r + 1)     # dx of operand rows
    gdx = [None] * (r + 1)     # dx of mid rows
    hdx = [None] * (r + 1)     # pair 2: dx of U at the mid rows
    fmat = [None] * (r + 1)    # materials of mid rows
    fU = [None] * WN           # pair 2: U of the mid rows, oldest first
    nk = nrows + 4 * r
    for k in range(nk):
        row = -2 * r + k
        win = win[1:] + [row]                      # phase 0
        fdx = fdx[1:] + [row]                      # phase 1: dx of row r
        if k >= 2 * r:
            m = row - r
            assert win == list(range(m - r, m + r + 1)) and fdx[0] == m
            assert win[r] == m                     # U(m) of pair 1
            wmid = wmid[1:] + [m]
            fmat = fmat[1:] + [m]
            fU = fU[1:] + [m]
            gdx = gdx[1:] + [m]                    # phase 2: dx of mid row m
            hdx = hdx[1:] + [m]                    #          and of U at row m
        if k >= 4 * r:
            o = -4 * r + k
            assert wmid == list(range(o - r, o + r + 1)), (order, k, wmid)
            assert gdx[0] == o and fmat[0] == o and wmid[r] == o and win[0] == o
            # pair 2 recomputes K1 at row o: U over rows o-R..o+R, dx of U at row o
            assert fU == list(range(o - r, o + r + 1)) and fU[r] == o and hdx[0] == o


def pair_pf(esz):
    """Rows of loads in flight of a two-stage tile, the deepest of the two
    launches (sbp_kernels.hpp pair_pf: FP64 pair 1 and every FP32 tile run
    three rows ahead, FP64 pair 2 two; the ensemble's FP64 tiles two)."""
    return 3


def check(nx, ny, order, rows=None, pf=None, esz=8):
    pf = pair_pf(esz) if pf is None else pf
    g = pair_geometry(nx, ny, order, rows, esz)
    NT = g["nt"]
    r, d, cls = R[order], g["d"], CLS[order]
    LW = NT + 2 * r
    count = np.zeros((ny, nx), dtype=np.int32)
    # frame patches: first stage on the grown patch, second on the patch
    rw = TAPS[order] - r - 1
    fp_n = (d + rw) * (PATCH_W + 2 * rw)
    for p in range(g["n_patches"]):
        (i0, i1, j0, j1), (ei0, ei1, ej0, ej1) = frame_patch(p, nx, ny, d, rw, TAPS[order])
        assert 0 <= ei0 <= i0 < i1 <= ei1 <= nx and 0 <= ej0 <= j0 < j1 <= ej1 <= ny
        assert (ei1 - ei0) * (ej1 - ej0) <= fp_n
        for j in range(ej0, ej1):       # first stage: every read inside the grid
            for i in range(ei0, ei1):
                lo, hi = dline_range(nx, i, order)
                assert 0 <= lo and hi < nx
                lo, hi = dline_range(ny, j, order)
                assert 0 <= lo and hi < ny
        for j in range(j0, j1):         # second stage: every read inside the grown patch
            for i in range(i0, i1):
                count[j, i] += 1
                lo, hi = dline_range(nx, i, order)
                assert ei0 <= lo and hi < ei1, (p, i, j, lo, hi, ei0, ei1)
                lo, hi = dline_range(ny, j, order)
                assert ej0 <= lo and hi < ej1, (p, i, j, lo, hi, ej0, ej1)
    assert g["n_edge_blocks"] >= g["n_patches"] and g["n_edge_blocks"] % NXCD == 0
    tiles = [xcd_tile(bi, g["n_tiles"]) for bi in range(g["n_tiles"])]
    assert sorted(tiles) == list(range(g["n_tiles"]))
    t = np.arange(NT)
    for tile in range(g["n_tiles"]):
        tx, tyb = tile % g["tiles_x"], tile // g["tiles_x"]
        x0 = d + tx * (NT - 2 * r)
        mc = x0 - r + t
        mcc = np.minimum(mc, nx - 1)
        hl, hr = t < r, t >= NT - r
        hc = np.where(hl, mc - r, np.minimum(mc + r, nx - 1))
        assert mcc.min() >= 0 and mcc.max() < nx
        assert hc[hl | hr].min() >= 0 and hc[hl | hr].max() < nx
        hslot = np.where(hl, t, t + 2 * r)
        assert hslot[hl | hr].min() >= 0 and hslot[hl | hr].max() < LW
        out = (t >= r) & (t < NT - r) & (mc < nx - d)
        j0 = d + tyb * g["rows"]
        j1 = min(j0 + g["rows"], ny - d)
        assert j1 > j0
        nk = (j1 - j0) + 4 * r
        # the row loop runs whole groups of pf + 1 iterations, each loading
        # row k + pf ahead: the last load is for k = groups * (pf + 1) - 1 + pf
        kmax = -(-nk // (pf + 1)) * (pf + 1) + pf
        for k in range(kmax):
            row = min(j0 - 2 * r + k, ny - 1)
            assert 0 <= row < ny and 0 <= max(row - r, 0) and 0 <= max(row - 2 * r, 0)
        # staged input columns of this tile (LDS position p <-> column x0 - 2R + p)
        staged = set(range(x0 - 2 * r, x0 - 2 * r + LW))
        for tt in t[out]:
            c = mc[tt]
            assert c >= d
            for s in range(2 * r + 1):
                mcol = c - r + s
                # mid column computed by thread tt - r + s at an unclamped column
                assert mc[tt - r + s] == mcol and mcol < nx
                assert cls <= mcol < nx - cls
                for s2 in range(2 * r + 1):
                    assert mcol - r + s2 in staged and 0 <= mcol - r + s2 < nx
        cols = mc[out]
        for o in range(j0, j1):
            for s in range(2 * r + 1):
                assert cls <= o - r + s < ny - cls
            count[o, cols] += 1
    assert (count == 1).all(), (nx, ny, order, np.argwhere(count != 1)[:5])
    return g


SHAPES = [(13, 14), (19, 19), (37, 29), (256, 40), (257, 257), (270, 33), (300, 45), (520, 521), (531, 270)]


@pytest.mark.parametrize("esz", [8, 4])
@pytest.mark.parametrize("order", [2, 4, 6])
@pytest.mark.parametrize("shape", SHAPES)
def test_pair_geometry_covers_grid_once(order, shape, esz):
    nx, ny = shape
    d = CLS[order] + R[order]
    if nx <= 2 * d or ny <= 2 * d:
        pytest.skip("no interior for the two-stage tiles")
    check(nx, ny, order, esz=esz)


@pytest.mark.parametrize("rows", [1, 3, 8, 64, 200])
def test_pair_geometry_tile_heights(rows):
    check(300, 300, 6, rows)
    check(300, 300, 6, rows, esz=4)
    check(270, 70, 4, rows)
    check(40, 300, 2, rows)


@pytest.mark.parametrize("order", [2, 4, 6])
@pytest.mark.parametrize("nrows", [1, 2, 5, 8, 64])
def test_pair_register_windows(order, nrows):
    simulate_windows(order, nrows)


def test_pair_geometry_large_grid_parameters():
    g = pair_geometry(4097, 4097, 4)
    assert g["d"] == 6 and g["tiles_x"] == 17 and g["rows"] == 40 and g["n_tiles"] == 17 * 103
    assert pair_geometry(4097, 4097, 6)["rows"] == 136 and pair_geometry(4097, 4097, 6)["nt"] == 256
    assert pair_geometry(4097, 4097, 6, esz=4)["rows"] == 48
    assert pair_geometry(4097, 4097, 4, esz=4)["rows"] == 16
    assert pair_geometry(1025, 1025, 4)["rows"] == 10
    assert pair_geometry(1025, 1025, 4, esz=4)["rows"] == 8
    check(4097, 300, 4)
    check(1025, 1025, 4)


def batch_traversal(counts, grid, sub):
    """Replays the persistent loops of the ensemble kernels (sbp_batch.hpp
    batch_stage_kernel / batch_pair_kernel: sub = 1; sub > 1 splits every
    item over sub blocks of grid * sub): block b works on XCD b % 8's chunk,
    from item (b / 8) / sub on, striding by the XCD's block count.  Returns
    the (xcd, item, part) triples visited."""
    assert grid % NXCD == 0
    nb = grid * sub
    seen = []
    for b in range(nb):
        xcd = b % NXCD
        n = counts[xcd] * sub
        v = b // NXCD
        while v < n:
            seen.append((xcd, v // sub, v % sub))
            v += nb // NXCD
    return seen


@pytest.mark.parametrize("sub", [1, 4])
@pytest.mark.parametrize("grid", [8, 256, 1024, 4096])
def test_batch_work_traversal_visits_every_item_once(grid, sub):
    """Every work item of every XCD chunk (and every part of an item split
    over several blocks) is taken by exactly one block, for chunks shorter
    and longer than the XCD's share of the grid, and empty ones."""
    rng = np.random.default_rng(grid + sub)
    for counts in ([0] * 8, [1, 0, 3, 0, 0, 7, 0, 2], list(rng.integers(0, 2000, 8))):
        seen = batch_traversal(counts, grid, sub)
        want = [(x, i, p) for x in range(NXCD) for i in range(counts[x]) for p in range(sub)]
        assert sorted(seen) == sorted(want)
