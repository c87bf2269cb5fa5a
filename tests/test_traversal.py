"""Host-side guard of the sweep's index math: the wave decode of gpe_device.h (decode for one
row per wave, decode_rows for x pairs over row groups) emulated with the constants of the
library's own traversal plan (gpe_plan_traversal, computed by the same host code that sizes the
launches).  For every wave of a sweep the unknown, plane, rows and x positions it touches lie
inside the box or are masked, the fast divisions agree with exact integer division, the waves
cover every (z, y, v, x) of the box exactly once, and both workgroup orders of consecutive
sweeps (the XCD remap and its reversal) are permutations of the launch.  The level layout's
strides place every element of the box at its own address inside the level, and x pairs spread
over a tile of the slab layout load from 128-byte boundaries."""
import ctypes

import numpy as np
import pytest

from spacetime_amd import gpe

U32 = np.uint64(0xFFFFFFFF)


_Plan = gpe._Plan  # the ctypes mirror of gpe_traversal_plan


def _plan(cfg, ty, remap, xpl):
    lib = gpe.load_library()
    p = _Plan()
    rc = lib.gpe_plan_traversal(cfg.to_c(), ty, remap, xpl, ctypes.byref(p))
    return rc, p


def _fdiv(n, div):
    """The device's n / d as multiply-high, add and shift (gpe_device.h fast_div) for n < 2^31,
    in 32-bit unsigned arithmetic."""
    d, m, l = (int(x) for x in div)
    n = n.astype(np.uint64)
    assert int(n.max(initial=0)) < 2 ** 31
    t = (n * np.uint64(m)) >> np.uint64(32)
    s = t + n
    assert int(s.max(initial=0)) < 2 ** 32  # no 32-bit wrap
    return s >> np.uint64(l)


def _magic(d):
    """The host's constants for d (gpe_host.cpp make_fast_div), restated."""
    l = 0
    while (1 << l) < d:
        l += 1
    m = 1 if d <= 1 else ((2 ** 32) * ((1 << l) - d)) // d + 1
    return d, m, l


@pytest.mark.parametrize("d", [1, 2, 3, 5, 7, 12, 41, 46, 64, 80, 97, 136, 1000, 4099, 65536, 1000003, 2 ** 30 + 3,
                               2 ** 31 - 1])
def test_fast_division_is_exact_below_2_31(d):
    rng = np.random.default_rng(d)
    k = rng.integers(0, (2 ** 31 - 1) // d + 1, 2000)  # multiples of d and their neighbours
    n = np.concatenate([rng.integers(0, 2 ** 31, 20000), np.arange(0, min(4 * d + 5, 5000)),
                        k * d, k * d + d - 1, np.maximum(k * d, 1) - 1,
                        [2 ** 31 - 1, 2 ** 31 - 2, (2 ** 31 - 1) // d * d, (2 ** 31 - 1) // d * d - 1]])
    n = n[(n >= 0) & (n < 2 ** 31)].astype(np.uint64)
    np.testing.assert_array_equal(_fdiv(n, _magic(d)).astype(np.int64), n.astype(np.int64) // d)


def _logical_block(b, nb):
    q, r, k = nb >> 3, nb & 7, b & 7
    return k * q + np.minimum(k, r) + (b >> 3)


def _reverse_in_xcd(b, nb):
    q, r, k = nb >> 3, nb & 7, b & 7
    length = q + (k < r)
    return k + 8 * (length - 1 - (b >> 3))


def _decode_all(p):
    """Every (workgroup, wave, lane) of a sweep: arrays of v, z, y, x, live flag."""
    nb = p.blocks
    b = np.arange(nb, dtype=np.int64)
    lb = _logical_block(b, nb) if p.xcd_remap & 1 else b
    w = (lb[:, None] * 4 + np.arange(4)[None, :]).reshape(-1).astype(np.uint64)
    assert int(w.max()) < 2 ** 31
    wi = w.astype(np.int64)
    if p.xcd_remap & 2:  # the unknown inside the x segment
        q0 = _fdiv(w, p.div_V)
        v = (w - q0 * np.uint64(p.V)).astype(np.int64)
        q1 = _fdiv(q0, p.div_nseg)
        xseg = (q0 - q1 * np.uint64(p.nseg)).astype(np.int64)
        assert np.array_equal(q0.astype(np.int64), wi // p.V)
    else:
        q0 = _fdiv(w, p.div_nseg)
        xseg = (w - q0 * np.uint64(p.nseg)).astype(np.int64)
        q1 = _fdiv(q0, p.div_V)
        v = (q0 - q1 * np.uint64(p.V)).astype(np.int64)
        assert np.array_equal(q0.astype(np.int64), wi // p.nseg)
    q2 = _fdiv(q1, p.div_rows)
    q3 = _fdiv(q2, p.div_nz)
    z = (q2 - q3 * np.uint64(p.nz)).astype(np.int64)
    # exact divisions for comparison
    assert np.array_equal(q1.astype(np.int64), wi // (p.nseg * p.V))
    assert np.array_equal(q2.astype(np.int64), (wi // (p.nseg * p.V)) // p.row_groups)
    lane = np.arange(64, dtype=np.int64)
    if p.rows == 1:
        yin = (q1 - q2 * np.uint64(p.tile_rows)).astype(np.int64)
        y0 = q3.astype(np.int64) * p.tile_rows + yin
        valid = (wi < p.waves) & (y0 < p.ny)
        x = ((xseg[:, None] * 64 + lane[None, :]) * p.x_per_lane)
        r = np.zeros_like(x)
    else:
        yg = (q1 - q2 * np.uint64(p.row_groups)).astype(np.int64)
        y0 = q3.astype(np.int64) * p.tile_rows + yg * p.rows
        valid = (wi < p.waves) & (y0 < p.ny)
        slot = (xseg[:, None] * 64 + lane[None, :]).astype(np.uint64)
        rr = _fdiv(slot.reshape(-1), p.div_pairs).reshape(slot.shape)
        assert np.array_equal(rr.astype(np.int64), slot.astype(np.int64) // p.pairs)
        r = rr.astype(np.int64)
        x = 2 * (slot.astype(np.int64) - r * p.pairs)
    return v, z, y0, x, r, valid


CASES = [  # lattice, ncomp, precision, dims
    ("bcc", 3, gpe.FP32, (64, 20, 7)), ("bcc", 3, gpe.FP32, (130, 9, 5)), ("bcc", 1, gpe.FP32, (192, 6, 3)),
    ("bcc", 5, gpe.FP64, (70, 11, 4)), ("cube", 5, gpe.FP32, (33, 13, 6)), ("fcc", 3, gpe.FP32, (7, 17, 5)),
    ("a15", 1, gpe.FP32, (57, 10, 3)), ("c15", 5, gpe.FP32, (29, 6, 3)), ("z", 3, gpe.FP32, (66, 9, 4)),
    ("fcc", 1, gpe.FP32, (160, 5, 3)), ("cube", 3, gpe.FP32, (293, 4, 2)),
    # row groups of 3, 6 and 7 rows (A15's 41 pairs, Z's 34, BCC's 35, C15's 10)
    ("a15", 5, gpe.FP32, (82, 13, 2)), ("z", 5, gpe.FP32, (68, 15, 2)), ("bcc", 3, gpe.FP32, (70, 9, 2)),
    ("c15", 1, gpe.FP32, (20, 7, 2)), ("a15", 5, gpe.FP32, (82, 7, 2)), ("z", 3, gpe.FP32, (68, 16, 2)),
    ("bcc", 1, gpe.FP32, (70, 9, 2)), ("c15", 3, gpe.FP32, (20, 7, 2)),
    # one component over long lines: several waves per line
    ("cube", 1, gpe.FP32, (293, 4, 2)), ("bcc", 1, gpe.FP32, (600, 3, 2)), ("z", 1, gpe.FP32, (131, 5, 2)),
    # planes of more than 64 rows (whole-plane tiles of 82 and 130 rows)
    ("a15", 1, gpe.FP32, (82, 82, 1)), ("cube", 3, gpe.FP32, (58, 130, 1)),
]


@pytest.mark.parametrize("lat,ncomp,precision,dims", CASES)
def test_decode_stays_in_range_and_covers_the_box(lat, ncomp, precision, dims):
    cfg = gpe.Config(lattice=lat, ncomp=ncomp, precision=precision, dims=dims)
    nx, ny, nz = dims
    xpls = (1, 2) if precision == gpe.FP32 else (1,)
    checked = 0
    for xpl in xpls:
        for ty in (1, 2, 3, 4, 5, 6, 7, 8, 12, 14, 16, 32, 64, 1024):  # 1024: the whole plane
            for remap in (0, 1, 2, 3):
                rc, p = _plan(cfg, ty, remap, xpl)
                assert rc == 0
                assert p.blocks * 4 >= p.waves and p.tile_rows <= ny and p.tile_rows % p.rows == 0
                assert p.lx >= nx and p.lx % 2 == 0 or p.lx == nx
                v, z, y0, x, r, valid = _decode_all(p)
                V = p.V
                # wave-uniform coordinates of valid waves are inside the box
                assert np.all((v[valid] >= 0) & (v[valid] < V))
                assert np.all((z[valid] >= 0) & (z[valid] < nz))
                assert np.all((y0[valid] >= 0) & (y0[valid] < ny))
                count = np.zeros((nz, ny, V, nx), dtype=np.int32)
                if p.rows == 1:
                    live = valid[:, None] & (x < nx)
                    zz, yy, vv = (np.broadcast_to(a[:, None], x.shape) for a in (z, y0, v))
                    for e in range(xpl):
                        ok = live & (x + e < nx)
                        np.add.at(count, (zz[ok], yy[ok], vv[ok], (x + e)[ok]), 1)
                else:
                    nrow = np.minimum(ny - y0, p.rows)
                    live = valid[:, None] & (r < nrow[:, None]) & (x < nx)
                    # the lane's row (clamped into the box for masked lanes) is inside the box
                    rr = np.where(r < nrow[:, None], r, nrow[:, None] - 1)
                    assert np.all(y0[valid, None] + rr[valid] < ny)
                    zz, vv = (np.broadcast_to(a[:, None], x.shape) for a in (z, v))
                    yy = y0[:, None] + r
                    for e in range(2):
                        ok = live & (x + e < nx)
                        np.add.at(count, (zz[ok], yy[ok], vv[ok], (x + e)[ok]), 1)
                    # an x pair that starts inside the line ends inside its padding at most
                    assert np.all(x[live] + 1 < p.lx)
                assert count.min() == 1 and count.max() == 1, (ty, remap, xpl)
                if p.rows > 1 and p.slabs and p.rows == p.tile_rows and (p.rows * p.lx % 16 == 0 or p.ntile == 1):
                    # the tile is one run of pairs: every wave's first lane starts a 128-byte line
                    first = (z * p.zs + v * p.vs + y0 * p.ys)[valid] + (r[valid, 0] * p.ys + x[valid, 0])
                    assert np.all(first % 16 == 0), (ty, remap)
                checked += 1
    assert checked >= 32


@pytest.mark.parametrize("lat,ncomp,precision,dims", CASES + [("c15", 5, gpe.FP32, (57, 9, 3))])
def test_layout_addresses_every_element_once_inside_the_level(lat, ncomp, precision, dims):
    """z zs + y ys + v vs + c cs + x over the box: distinct, inside the level (nz zs elements),
    offsets within a plane and its neighbours in int range; slabs exactly for FP32 lines that
    are not whole 128-byte lines, with contiguous rows."""
    cfg = gpe.Config(lattice=lat, ncomp=ncomp, precision=precision, dims=dims)
    nx, ny, nz = dims
    rc, p = _plan(cfg, 1, 1, 1)
    assert rc == 0
    align = 8 if precision == gpe.FP64 else 16
    assert p.slabs == (precision == gpe.FP32 and (nx + nx % 2) % align != 0)
    if p.slabs or precision == gpe.FP32:
        assert p.lx == nx + nx % 2  # FP32: no padding beyond an even length
    if p.slabs:
        assert p.ys == p.lx and p.cs % align == 0 and p.cs >= ny * p.lx
    else:
        assert p.cs == p.lx and p.ys == p.V * ncomp * p.lx
    z, y, v, c, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(p.V), np.arange(ncomp), np.arange(nx),
                                indexing="ij")
    addr = z * p.zs + y * p.ys + v * p.vs + c * p.cs + x
    assert addr.min() == 0 and addr.max() < nz * p.zs
    assert np.unique(addr).size == addr.size
    assert 4 * p.zs < 2 ** 31


@pytest.mark.parametrize("lat,ncomp,dims,ty,rows", [
    ("a15", 5, (82, 13, 2), 12, 3),  # 41 pairs: 3 rows fill 96% of two waves
    ("a15", 5, (82, 13, 2), 4, 1),   # 4 rows fill 85%: below the large-cell floor of 90%
    ("a15", 5, (82, 13, 2), 8, 8),   # slabs: 8 rows of 82 values start on 128 bytes, the tile is the group
    ("a15", 5, (82, 13, 2), 16, 13), # one tile
    ("z", 5, (68, 15, 2), 14, 7),    # 34 pairs: 7 rows fill 93% of four waves
    ("z", 5, (68, 15, 2), 8, 8),
    ("z", 5, (68, 15, 2), 6, 1),
    ("bcc", 3, (70, 9, 2), 7, 7),    # 35 pairs: 7 rows fill 96%
    ("bcc", 3, (70, 9, 2), 6, 3),    # small cells take the best group of 1-8 rows (82%)
    ("bcc", 3, (70, 9, 2), 8, 8),
    ("c15", 1, (20, 7, 2), 6, 6),    # 10 pairs: 6 rows fill 94% of one wave
    ("c15", 5, (57, 9, 2), 8, 8),
    ("c15", 5, (57, 9, 2), 2, 2),    # lines of <= 64 values: 2 rows
    ("cube", 5, (90, 40, 2), 24, 24),
    ("cube", 5, (90, 40, 2), 12, 4),
    ("cube", 5, (293, 40, 2), 24, 24), # slabs of 294-value rows, not x-lines padded to 304
    ("cube", 5, (293, 40, 2), 12, 3),
    ("bcc", 3, (128, 40, 2), 16, 1), # whole 128-byte lines keep the line layout: one row per wave
])
def test_row_group_choice(lat, ncomp, dims, ty, rows):
    """The row group a tile height gets, and that the cases above cover groups of 3, 6 and 7 and
    tiles as groups."""
    cfg = gpe.Config(lattice=lat, ncomp=ncomp, precision=gpe.FP32, dims=dims)
    rc, p = _plan(cfg, ty, 1, 2)
    assert rc == 0 and p.rows == rows and p.tile_rows == min(ty, dims[1])


@pytest.mark.parametrize("nb", [1, 7, 8, 9, 63, 64, 1000, 4097])
def test_workgroup_orders_are_permutations(nb):
    b = np.arange(nb)
    assert np.array_equal(np.sort(_logical_block(b, nb)), b)
    rev = _reverse_in_xcd(b, nb)
    assert np.array_equal(np.sort(rev), b)
    assert np.all(rev % 8 == b % 8)  # a reversed workgroup stays on its XCD
    # within each XCD's range the remapped order runs backwards
    lb = _logical_block(rev, nb)
    for k in range(min(8, nb)):
        sel = b % 8 == k
        assert np.all(np.diff(lb[sel]) < 0) or sel.sum() <= 1


def test_plan_keeps_wave_indices_below_2_31():
    """The decode's divisions are exact for wave indices below 2^31: a launch of 2^29 workgroups or
    more (four waves each) is refused; one just below is planned."""
    ok = gpe.Config(lattice="cube", ncomp=1, precision=gpe.FP64, dims=(64, 32768, 65535))
    rc, p = _plan(ok, 1, 1, 1)
    assert rc == 0 and p.waves < 2 ** 31 and p.blocks < 2 ** 29
    big = gpe.Config(lattice="cube", ncomp=1, precision=gpe.FP64, dims=(64, 32768, 65537))
    assert _plan(big, 1, 1, 1)[0] == -22


def test_plan_rejects_bad_arguments():
    cfg = gpe.Config(lattice="bcc", ncomp=3, precision=gpe.FP64, dims=(8, 8, 8))
    assert _plan(cfg, 8, 1, 2)[0] == -22  # FP64 has one x per lane
    assert _plan(gpe.Config(lattice="bcc", ncomp=1, dims=(8, 8, 8)), 8, 1, 4)[0] == -22
    assert _plan(cfg, 0, 1, 1)[0] == -22 and _plan(cfg, 1025, 1, 1)[0] == -22  # tile rows 1..1024
    rc, p = _plan(cfg, 1024, 1, 1)
    assert rc == 0 and p.tile_rows == 8 and p.ntile == 1  # clamped to the plane
    assert _plan(gpe.Config(lattice="bcc", dims=(0, 8, 8)), 8, 1, 1)[0] == -22
    assert _plan(cfg, 8, 4, 1)[0] == -22 and _plan(cfg, 8, -1, 1)[0] == -22  # order flags 0..3
    rc, p = _plan(cfg, 8, 3, 1)
    assert rc == 0 and p.xcd_remap == 3


def test_level_layout_wrapper():
    """gpe.level_layout: slabs for FP32 lines whose even length is not a whole number of 128-byte
    lines, x-lines otherwise and for FP64; strides as the plan gives them."""
    lay = gpe.level_layout(gpe.Config(lattice="a15", ncomp=5, precision=gpe.FP32, dims=(82, 82, 4)))
    assert lay["layout"] == "slabs" and lay["lx"] == 82 and lay["ys"] == 82 and lay["cs"] % 16 == 0
    assert lay["cs"] >= 82 * 82 and lay["vs"] == 5 * lay["cs"] and lay["zs"] == 46 * lay["vs"]  # A15: 46 unknowns a block
    assert gpe.level_layout(gpe.Config(lattice="bcc", ncomp=3, precision=gpe.FP32, dims=(128, 8, 4)))["layout"] == "lines"
    lay = gpe.level_layout(gpe.Config(lattice="a15", ncomp=5, precision=gpe.FP64, dims=(82, 82, 4)))
    assert lay["layout"] == "lines" and lay["cs"] == lay["lx"] and lay["ys"] == 46 * 5 * lay["lx"]
