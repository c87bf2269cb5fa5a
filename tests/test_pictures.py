"""BMP density pictures written by libgpe's host code from host states (no device needed).

Each picture is compared byte for byte with a pixel-by-pixel restatement of the reference's
drawing loops (Spin1GpeBCC/utils.h:151-580, Spin2GpeBCC/utils.h:131-796,
GPUGrossPitaevskii_BCC.cu:700-723) and Picture::save (Output/Picture.cpp:83-138), written here
in plain Python over the padded block array the reference draws from; what those loops leave
open is pinned by properties: where a single occupied block lands (the mirrored panels), that
an eigenstate of the basis axis lights exactly one component panel, the channel encoding of a
known value, and the spin-2 YZ panels' use of the XZ scale."""
import struct

import numpy as np
import pytest

from spacetime_amd import gpe, pictures

RNG = np.random.default_rng(5)
S2 = np.sqrt(2.0)


def _cfg(lat, ncomp, dims, cell=0.7):
    return gpe.Config.centred_box(lat, dims, cell, ncomp=ncomp)


def _random(cfg):
    return RNG.normal(size=cfg.shape) + 1j * RNG.normal(size=cfg.shape)


# ------------------------------------------------------------ restatement of the reference --

def _padded(cfg, psi):
    """BlockPsis array of the reference: [z][y][x] padded blocks of V unknowns x C components."""
    nx, ny, nz = cfg.dims
    C, V = psi.shape[0], psi.shape[-1]
    blocks = np.zeros((nz + 2, ny + 2, nx + 2, V, C), complex)
    blocks[1:-1, 1:-1, 1:-1] = np.moveaxis(psi, 0, -1)
    return blocks


def _basis(s, basis):
    """The reference's BASIS rotations, term by term (Spin1GpeBCC/utils.h:176-190,
    Spin2GpeBCC/utils.h:312-343)."""
    if basis == gpe.BASIS_Z:
        return s
    if len(s) == 3:
        s1, s0, sm = s
        if basis == gpe.BASIS_X:
            return [0.5 * (s1 + sm) - s0 / S2, (s1 - sm) / S2, 0.5 * (sm + s1) + s0 / S2]
        mi = complex(0, -1)
        return [0.5 * (s1 - sm) + s0 * mi / S2, (s1 + sm) * mi / S2, 0.5 * (sm - s1) + s0 * mi / S2]
    s2, s1, s0, sm1, sm2 = s
    c = np.sqrt(6) * 0.25
    if basis == gpe.BASIS_X:
        return [0.25 * s2 + 0.5 * s1 + c * s0 + 0.5 * sm1 + 0.25 * sm2,
                -0.5 * s2 - 0.5 * s1 + 0.5 * sm1 + 0.5 * sm2,
                c * s2 - 0.5 * s0 + c * sm2,
                -0.5 * s2 + 0.5 * s1 - 0.5 * sm1 + 0.5 * sm2,
                0.25 * s2 - 0.5 * s1 + c * s0 - 0.5 * sm1 + 0.25 * sm2]
    im = 1j
    return [0.25 * s2 - im * 0.5 * s1 - c * s0 + im * 0.5 * sm1 + 0.25 * sm2,
            -im * 0.5 * s2 - 0.5 * s1 - 0.5 * sm1 + im * 0.5 * sm2,
            -c * s2 - 0.5 * s0 - c * sm2,
            im * 0.5 * s2 - 0.5 * s1 - 0.5 * sm1 - im * 0.5 * sm2,
            0.25 * s2 + im * 0.5 * s1 - c * s0 - im * 0.5 * sm1 + 0.25 * sm2]


class _Picture:
    def __init__(self, w, h):
        self.w, self.h = w, h
        self.px = [[(0.0, 0.0, 0.0)] * w for _ in range(h)]

    def set(self, x, y, col):
        if x < self.w:  # the reference checks x only; these tests stay inside in y
            self.px[y][x] = col

    def bytes(self):
        def ch(v):
            c = 256.0 * v
            return 0 if not c >= 1 else (255 if c > 255 else int(c))
        data = 4 * self.w * self.h
        out = b"BM" + struct.pack("<III", 54 + data, 0, 54)
        out += struct.pack("<IiiHHIIIIII", 40, self.w, self.h, 1, 24, 0, data, 0, 0, 0, 0)
        pad = 3 - ((3 * self.w + 3) % 4)
        for row in self.px:
            for r, g, b in row:
                out += bytes([ch(b), ch(g), ch(r)])
            out += b"\0" * pad
        return out


def _sums(blocks, coord, n_s, basis, ri):
    """Column sums of one pixel: the third axis visited pixel by pixel (coord(p) -> block of
    pixel p along it), each block's unknowns in turn."""
    C = blocks.shape[-1]
    acc = [complex(0, 0) if ri else 0.0 for _ in range(C)]
    for p in range(n_s):
        z, y, x = coord(p)
        for v in range(blocks.shape[3]):
            s = _basis(list(blocks[z, y, x, v]), basis)
            for c in range(C):
                if ri:
                    acc[c] += complex(s[c].real * s[c].real, s[c].imag * s[c].imag)
                else:
                    acc[c] += s[c].real * s[c].real + s[c].imag * s[c].imag
    return acc


def _max_fold(vals):
    m = vals[-1]
    for v in reversed(vals[:-1]):
        m = max(v, m) if not v < m else m
    return m


def _reference_picture(cfg, psi, style, basis=gpe.BASIS_Z):
    blocks = _padded(cfg, psi)
    nx, ny, nz = cfg.dims
    W, H, D = 2 * (nx + 2), 2 * (ny + 2), 2 * (nz + 2)
    C = psi.shape[0]
    if style == pictures.PICTURE_SLICE:
        pic = _Picture(W, H)
        k = nz // 2 + 1
        for j in range(H):
            for i in range(W):
                s = blocks[k, j // 2, i // 2, 0, 0]
                pic.set(i, j, (20.0 * (s.real * s.real), 20.0 * (s.imag * s.imag), 0.0))
        return pic.bytes()
    ri = style == pictures.PICTURE_RI
    xz = lambda i, k: (lambda j: (k // 2, j // 2, i // 2))
    xy = lambda i, j: (lambda k: (k // 2, j // 2, i // 2))
    yz = lambda j, k: (lambda i: (k // 2, j // 2, i // 2))

    def plane_max(coords):
        mx = 0.0
        for cf, n in coords:
            mx = max(mx, _max_fold(_sums(blocks, cf, n, basis, False)))
        return mx

    def col(v, I):
        if ri:
            return (I * v.real, I * v.imag, 0.0)
        g = I * v
        return (g, g, g)

    if C == 3:
        pic = _Picture(3 * W, 2 * H)
        num = 2.0 if ri else 1.0
        I = num / plane_max([(xz(i, k), H) for k in range(D) for i in range(W)])
        for k in range(D):
            for i in range(W):
                s = _sums(blocks, xz(i, k), H, basis, ri)
                u = i if ri else W - i - 1
                for c in range(3):
                    pic.set(c * W + u, k, col(s[c], I))
        I = num / plane_max([(xy(i, j), D) for j in range(H) for i in range(W)])
        for j in range(H):
            for i in range(W):
                s = _sums(blocks, xy(i, j), D, basis, ri)
                u, w = (i, j) if ri else (W - i - 1, H - j - 1)
                for c in range(3):
                    pic.set(c * W + u, H + w, col(s[c], I))
        for x in range(3 * W):
            pic.set(x, H, (0.5, 0.5, 0.5))
        for y in range(2 * H):
            pic.set(W, y, (0.5, 0.5, 0.5))
            pic.set(2 * W, y, (0.5, 0.5, 0.5))
        return pic.bytes()
    if ri:
        pic = _Picture(5 * W, 2 * H)
        for k in range(D):
            for i in range(W):
                s = _sums(blocks, xz(i, k), H, basis, True)
                for c in range(5):
                    pic.set(c * W + i, k, col(s[c], 1.0))
        for j in range(H):
            for i in range(W):
                s = _sums(blocks, xy(i, j), D, basis, True)
                for c in range(5):
                    pic.set(c * W + i, H + j, col(s[c], 1.0))
        for x in range(5 * W):
            pic.set(x, H, (0.5, 0.5, 0.5))
        for y in range(2 * H):
            for k in range(1, 5):
                pic.set(k * W, y, (0.5, 0.5, 0.5))
        return pic.bytes()
    pic = _Picture(5 * W, 3 * H)
    I = 1.0 / plane_max([(xz(i, k), H) for k in range(D) for i in range(W)])
    for k in range(D):
        for i in range(W):
            s = _sums(blocks, xz(i, k), H, basis, False)
            for c in range(5):
                pic.set(c * W + i, k, col(s[c], I))
    for k in range(D):
        for j in range(H):
            s = _sums(blocks, yz(j, k), W, basis, False)
            for c in range(5):
                pic.set(c * W + j, H + k, col(s[c], I))
    I = 1.0 / plane_max([(xy(i, j), D) for j in range(H) for i in range(W)])
    for j in range(H):
        for i in range(W):
            s = _sums(blocks, xy(i, j), D, basis, False)
            for c in range(5):
                pic.set(c * W + i, 2 * H + j, col(s[c], I))
    for x in range(5 * W):
        pic.set(x, H, (0.5, 0.5, 0.5))
        pic.set(x, 2 * H, (0.5, 0.5, 0.5))
    for y in range(3 * H):
        for k in range(1, 5):
            pic.set(k * W, y, (0.5, 0.5, 0.5))
    return pic.bytes()


# --------------------------------------------------------------------------------- tests --

CASES = [
    ("bcc", 3, (2, 3, 2), pictures.PICTURE_DENSITY, gpe.BASIS_Z),
    ("bcc", 3, (3, 2, 2), pictures.PICTURE_DENSITY, gpe.BASIS_X),
    ("bcc", 3, (2, 2, 3), pictures.PICTURE_DENSITY, gpe.BASIS_Y),
    ("bcc", 3, (2, 3, 2), pictures.PICTURE_RI, gpe.BASIS_Z),
    ("bcc", 3, (2, 2, 2), pictures.PICTURE_RI, gpe.BASIS_Y),
    ("cube", 5, (3, 2, 2), pictures.PICTURE_DENSITY, gpe.BASIS_Z),
    ("cube", 5, (2, 2, 3), pictures.PICTURE_DENSITY, gpe.BASIS_X),
    ("bcc", 5, (2, 2, 2), pictures.PICTURE_DENSITY, gpe.BASIS_Y),
    ("cube", 5, (2, 3, 2), pictures.PICTURE_RI, gpe.BASIS_Z),
    ("bcc", 1, (3, 4, 3), pictures.PICTURE_SLICE, gpe.BASIS_Z),
    ("a15", 1, (2, 2, 2), pictures.PICTURE_SLICE, gpe.BASIS_Z),
]


@pytest.mark.parametrize("lat,ncomp,dims,style,basis", CASES)
def test_picture_matches_the_reference_loops_byte_for_byte(tmp_path, lat, ncomp, dims, style, basis):
    cfg = _cfg(lat, ncomp, dims)
    psi = _random(cfg) * (0.3 if style == pictures.PICTURE_SLICE else 1.0)
    path = tmp_path / "p.bmp"
    pictures.write(path, cfg, psi, style, basis)
    got = path.read_bytes()
    assert got == _reference_picture(cfg, psi, style, basis)
    header, rgb = pictures.read(path)
    W, H = 2 * (dims[0] + 2), 2 * (dims[1] + 2)
    w_expect = {pictures.PICTURE_SLICE: W}.get(style, (3 if ncomp == 3 else 5) * W)
    h_expect = H if style == pictures.PICTURE_SLICE else (3 * H if ncomp == 5 and style == pictures.PICTURE_DENSITY else 2 * H)
    assert (header["width"], header["height"]) == (w_expect, h_expect)
    assert header["size"] == 54 + 4 * w_expect * h_expect == header["offset"] + header["image_size"]
    assert (header["offset"], header["info"], header["planes"], header["bits"]) == (54, 40, 1, 24)
    assert rgb.shape == (h_expect, w_expect, 3)
    if style != pictures.PICTURE_SLICE:
        assert np.all(rgb[H, :, :] == 128)  # grey separator row
        if style == pictures.PICTURE_DENSITY:
            assert rgb.max() == 255  # each plane scaled to its largest column sum


def test_single_block_lands_in_the_mirrored_panels(tmp_path):
    """One occupied block at interior (x, y, z) = (0, 1, 0): padded block (1, 2, 1), pixels 2-3
    along x, 4-5 along y, 2-3 along z; spin-1 density pictures mirror x (XZ and XY panels) and y
    (XY panels), so its four XZ pixels sit at x = W-4..W-3, y = 2..3 of the m = 0 panel."""
    cfg = _cfg("bcc", 3, (2, 3, 2))
    psi = np.zeros(cfg.shape, complex)
    psi[1, 0, 1, 0, :] = 0.25
    path = tmp_path / "one.bmp"
    pictures.write(path, cfg, psi)
    _, rgb = pictures.read(path)
    W, H = 8, 10
    lit = np.argwhere(rgb[:, :, 0] == 255)
    xz = {(y, x - W) for y, x in lit if y < H}
    xy = {(y - H, x - W) for y, x in lit if y > H}
    assert xz == {(2, W - 4), (2, W - 3), (3, W - 4), (3, W - 3)}
    assert xy == {(H - 5, W - 4), (H - 5, W - 3), (H - 6, W - 4), (H - 6, W - 3)}
    grey = (rgb == 128).all(axis=2)
    assert np.all(rgb[~grey] % 255 == 0)  # everything else black


def _spin_matrix(F, axis):
    m = np.arange(F, -F - 1, -1).astype(float)
    fp = np.zeros((2 * F + 1, 2 * F + 1))
    for i in range(2 * F):
        fp[i, i + 1] = np.sqrt(F * (F + 1) - m[i + 1] * (m[i + 1] + 1))
    return [0.5 * (fp + fp.T), -0.5j * (fp - fp.T), np.diag(m)][axis].astype(complex)


# (ncomp, basis, axis, eigenvalue) -> panel index the reference's rotation sends the eigenstate
# to: its spin-1 x rotation maps f_x = -1 to the m = 1 slot, its y rotation f_y = +1 to m = 1;
# its spin-2 rotations map the +2 eigenstate of the axis to m = 2.
EIGEN = [(3, gpe.BASIS_X, 0, -1, 0), (3, gpe.BASIS_X, 0, 1, 2), (3, gpe.BASIS_X, 0, 0, 1),
         (3, gpe.BASIS_Y, 1, 1, 0), (3, gpe.BASIS_Y, 1, -1, 2), (3, gpe.BASIS_Z, 2, 1, 0),
         (5, gpe.BASIS_X, 0, 2, 0), (5, gpe.BASIS_Y, 1, 2, 0), (5, gpe.BASIS_Z, 2, -2, 4)]


@pytest.mark.parametrize("ncomp,basis,axis,ev,panel", EIGEN)
def test_axis_eigenstate_lights_one_panel(tmp_path, ncomp, basis, axis, ev, panel):
    cfg = _cfg("bcc", ncomp, (2, 2, 2))
    w, vec = np.linalg.eigh(_spin_matrix((ncomp - 1) // 2, axis))
    state = vec[:, int(np.argmin(np.abs(w - ev)))]
    psi = np.zeros(cfg.shape, complex)
    psi[:] = (0.3 + 0.4j) * state[:, None, None, None, None]
    psi *= RNG.uniform(0.5, 1.0, size=cfg.shape[1:])  # same spinor, varying density
    path = tmp_path / "e.bmp"
    pictures.write(path, cfg, psi, pictures.PICTURE_DENSITY, basis)
    _, rgb = pictures.read(path)
    W = 2 * (2 + 2)
    grey = (rgb == 128).all(axis=2)
    lit_panels = {x // W for y, x in np.argwhere((rgb[:, :, 0] > 0) & ~grey)}
    assert lit_panels == {panel}


def test_ri_channels_swap_under_multiplication_by_i(tmp_path):
    """Red carries sum Re^2, green sum Im^2: psi and i psi swap the two channels; a real state
    has no green; both are scaled by 2 / the largest density sum of the plane."""
    cfg = _cfg("bcc", 3, (2, 2, 2))
    psi = _random(cfg)
    a, b, c = tmp_path / "a.bmp", tmp_path / "b.bmp", tmp_path / "c.bmp"
    pictures.write(a, cfg, psi, pictures.PICTURE_RI)
    pictures.write(b, cfg, 1j * psi, pictures.PICTURE_RI)
    pictures.write(c, cfg, psi.real.astype(complex), pictures.PICTURE_RI)
    _, ra = pictures.read(a)
    _, rb = pictures.read(b)
    _, rc = pictures.read(c)
    grey = (ra == 128).all(axis=2)
    np.testing.assert_array_equal(ra[~grey][:, 0], rb[~grey][:, 1])
    np.testing.assert_array_equal(ra[~grey][:, 1], rb[~grey][:, 0])
    assert not np.any(rc[~grey][:, 1]) and not np.any(ra[:, :, 2][~grey])


def test_slice_channel_encoding(tmp_path):
    """Scalar slice: layer nz/2 + 1 of the padded grid, unknown 0, red = 20 Re^2 stored as
    floor(256 * 20 Re^2), clamped at 255; other layers do not show."""
    cfg = _cfg("bcc", 1, (3, 3, 4))
    psi = np.zeros(cfg.shape, complex)
    psi[0, 2, 1, 0, 0] = 0.1 + 0.02j   # interior z = 2 = padded layer 3 = nz / 2 + 1
    psi[0, 2, 1, 1, 0] = 0.5           # 20 * 0.25 * 256 > 255
    psi[0, 2, 1, 2, 1] = 0.3           # unknown 1: not drawn
    psi[0, 1, 1, 2, 0] = 0.3           # another layer: not drawn
    path = tmp_path / "s.bmp"
    pictures.write(path, cfg, psi, pictures.PICTURE_SLICE)
    _, rgb = pictures.read(path)
    assert rgb.shape == (10, 10, 3)
    r_expect = int(256.0 * (20.0 * (0.1 * 0.1)))
    g_expect = int(256.0 * (20.0 * (0.02 * 0.02)))
    np.testing.assert_array_equal(rgb[4:6, 2:4, 0], r_expect)
    np.testing.assert_array_equal(rgb[4:6, 2:4, 1], g_expect)
    np.testing.assert_array_equal(rgb[4:6, 4:6, 0], 255)
    mask = np.zeros((10, 10), bool)
    mask[4:6, 2:6] = True
    assert not np.any(rgb[~mask])


def test_spin2_yz_panels_use_the_xz_scale(tmp_path):
    """The reference scales the YZ panels by the XZ maximum.  With |psi_2|^2 = (x + 1)(y + 1) on
    8 x 2 x 2 blocks the XZ column sums (over y) peak at 3 * 8, the YZ sums (over x) are 36 and 72:
    on the XZ scale every interior YZ pixel saturates (on their own scale the y = 0 half would
    read 128), while the XZ panel shows its gradient."""
    cfg = _cfg("cube", 5, (8, 2, 2))
    psi = np.zeros(cfg.shape, complex)
    x = np.arange(8)[None, None, :, None]
    y = np.arange(2)[None, :, None, None]
    psi[0] = np.sqrt((x + 1.0) * (y + 1.0))
    path = tmp_path / "l.bmp"
    pictures.write(path, cfg, psi)
    _, rgb = pictures.read(path)
    W, H = 20, 8
    np.testing.assert_array_equal(rgb[H + 2:H + 6, 2:6, 0], 255)  # YZ interior, m = 2
    xz = rgb[2:6, 2:18, 0]
    np.testing.assert_array_equal(xz[:, -2:], 255)
    np.testing.assert_array_equal(xz[:, :2], int(256.0 * (3.0 * 2 / 48.0)))


def test_bad_styles_are_rejected(tmp_path):
    scalar = _cfg("bcc", 1, (2, 2, 2))
    spin2 = _cfg("bcc", 5, (2, 2, 2))
    spin1 = _cfg("bcc", 3, (2, 2, 2))
    with pytest.raises(gpe.GpeError):
        pictures.write(tmp_path / "a.bmp", scalar, _random(scalar), pictures.PICTURE_DENSITY)
    with pytest.raises(gpe.GpeError):
        pictures.write(tmp_path / "b.bmp", spin1, _random(spin1), pictures.PICTURE_SLICE)
    with pytest.raises(gpe.GpeError):
        pictures.write(tmp_path / "c.bmp", spin2, _random(spin2), pictures.PICTURE_RI, gpe.BASIS_X)
    with pytest.raises(gpe.GpeError):
        pictures.write(tmp_path / "d.bmp", spin1, _random(spin1), 7)
    with pytest.raises(gpe.GpeError):
        pictures.write(tmp_path / "no_such_dir" / "e.bmp", spin1, _random(spin1))


def test_picture_path_is_the_reference_name():
    assert pictures.picture_path("out", "GS", 0.25) == "out/GS0.250000000000000000ms.bmp"
