"""VTK files in the reference's layouts, written by libgpe's host code from host states (no
device needed).  The reference's viewers take the point arrays by index
(ferro_skyrmion.py:105-110 spin-1 r/i m=1,0,-1 as arrays 0..5; bn_skyrmion.py:234-243 and
integrate_dens.py:198-202 spin-2 arrays 0..9), so the tests read the files back by array index
and check what those indices hold, the padded point order (z, x, y, v), the zero halo, and
properties the spinor algebra fixes (unit-norm normalised spinors, basis rotations that turn
an x- or y-polarised state into m = +2, spin directions of the pre-image selection)."""
import numpy as np
import pytest

from spacetime_amd import gpe, lattice, states, vtk

RNG = np.random.default_rng(11)


def _cfg(lat, ncomp, dims, cell=0.7):
    return gpe.Config.centred_box(lat, dims, cell, ncomp=ncomp)


def _random(cfg, scale=1.0):
    psi = RNG.normal(size=cfg.shape) + 1j * RNG.normal(size=cfg.shape)
    return scale * psi


def _padded_points(cfg):
    """Positions and interior unknown of every padded point in the reference's order."""
    lat = lattice.get(cfg.lattice)
    nx, ny, nz = cfg.dims
    pts, idx = [], []
    for z in range(nz + 2):
        for x in range(nx + 2):
            for y in range(ny + 2):
                for v in range(lat.V):
                    p = [cfg.p0[a] + cfg.block_scale * ((x, y, z)[a] * lat.block_dims[a] + lat.pos[v, a]) for a in range(3)]
                    pts.append(p)
                    inside = 0 < x <= nx and 0 < y <= ny and 0 < z <= nz
                    idx.append((z - 1, y - 1, x - 1, v) if inside else None)
    return np.array(pts, dtype=np.float32), idx


@pytest.mark.parametrize("lat,ncomp,dims", [("bcc", 3, (2, 3, 2)), ("cube", 5, (3, 2, 4)), ("a15", 3, (1, 2, 1))])
def test_spinor_file_by_array_index(tmp_path, lat, ncomp, dims):
    cfg = _cfg(lat, ncomp, dims)
    psi = _random(cfg)
    psi[:, 0, 0, 0, 0] *= 1e-3  # one unknown below the density threshold
    path = tmp_path / "s.vtk"
    vtk.write_spinor(path, cfg, psi)
    head = path.read_bytes()[:100]
    assert head.startswith(b"# vtk DataFile Version 3.0\nComment if needed\nBINARY\nDATASET POLYDATA\nPOINTS ")
    points, arrays = vtk.read(path)
    F = (ncomp - 1) // 2
    names = [n for n, _ in arrays]
    expect = []
    for m in range(F, -F - 1, -1):
        expect += ["r_m=%d" % m, "i_m=%d" % m]
    assert names == expect  # the viewers' array indices
    pts, idx = _padded_points(cfg)
    np.testing.assert_array_equal(points, pts)
    spinor = np.zeros((len(idx), ncomp), complex)
    for c in range(ncomp):
        spinor[:, c] = arrays[2 * c][1] + 1j * arrays[2 * c + 1][1]
    n = np.sum(np.abs(psi) ** 2, axis=0)
    live = 0
    for k, u in enumerate(idx):
        if u is None:
            assert not np.any(spinor[k])  # zero halo
            continue
        if n[u] <= vtk.DENSITY_THRESHOLD:
            assert not np.any(spinor[k])
            continue
        live += 1
        # the stored spinor is psi / sqrt(n) at that unknown: unit norm, parallel to psi
        assert np.sum(np.abs(spinor[k]) ** 2) == pytest.approx(1.0, abs=1e-6)
        np.testing.assert_allclose(spinor[k] * np.sqrt(n[u]), psi[(slice(None),) + u], rtol=0, atol=1e-5 * np.sqrt(n[u]))
    assert live == np.prod(dims) * cfg.V - 1


def _spin_matrix(F, axis):
    m = np.arange(F, -F - 1, -1).astype(float)
    fp = np.zeros((2 * F + 1, 2 * F + 1))
    for i in range(2 * F):
        fp[i, i + 1] = np.sqrt(F * (F + 1) - m[i + 1] * (m[i + 1] + 1))
    return [0.5 * (fp + fp.T), -0.5j * (fp - fp.T), np.diag(m)][axis].astype(complex)


@pytest.mark.parametrize("basis,axis", [(gpe.BASIS_X, 0), (gpe.BASIS_Y, 1), (gpe.BASIS_Z, 2)])
def test_spin2_basis_turns_the_axis_eigenstate_into_m2(tmp_path, basis, axis):
    """A spinor polarised along the basis axis (eigenvalue +2 of f_axis) reads as |m=2| = 1 in
    that basis, and every rotated spinor keeps unit norm (the basis change is unitary)."""
    cfg = _cfg("cube", 5, (2, 1, 1))
    w, vec = np.linalg.eigh(_spin_matrix(2, axis))
    top = vec[:, np.argmax(w)]
    psi = np.zeros(cfg.shape, complex)
    psi[:, 0, 0, 0, 0] = 0.7 * top
    psi[:, 0, 0, 1, 0] = 0.9 * _random(cfg)[:, 0, 0, 1, 0] / 3
    path = tmp_path / "b.vtk"
    vtk.write_spinor(path, cfg, psi, basis=basis)
    _, arrays = vtk.read(path)
    _, idx = _padded_points(cfg)
    k0 = idx.index((0, 0, 0, 0))
    k1 = idx.index((0, 0, 1, 0))
    s0 = np.array([arrays[2 * c][1][k0] + 1j * arrays[2 * c + 1][1][k0] for c in range(5)])
    s1 = np.array([arrays[2 * c][1][k1] + 1j * arrays[2 * c + 1][1][k1] for c in range(5)])
    assert abs(s0[0]) == pytest.approx(1.0, abs=1e-6)
    assert np.sum(np.abs(s1) ** 2) == pytest.approx(1.0, abs=1e-6)


def test_preimage_selects_y_polarised_interior_points(tmp_path):
    cfg = _cfg("bcc", 3, (2, 2, 1))
    w, vec = np.linalg.eigh(_spin_matrix(1, 1))
    up_y = vec[:, np.argmax(w)]
    psi = np.zeros(cfg.shape, complex)
    psi[0] = 0.5  # polarised along z elsewhere
    sel = [(0, 1, 0, 3), (0, 0, 1, 7), (0, 1, 1, 11)]
    for u in sel:
        psi[(slice(None),) + u] = 0.8 * up_y
    path = tmp_path / "p.vtk"
    vtk.write_preimage(path, cfg, psi)
    points, arrays = vtk.read(path)
    assert [n for n, _ in arrays] == ["r_m=1", "i_m=1", "r_m=0", "i_m=0", "r_m=-1", "i_m=-1"]
    pts, idx = _padded_points(cfg)
    order = [k for k, u in enumerate(idx) if u in sel]
    np.testing.assert_array_equal(points, pts[order])
    for j in range(len(sel)):
        s = np.array([arrays[2 * c][1][j] + 1j * arrays[2 * c + 1][1][j] for c in range(3)])
        np.testing.assert_allclose(s, up_y, atol=1e-6)


def test_spin1_volume_fields(tmp_path):
    """density, spinNorm, theta, localAvgSpin, u (arrays 0-4).  The reference's local spin
    divides the transverse part by n but not the z part (Spin1GpeBCC.cu:251-261): a z-polarised
    unknown of density n reads spinNorm n, an x-polarised one 1; a polar spinor e^{i phi}(0, a, 0)
    has theta = pi - phi and u = (0, 0, -a)."""
    cfg = _cfg("bcc", 3, (1, 1, 1))
    psi = np.zeros(cfg.shape, complex)
    n_z, n_x, a, phi = 0.36, 0.49, 0.6, 0.3
    psi[0, 0, 0, 0, 0] = np.sqrt(n_z)                              # z-polarised
    psi[:, 0, 0, 0, 1] = np.sqrt(n_x) * np.array([0.5, 2 ** -0.5, 0.5])  # x-polarised
    psi[1, 0, 0, 0, 2] = a * np.exp(1j * phi)                        # polar
    psi[1, 0, 0, 0, 3] = 1e-3                                        # below the threshold
    path = tmp_path / "v.vtk"
    vtk.write_volume(path, cfg, psi)
    _, arrays = vtk.read(path)
    assert [n for n, _ in arrays] == ["density", "spinNorm", "theta", "localAvgSpin", "u"]
    _, idx = _padded_points(cfg)
    k = [idx.index((0, 0, 0, v)) for v in range(4)]
    dens, snorm, theta, spin, u = (arr for _, arr in arrays)
    np.testing.assert_allclose(dens[k], [n_z, n_x, a * a, 1e-6], rtol=1e-6)
    assert snorm[k[0]] == pytest.approx(n_z, rel=1e-6)
    assert snorm[k[1]] == pytest.approx(1.0, rel=1e-6)
    np.testing.assert_allclose(spin[k[1]], [1, 0, 0], atol=1e-6)
    assert theta[k[2]] == pytest.approx(np.pi - phi, rel=1e-6)
    np.testing.assert_allclose(u[k[2]], [0, 0, -a], atol=1e-6)
    assert snorm[k[3]] == 0 and theta[k[3]] == 0 and not np.any(u[k[3]])  # below the threshold
    assert not np.any(dens[[j for j, q in enumerate(idx) if q is None]])


def test_spin2_volume_densities(tmp_path):
    cfg = _cfg("fcc", 5, (1, 2, 1))
    psi = _random(cfg)
    path = tmp_path / "v2.vtk"
    vtk.write_volume(path, cfg, psi)
    _, arrays = vtk.read(path)
    assert [n for n, _ in arrays] == ["m=2", "m=1", "m=0", "m=-1", "m=-2"]
    _, idx = _padded_points(cfg)
    live = [(k, u) for k, u in enumerate(idx) if u is not None]
    for c in range(5):
        got = arrays[c][1][[k for k, _ in live]]
        want = np.array([abs(psi[(c,) + u]) ** 2 for _, u in live])
        np.testing.assert_allclose(got, want, rtol=1e-6)


def test_writers_refuse_scalar_and_bad_basis(tmp_path):
    cfg = _cfg("bcc", 1, (1, 1, 1))
    with pytest.raises(gpe.GpeError):
        vtk.write_spinor(tmp_path / "x.vtk", cfg, np.zeros(cfg.shape, complex))
    cfg3 = _cfg("bcc", 3, (1, 1, 1))
    with pytest.raises(gpe.GpeError):
        vtk.write_spinor(tmp_path / "x.vtk", cfg3, np.zeros(cfg3.shape, complex), basis=gpe.BASIS_X)


def test_dat_files_of_spin2_and_scalar_round_trip(tmp_path):
    for lat, ncomp, dims in (("c15", 5, (1, 1, 2)), ("cube", 1, (4, 3, 2))):
        cfg = _cfg(lat, ncomp, dims)
        psi = _random(cfg)
        p = tmp_path / ("%s.dat" % lat)
        states.write_reference_dat(p, cfg, psi)
        assert p.stat().st_size == (dims[0] + 2) * (dims[1] + 2) * (dims[2] + 2) * cfg.V * ncomp * 16
        np.testing.assert_array_equal(states.read_reference_dat(p, cfg), psi)
