"""Unit-cell tables: structure, metric and spectral checks that do not rely on the tables'
own formula (symmetry, volumes, Bloch symbol -> |k|^2/2, BCC spectrum, cross-source match)."""
import json
import os

import numpy as np
import pytest

from spacetime_amd import lattice

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "gpblocks_scalar_bcc.json")


@pytest.mark.parametrize("name", lattice.LATTICES)
def test_table_shape_and_symmetry(name):
    lat = lattice.get(name)
    expected_v = {"cube": 1, "bcc": 12, "fcc": 12, "a15": 46, "c15": 136, "z": 80}[name]
    assert lat.V == expected_v
    assert np.all(np.abs(lat.face_d) <= 1)
    assert np.all((lat.face_tgt >= 0) & (lat.face_tgt < lat.V))
    assert np.all(lat.face_hodge > 0)
    faces = {(int(lat.face_src[f]), int(lat.face_tgt[f]), tuple(int(x) for x in lat.face_d[f])): lat.face_hodge[f]
             for f in range(lat.nfaces)}
    for (i, j, d), h in faces.items():
        back = faces[(j, i, tuple(-x for x in d))]
        # volume-weighted symmetry vol_i h_ij == vol_j h_ji
        assert abs(lat.vol[i] * h - lat.vol[j] * back) < 1e-12 * lat.vol[i] * h


@pytest.mark.parametrize("name", lattice.LATTICES)
def test_volumes(name):
    lat = lattice.get(name)
    block = float(np.prod(lat.block_dims))
    assert abs(lat.vol.sum() - block) < 1e-10 * block
    # the reference's single VOLUME constant is the volume of unknown 0's body
    assert abs(lat.vol[0] - lat.ref_volume) < 1e-12 * lat.ref_volume
    if name in ("cube", "bcc"):
        assert np.allclose(lat.vol, lat.ref_volume, rtol=1e-12)


@pytest.mark.parametrize("name", lattice.LATTICES)
def test_bloch_symbol_is_half_k_squared(name):
    """Lowest band of -1/2 L tends to |k|^2/2 (the continuum kinetic energy)."""
    lat = lattice.get(name)
    rng = np.random.default_rng(7)
    for _ in range(3):
        direction = rng.normal(size=3)
        direction /= np.linalg.norm(direction)
        for kk in (1e-3, 3e-3):
            k = kk * direction
            ev = np.linalg.eigvalsh(lat.laplacian_symbol(k))
            assert abs(ev[0] - 0.5 * kk * kk) < 2e-3 * 0.5 * kk * kk
            # k = 0: constants are in the kernel
    ev0 = np.linalg.eigvalsh(lat.laplacian_symbol(np.zeros(3)))
    assert abs(ev0[0]) < 1e-10


def test_bcc_spectrum_max_is_6_over_h2():
    """Spectrum of -1/2 L for the BCC table is [0, 6/h^2] with h = dual edge = block_scale."""
    lat = lattice.get("bcc")
    s = 0.1
    best = 0.0
    a = s * lat.block_width
    for kx in np.linspace(0, 2 * np.pi / a, 9):
        for ky in np.linspace(0, 2 * np.pi / a, 9):
            for kz in np.linspace(0, 2 * np.pi / a, 9):
                best = max(best, np.linalg.eigvalsh(lat.laplacian_symbol([kx, ky, kz], s))[-1])
    assert abs(best * s * s - 6.0) < 1e-9


def test_bcc_table_matches_gpblocks_uint_table():
    """Spin1GpeBCC/mesh.h int4 table == GPBlocks.cpp:397-449 uint-offset table (printed by the
    compiled reference) and positions == GPBlocks.cpp:380-396."""
    with open(GOLDEN) as f:
        g = json.load(f)
    lat = lattice.get("bcc")
    st = g["bcc_ind_strides"]
    ref = g["cases"][0]["bcc_ind"]
    for f in range(lat.nfaces):
        dx, dy, dz = (int(x) for x in lat.face_d[f])
        val = (int(lat.face_tgt[f]) + dx * st["nx"] + dy * st["ny"] + dz * st["nz"]) % 2 ** 32
        assert val == ref[f]
    assert np.max(np.abs(np.array(g["cases"][0]["bcc_pos"]) - lat.pos)) < 1e-14
    assert np.allclose(lat.face_hodge, 1.5, rtol=1e-14)


def test_cube_table_is_the_seven_point_stencil():
    lat = lattice.get("cube")
    dirs = sorted(tuple(int(x) for x in d) for d in lat.face_d)
    assert dirs == sorted([(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)])
    assert np.all(lat.face_hodge == 1.0) and lat.block_width == 1.0


def test_fcc_degrees():
    lat = lattice.get("fcc")
    assert sorted(set(lat.degree.tolist())) == [4, 8]
    assert int((lat.degree == 8).sum()) == 4  # octahedra


def test_packed_faces_roundtrip():
    for name in lattice.LATTICES:
        lat = lattice.get(name)
        p = lat.packed_faces()
        assert np.array_equal((p & 3) - 1, lat.face_d[:, 0])
        assert np.array_equal(((p >> 2) & 3) - 1, lat.face_d[:, 1])
        assert np.array_equal(((p >> 4) & 3) - 1, lat.face_d[:, 2])
        assert np.array_equal(p >> 8, lat.face_tgt)


@pytest.mark.parametrize("name", ["cube", "bcc", "fcc", "a15", "c15", "z"])
def test_kinetic_bound_bounds_the_bloch_spectrum(name):
    """Lattice.kinetic_bound (Gershgorin) is >= every Bloch eigenvalue of -1/2 L, and is the
    exact maximum 6/s^2 on BCC (spectrum [0, 6/h^2])."""
    L = lattice.get(name)
    s = 0.37
    bound = L.kinetic_bound(s)
    lam = 0.0
    for k in np.random.default_rng(3).uniform(-np.pi, np.pi, size=(60, 3)) / (s * L.block_width):
        lam = max(lam, np.linalg.eigvalsh(L.laplacian_symbol(k, s)).max())
    assert lam <= bound * (1 + 1e-12)
    if name == "bcc":
        assert bound == pytest.approx(6.0 / s ** 2)


@pytest.mark.parametrize("name", ["cube", "bcc", "fcc", "a15", "c15", "z"])
def test_tables_match_an_independent_dec_construction(name):
    """The converted reference tables are the DEC discretisation of the lattice they name,
    rebuilt here from the textbook atom positions without the reference's builder: unknowns at
    the Voronoi vertices (Delaunay circumcentres; FCC octahedra and simple-cubic cubes merged),
    Hodge ratio |face| / (|dual edge| |cell|).  Positions agree up to a translation, every face
    (block offset and target) agrees exactly, the Hodge weights to 1e-9 (the precision the
    reference printed them with), and the per-unknown dual volumes solved from Hodge symmetry
    equal the rebuilt cell volumes."""
    from spacetime_amd.tables import dec_builder
    r = dec_builder.compare(name, lattice.get(name))
    assert r["topology"]
    assert r["max_hodge_diff"] < 1e-9
    assert r["max_volume_diff"] < 1e-12
