"""VTK volume files in the reference's saveVolume layout (Spin1GpeBCC/utils.h:792-1080): the
exact bytes of a two-point file, written out by hand with struct, and a round trip of a
random file with the density threshold applied."""
import struct

import numpy as np

from spacetime_amd import vtk


def _be(*xs):
    return b"".join(struct.pack(">f", x) for x in xs)


def test_two_point_file_bytes(tmp_path):
    p = tmp_path / "v.vtk"
    pos = [[0.5, -1.0, 2.0], [3.0, 0.25, -0.125]]
    dens = [2.0, 5e-5]                 # the second point is below the density threshold
    spin = [[0.0, 0.6, 0.8], [1.0, 0.0, 0.0]]
    theta = [1.5, 2.5]
    vtk.write_volume(p, pos, dens, spin, theta)
    want = (b"# vtk DataFile Version 3.0\nComment if needed\nBINARY\nDATASET POLYDATA\nPOINTS 2 float\n"
            + _be(0.5, -1.0, 2.0, 3.0, 0.25, -0.125)
            + b"\nPOINT_DATA 2\nSCALARS density float 1\nLOOKUP_TABLE default\n" + _be(2.0, 5e-5)
            + b"\nSCALARS spinNorm float 1\nLOOKUP_TABLE default\n" + _be(1.0, 0.0)
            + b"\nSCALARS theta float 1\nLOOKUP_TABLE default\n" + _be(1.5, 0.0)
            + b"\nVECTORS localAvgSpin float\n" + _be(0.0, 0.6, 0.8, 0.0, 0.0, 0.0) + b"\n")
    assert p.read_bytes() == want


def test_round_trip_and_threshold(tmp_path):
    rng = np.random.default_rng(3)
    n = (3, 4, 5, 12)
    pos = rng.normal(size=n + (3,))
    dens = rng.uniform(0, 2e-4, size=n)
    spin = rng.normal(size=n + (3,))
    theta = rng.uniform(-np.pi, np.pi, size=n)
    p = tmp_path / "r.vtk"
    vtk.write_volume(p, pos, dens, spin, theta)
    got = vtk.read_volume(p)
    live = dens.reshape(-1) > vtk.DENSITY_THRESHOLD
    assert 0 < live.sum() < live.size
    f32 = lambda a: np.asarray(a, dtype=np.float32)  # noqa: E731
    np.testing.assert_array_equal(got["points"], f32(pos.reshape(-1, 3)))
    np.testing.assert_array_equal(got["density"], f32(dens.reshape(-1)))
    s = np.where(live[:, None], spin.reshape(-1, 3), 0.0)
    np.testing.assert_array_equal(got["localAvgSpin"], f32(s))
    np.testing.assert_allclose(got["spinNorm"], np.linalg.norm(s, axis=1), rtol=1e-6)
    np.testing.assert_array_equal(got["theta"], f32(np.where(live, theta.reshape(-1), 0.0)))


def test_density_only_file(tmp_path):
    p = tmp_path / "d.vtk"
    vtk.write_volume(p, np.zeros((4, 3)), np.arange(4.0))
    got = vtk.read_volume(p)
    assert set(got) == {"points", "density"}
    np.testing.assert_array_equal(got["density"], np.arange(4.0, dtype=np.float32))
