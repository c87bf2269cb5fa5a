"""Legacy binary VTK volume files laid out as the reference's `saveVolume` writes them
(Spin1GpeBCC/utils.h:792-1080), so its Python viewers (ferro_skyrmion.py, integrate_dens.py:
`tvtk.PolyDataReader`, fields looked up by name) read them unchanged:

    # vtk DataFile Version 3.0 / comment / BINARY / DATASET POLYDATA
    POINTS n float                      one point per unknown, big-endian float32 x, y, z
    POINT_DATA n
    SCALARS density float 1             |psi|^2 summed over the components
    SCALARS spinNorm float 1            |<F>|              (spinor states)
    SCALARS theta float 1               spin-1 phase theta  (spin-1 states)
    VECTORS localAvgSpin float          <F> = psi^dag F psi / |psi|^2 (spinor states)

Every field but the density is 0 where the density is at most DENSITY_THRESHOLD (utils.h:788,
903-935).  The points follow this package's unknown order (z, y, x, v); the reference walks
(z, x, y, v) over its halo-padded grid (utils.h:806-827), whose halo points carry zeros.
"""
import numpy as np

DENSITY_THRESHOLD = 1e-4  # utils.h:788

_HEADER = b"# vtk DataFile Version 3.0\nComment if needed\nBINARY\n"


def write_volume(path, positions, density, spin=None, theta=None):
    """Write one volume file.  positions (..., 3); density (...); spin (..., 3) or None;
    theta (...) or None; all over the same unknowns."""
    pos = np.asarray(positions, dtype=np.float64).reshape(-1, 3)
    n = pos.shape[0]
    dens = np.asarray(density, dtype=np.float64).reshape(-1)
    if dens.shape[0] != n:
        raise ValueError("density has %d values for %d points" % (dens.shape[0], n))
    live = dens > DENSITY_THRESHOLD

    def be(a):
        return np.ascontiguousarray(a, dtype=">f4").tobytes()

    with open(path, "wb") as f:
        f.write(_HEADER)
        f.write(b"DATASET POLYDATA\nPOINTS %d float\n" % n)
        f.write(be(pos))
        f.write(b"\nPOINT_DATA %d\nSCALARS density float 1\nLOOKUP_TABLE default\n" % n)
        f.write(be(dens))
        if spin is not None:
            s = np.asarray(spin, dtype=np.float64).reshape(-1, 3)
            if s.shape[0] != n:
                raise ValueError("spin has %d vectors for %d points" % (s.shape[0], n))
            s = np.where(live[:, None], s, 0.0)
            f.write(b"\nSCALARS spinNorm float 1\nLOOKUP_TABLE default\n")
            f.write(be(np.sqrt(np.sum(s * s, axis=1))))
        if theta is not None:
            th = np.asarray(theta, dtype=np.float64).reshape(-1)
            if th.shape[0] != n:
                raise ValueError("theta has %d values for %d points" % (th.shape[0], n))
            f.write(b"\nSCALARS theta float 1\nLOOKUP_TABLE default\n")
            f.write(be(np.where(live, th, 0.0)))
        if spin is not None:
            f.write(b"\nVECTORS localAvgSpin float\n")
            f.write(be(s))
        f.write(b"\n")


def read_volume(path):
    """Read a file written by write_volume (or the reference's saveVolume): dict with
    "points" (n, 3) and each named field ((n,) scalars, (n, 3) vectors), as float32."""
    with open(path, "rb") as f:
        data = f.read()
    out = {}
    pos = 0

    def line():
        nonlocal pos
        end = data.index(b"\n", pos)
        text = data[pos:end].decode("ascii")
        pos = end + 1
        return text

    def floats(count):
        nonlocal pos
        a = np.frombuffer(data, dtype=">f4", count=count, offset=pos).astype(np.float32)
        pos += 4 * count
        return a

    n = None
    while pos < len(data):
        words = line().split()
        if not words:
            continue
        if words[0] == "POINTS":
            n = int(words[1])
            out["points"] = floats(3 * n).reshape(n, 3)
        elif words[0] == "SCALARS":
            if line().split()[0] != "LOOKUP_TABLE":
                raise ValueError("SCALARS without LOOKUP_TABLE")
            out[words[1]] = floats(n)
        elif words[0] == "VECTORS":
            out[words[1]] = floats(3 * n).reshape(n, 3)
    return out


def save_volume(stepper, path, level=0):
    """Write the volume file of one time level of a Stepper: density, and for spinor states
    <F> with its norm, and for spin-1 the phase theta (field maps from gpe_local_fields)."""
    from . import gpe, states
    cfg = stepper.config
    dens = stepper.local_fields(gpe.FIELD_DENSITY, level)[..., 0]
    spin = stepper.local_fields(gpe.FIELD_SPIN, level) if cfg.ncomp > 1 else None
    theta = stepper.local_fields(gpe.FIELD_UV_THETA, level)[..., 6] if cfg.ncomp == 3 else None
    write_volume(path, states.positions(cfg), dens, spin, theta)
