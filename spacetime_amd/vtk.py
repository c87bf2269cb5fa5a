"""Legacy binary VTK files in the reference's layouts, written by libgpe's host code.

The reference's viewers read these files with ``tvtk.PolyDataReader`` and take the point
arrays **by index** (``point_data.get_array(i)``), so the writers reproduce the reference's
files byte for byte where its arithmetic is exact:

* spinor (``saveSpinor``, Spin1GpeBCC/utils.h:1079-1314, Spin2GpeBCC/utils.h:1071-1744): arrays
  ``r_m=F, i_m=F, ..., r_m=-F, i_m=-F`` = Re / Im psi_m / sqrt(n) -- what ferro_skyrmion.py
  (:105-110), bn_skyrmion.py (:234-243) and integrate_dens.py (:198-202) read as arrays 0..2C-1;
* pre-image (``savePreImageSpinor``, Spin1GpeBCC/utils.h:1351-1715): the spinor arrays on the
  points whose spin direction has a y component above 0.9996 (ferro_skyrmion.py:112-123);
* volume (``saveVolume``): spin-1 density, spinNorm, theta, localAvgSpin, u
  (Spin1GpeBCC/utils.h:792-1077); spin-2 densities m=2..-2 (Spin2GpeBCC/utils.h:809-1069).

Points: every unknown of the one-block-padded grid in the order z, x, y, v (the halo points
carry zeros) at p0 + s (i_padded * block edge + pos_v), as the reference writes them.  Values
are zero where the density is at most DENSITY_THRESHOLD.  Writers take a Stepper (a level on
the device) or a Config and a host state in the ABI layout.
"""
import os

import numpy as np

from . import gpe

DENSITY_THRESHOLD = 1e-4  # Spin1GpeBCC/utils.h:789


def _host(name):
    lib = gpe.load_library()
    f = getattr(lib, name)
    return lib, f


def _state(cfg, psi):
    psi = np.ascontiguousarray(psi, dtype=np.complex128)
    if psi.shape != cfg.shape:
        raise ValueError("state shape %s, expected %s" % (psi.shape, cfg.shape))
    return psi


def _call(rc, what):
    if rc != 0:
        raise gpe.GpeError("%s failed: %s" % (what, gpe.ERRORS.get(rc, rc)))


def write_spinor(path, cfg, psi, basis=gpe.BASIS_Z):
    """saveSpinor layout for a host state (gpe_host_write_vtk_spinor)."""
    psi = _state(cfg, psi)
    lib, f = _host("gpe_host_write_vtk_spinor")
    c = cfg.to_c()
    _call(f(c, psi.ctypes.data, os.fsencode(str(path)), basis), "gpe_host_write_vtk_spinor")


def write_preimage(path, cfg, psi, axis=1, threshold=0.9996):
    """savePreImageSpinor layout for a host spin-1 state (gpe_host_write_vtk_preimage)."""
    psi = _state(cfg, psi)
    lib, f = _host("gpe_host_write_vtk_preimage")
    _call(f(cfg.to_c(), psi.ctypes.data, os.fsencode(str(path)), axis, float(threshold)),
          "gpe_host_write_vtk_preimage")


def write_volume(path, cfg, psi, basis=gpe.BASIS_Z):
    """saveVolume layout for a host state (gpe_host_write_vtk_volume)."""
    psi = _state(cfg, psi)
    lib, f = _host("gpe_host_write_vtk_volume")
    _call(f(cfg.to_c(), psi.ctypes.data, os.fsencode(str(path)), basis), "gpe_host_write_vtk_volume")


def save_spinor(stepper, path, level=0, basis=gpe.BASIS_Z):
    stepper.write_vtk_spinor(str(path), level, basis)


def save_preimage(stepper, path, level=0, axis=1, threshold=0.9996):
    stepper.write_vtk_preimage(str(path), level, axis, threshold)


def save_volume(stepper, path, level=0, basis=gpe.BASIS_Z):
    stepper.write_vtk_volume(str(path), level, basis)


def read(path):
    """Read a legacy binary POLYDATA file: (points (n, 3), [(name, array), ...]) with the point
    arrays in file order -- the indices under which tvtk's point_data.get_array(i) returns
    them -- scalars (n,), vectors (n, 3), as float32."""
    with open(path, "rb") as f:
        data = f.read()
    arrays = []
    points = None
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
            points = floats(3 * n).reshape(n, 3)
        elif words[0] == "SCALARS":
            if line().split()[0] != "LOOKUP_TABLE":
                raise ValueError("SCALARS without LOOKUP_TABLE")
            arrays.append((words[1], floats(n)))
        elif words[0] == "VECTORS":
            arrays.append((words[1], floats(3 * n).reshape(n, 3)))
    return points, arrays


def read_volume(path):
    """Named access to a file's arrays: dict of "points" and each array by name."""
    points, arrays = read(path)
    out = {"points": points}
    out.update(dict(arrays))
    return out
