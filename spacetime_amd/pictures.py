"""The reference's BMP density pictures, written by libgpe's host code.

Three kinds, byte for byte as the reference writes them (Picture::save, Output/Picture.cpp:83-138):

* ``PICTURE_DENSITY``: per spin component, the density summed along one axis of the padded grid,
  grey, two pixels per block edge -- spin-1 ``drawDensity`` (Spin1GpeBCC/utils.h:151-366: XZ and
  XY panels, mirrored) and spin-2 ``drawDensity`` (Spin2GpeBCC/utils.h:287-796: XZ, YZ, XY);
* ``PICTURE_RI``: red = sum of Re^2, green = sum of Im^2 -- spin-1 ``drawDensityRI``
  (Spin1GpeBCC/utils.h:368-580) and spin-2 ``drawIandR`` (Spin2GpeBCC/utils.h:131-285);
* ``PICTURE_SLICE``: a scalar state's middle block layer, red = 20 Re^2, green = 20 Im^2
  (GPUGrossPitaevskii_BCC.cu:700-723).

Writers take a Stepper (a level on the device) or a Config and a host state in the ABI layout;
``read`` parses a file back into its header fields and an RGB array indexed [y][x].
"""
import os
import struct

import numpy as np

from . import gpe

PICTURE_DENSITY, PICTURE_RI, PICTURE_SLICE = gpe.PICTURE_DENSITY, gpe.PICTURE_RI, gpe.PICTURE_SLICE


def picture_path(folder, name, t):
    """The reference's picture file name: folder/name + t (fixed, 18 decimals) + "ms.bmp"
    (Spin1GpeBCC/utils.h:365 with toString, :102-108)."""
    return os.path.join(str(folder), "%s%.18fms.bmp" % (name, t))


def write(path, cfg, psi, style=PICTURE_DENSITY, basis=gpe.BASIS_Z):
    """Picture of a host state (gpe_host_write_bmp)."""
    psi = np.ascontiguousarray(psi, dtype=np.complex128)
    if psi.shape != cfg.shape:
        raise ValueError("state shape %s, expected %s" % (psi.shape, cfg.shape))
    lib = gpe.load_library()
    rc = lib.gpe_host_write_bmp(cfg.to_c(), psi.ctypes.data, os.fsencode(str(path)), style, basis)
    if rc != 0:
        raise gpe.GpeError("gpe_host_write_bmp failed: %s" % gpe.ERRORS.get(rc, rc))


def save(stepper, path, level=0, style=PICTURE_DENSITY, basis=gpe.BASIS_Z):
    """Picture of a level of a Stepper (gpe_write_bmp)."""
    stepper.write_bmp(str(path), level, style, basis)


def read(path):
    """(header, rgb): the header's fields and the pixels as uint8 [height][width][r, g, b], row 0
    being the first row of the file."""
    data = open(path, "rb").read()
    if data[:2] != b"BM":
        raise ValueError("%s: not a BMP file" % path)
    size, reserved, offset = struct.unpack_from("<III", data, 2)
    info, w, h, planes, bits, compression, image = struct.unpack_from("<IiiHHII", data, 14)
    header = dict(size=size, reserved=reserved, offset=offset, info=info, width=w, height=h, planes=planes,
                  bits=bits, compression=compression, image_size=image, tail=struct.unpack_from("<IIII", data, 38))
    if bits != 24 or compression != 0:
        raise ValueError("%s: only uncompressed 24-bit pictures are read" % path)
    stride = 3 * w + (3 - (3 * w + 3) % 4)
    if len(data) != offset + stride * h:
        raise ValueError("%s: %d bytes, expected %d" % (path, len(data), offset + stride * h))
    rows = np.frombuffer(data, np.uint8, count=stride * h, offset=offset).reshape(h, stride)
    bgr = rows[:, :3 * w].reshape(h, w, 3)
    return header, np.ascontiguousarray(bgr[:, :, ::-1])
