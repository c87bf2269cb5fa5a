"""Convert the reference's per-lattice ``mesh.h`` unit-cell tables into repo-owned JSON.

This is the one-off recipe that produced ``cube.json``, ``bcc.json``, ``fcc.json``,
``a15.json``, ``c15.json`` and ``z.json`` next to it.  It reads the reference checkout (default
``/root/reference``) as text, so it only runs where that checkout exists; the JSON
files it writes are the data every other part of the repository uses.

What is extracted from each ``mesh.h`` (the tables are generated constants of the
reference's weighted-Delaunay builder, ``GrossPitaevskii/GPBlocks.cpp:38-301``):

* ``VALUES_IN_BLOCK`` unknowns (dual 0-cells = circumcentres) per cubic block,
* ``BLOCK_WIDTH`` (edges of the block in table units; cubic except Z), ``VOLUME`` (dual volume
  of one unknown in table units), ``DUAL_EDGE_LENGTH``,
* the unknowns' positions inside the block,
* for every unknown its faces: neighbour block offset (dx, dy, dz) in {-1,0,1}^3,
  target unknown and the Hodge ratio ``hodge`` (star-dual-edge / star-edge).

Sources and formats:

* Cube  ``Spin1GpeCube/mesh.h:15-42``  int4 (dx, dy, dz, target) offsets
* BCC   ``Spin1GpeBCC/mesh.h:15-161``  int4 offsets, block width 2*sqrt(2), hodge 1.5
* FCC   ``GrossPitaevskiiGpuFcc/mesh.h:8-172`` int2 (byte offset expression in
  nx/ny/nz, target) plus ``indicesAndFaceCounts`` (variable degree 4 or 8)
* A15   ``GrossPitaevskiiGpuA15/mesh.h:8-432`` int2 offsets, degree 4
* C15   ``GrossPitaevskiiGpuC15/mesh.h:8-1242`` int2 offsets, degree 4
* Z     ``GrossPitaevskiiGpuZ/mesh.h:8-738`` int2 offsets, degree 4; the block is not cubic:
  BLOCK_WIDTH = (w, sqrt(3) w, w) (the hexagonal Z phase in an orthorhombic cell)

Usage: ``python convert_mesh_tables.py [reference_root]``
"""
import hashlib
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))

SOURCES = {
    "cube": ("Spin1GpeCube/mesh.h", "int4"),
    "bcc": ("Spin1GpeBCC/mesh.h", "int4"),
    "fcc": ("GrossPitaevskiiGpuFcc/mesh.h", "int2"),
    "a15": ("GrossPitaevskiiGpuA15/mesh.h", "int2"),
    "c15": ("GrossPitaevskiiGpuC15/mesh.h", "int2"),
    "z": ("GrossPitaevskiiGpuZ/mesh.h", "int2"),
}

NUM = r"[-+]?\d+(?:\.\d*)?(?:[eE][-+]?\d+)?"


def _define(text, name):
    m = re.search(r"#define\s+%s\s+(%s)" % (name, NUM), text)
    if not m:
        raise ValueError("missing #define " + name)
    return m.group(1)


def _positions(text, nvals):
    pos = {}
    # Vector3 form used by getPositions (Cube / FCC / A15 / C15 files)
    for m in re.finditer(r"pos\[(\d+)\]\s*=\s*Vector3\((%s),\s*(%s),\s*(%s)\)" % (NUM, NUM, NUM), text):
        pos[int(m.group(1))] = [float(m.group(2)), float(m.group(3)), float(m.group(4))]
    # brace form used by getLocalPos (BCC / Cube spin files)
    for m in re.finditer(r"pos\[(\d+)\]\s*=\s*\{\s*(%s),\s*(%s),\s*(%s)\s*\}" % (NUM, NUM, NUM), text):
        pos.setdefault(int(m.group(1)), [float(m.group(2)), float(m.group(3)), float(m.group(4))])
    for m in re.finditer(r"pos\[(\d+)\]\s*=\s*make_double3\((%s),\s*(%s),\s*(%s)\)" % (NUM, NUM, NUM), text):
        pos.setdefault(int(m.group(1)), [float(m.group(2)), float(m.group(3)), float(m.group(4))])
    if sorted(pos) != list(range(nvals)):
        raise ValueError("positions incomplete: %d of %d" % (len(pos), nvals))
    return [pos[i] for i in range(nvals)]


def _offset_expr(expr):
    """Parse a byte-offset expression such as ``-nx - ny`` into (dx, dy, dz)."""
    expr = expr.replace(" ", "")
    if expr in ("0", ""):
        return (0, 0, 0)
    d = {"nx": 0, "ny": 0, "nz": 0}
    for sign, name in re.findall(r"([-+]?)(nx|ny|nz)", expr):
        d[name] += -1 if sign == "-" else 1
    rest = re.sub(r"[-+]?(nx|ny|nz)", "", expr)
    if rest not in ("", "+0", "-0", "0"):
        raise ValueError("unparsed offset expression: " + expr)
    return (d["nx"], d["ny"], d["nz"])


def _faces(text, fmt, nind):
    ind = {}
    if fmt == "int4":
        for m in re.finditer(r"ind\[(\d+)\]\s*=\s*make_int4\(\s*(-?\d+),\s*(-?\d+),\s*(-?\d+),\s*(\d+)\s*\)", text):
            ind[int(m.group(1))] = (int(m.group(2)), int(m.group(3)), int(m.group(4)), int(m.group(5)))
    else:
        for m in re.finditer(r"ind\[(\d+)\]\s*=\s*make_int2\(\s*([^,]+),\s*(\d+)\s*\)", text):
            dx, dy, dz = _offset_expr(m.group(2))
            ind[int(m.group(1))] = (dx, dy, dz, int(m.group(3)))
    if sorted(ind) != list(range(nind)):
        raise ValueError("index table incomplete: %d of %d" % (len(ind), nind))
    hod = {}
    for m in re.finditer(r"hodges\[(\d+)\]\s*=\s*(%s)\s*;" % NUM, text):
        hod[int(m.group(1))] = float(m.group(2))
    if sorted(hod) != list(range(nind)):
        raise ValueError("hodge table incomplete: %d of %d" % (len(hod), nind))
    return [ind[i] for i in range(nind)], [hod[i] for i in range(nind)]


def _face_ranges(text, nvals, face_count, nind):
    ranges = {}
    for m in re.finditer(r"indicesAndFaceCounts\[(\d+)\]\s*=\s*make_int2\(\s*(\d+),\s*(\d+)\s*\)", text):
        ranges[int(m.group(1))] = (int(m.group(2)), int(m.group(3)))
    if ranges:
        if sorted(ranges) != list(range(nvals)):
            raise ValueError("indicesAndFaceCounts incomplete")
        return [ranges[i] for i in range(nvals)]
    if face_count * nvals != nind:
        raise ValueError("fixed degree table size mismatch")
    return [(i * face_count, face_count) for i in range(nvals)]


def convert(name, ref_root):
    rel, fmt = SOURCES[name]
    path = os.path.join(ref_root, rel)
    with open(path, "r") as f:
        text = f.read()
    nvals = int(_define(text, "VALUES_IN_BLOCK"))
    nind = int(_define(text, "INDICES_PER_BLOCK"))
    face_count = int(_define(text, "FACE_COUNT"))
    dual_edge = float(_define(text, "DUAL_EDGE_LENGTH"))
    m = re.search(r"BLOCK_WIDTH\s*=\s*Vector3\((%s),\s*(%s),\s*(%s)\)" % (NUM, NUM, NUM), text)
    block_width = [float(m.group(i)) for i in (1, 2, 3)]
    m = re.search(r"VOLUME\s*=\s*(%s)\s*;" % NUM, text)
    volume = float(m.group(1))
    m = re.search(r"return\s+(%s)\s*;\s*\}\s*(?:#endif.*)?\s*$" % NUM, text.strip() + "\n", re.S)
    ret = float(m.group(1)) if m else None

    positions = _positions(text, nvals)
    ind, hod = _faces(text, fmt, nind)
    ranges = _face_ranges(text, nvals, face_count, nind)
    nodes = []
    for v in range(nvals):
        start, count = ranges[v]
        faces = []
        for f in range(start, start + count):
            dx, dy, dz, tgt = ind[f]
            faces.append({"d": [dx, dy, dz], "target": tgt, "hodge": hod[f]})
        nodes.append({"position": positions[v], "faces": faces})
    sha = hashlib.sha256(text.encode()).hexdigest()
    return {
        "lattice": name,
        "values_in_block": nvals,
        "block_width": block_width,
        "volume": volume,
        "dual_edge_length": dual_edge,
        "max_degree": max(len(n["faces"]) for n in nodes),
        "laplacian_scale": ret,
        "nodes": nodes,
        "provenance": {
            "source": rel,
            "format": fmt,
            "sha256": sha,
            "note": "generated constants of the reference's mesh builder, converted verbatim",
        },
    }


def main():
    ref_root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    for name in SOURCES:
        table = convert(name, ref_root)
        out = os.path.join(HERE, name + ".json")
        with open(out, "w") as f:
            json.dump(table, f, indent=1)
            f.write("\n")
        print("%s: %d unknowns, max degree %d -> %s" % (name, table["values_in_block"], table["max_degree"], out))


if __name__ == "__main__":
    main()
