"""DEC unit-cell tables for the Cube, BCC, FCC, A15, C15 and Z lattices.

The tables are the reference's generated ``mesh.h`` constants, converted verbatim to
JSON by ``tables/convert_mesh_tables.py`` (sources: ``Spin1GpeCube/mesh.h``,
``Spin1GpeBCC/mesh.h``, ``GrossPitaevskiiGpuFcc/mesh.h``, ``GrossPitaevskiiGpuA15/mesh.h``,
``GrossPitaevskiiGpuC15/mesh.h``, ``GrossPitaevskiiGpuZ/mesh.h``).

Each lattice has ``V`` unknowns (dual 0-cells) per block of edges ``block_dims`` (table units;
physical length = ``block_scale`` * table length).  The blocks are cubes of edge
``block_width`` except Z's, whose edges are (w, sqrt(3) w, w) (``block_width`` = w, the x
edge).  The Laplacian of the
reference (``Spin1GpeBCC/Spin1GpeBCC.cu:827-865``) is

    (L psi)_i = sum_f hodge_f / s^2 * (psi_{t(f)} - psi_i)

over the faces f of unknown i, the neighbour living in block offset d(f) in {-1,0,1}^3.

``vol`` is the dual volume of each unknown (table units).  The reference multiplies every
integral by a single ``VOLUME`` constant, which equals the per-unknown volume only for
Cube and BCC; here the volumes are solved from the Hodge symmetry
vol_i * hodge(i->j) = vol_j * hodge(j->i) and normalised to the block volume, which keeps
the discrete Laplacian self-adjoint in the volume-weighted inner product.
"""
import json
import os

import numpy as np

LATTICES = ("cube", "bcc", "fcc", "a15", "c15", "z")
LATTICE_IDS = {name: i for i, name in enumerate(LATTICES)}
_TABLE_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tables")


class Lattice:
    """Unit-cell topology and metric of one lattice (arrays in table units)."""

    def __init__(self, name):
        if name not in LATTICES:
            raise ValueError("unknown lattice %r (expected one of %s)" % (name, ", ".join(LATTICES)))
        with open(os.path.join(_TABLE_DIR, name + ".json")) as f:
            t = json.load(f)
        self.name = name
        self.id = LATTICE_IDS[name]
        self.V = int(t["values_in_block"])
        self.block_dims = np.array(t["block_width"], dtype=np.float64)
        self.block_width = float(self.block_dims[0])
        self.ref_volume = float(t["volume"])
        self.dual_edge_length = float(t["dual_edge_length"])
        self.provenance = t["provenance"]
        nodes = t["nodes"]
        self.pos = np.array([n["position"] for n in nodes], dtype=np.float64)
        counts = [len(n["faces"]) for n in nodes]
        self.degree = np.array(counts, dtype=np.int32)
        self.max_degree = int(max(counts))
        self.face_start = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
        faces = [f for n in nodes for f in n["faces"]]
        self.nfaces = len(faces)
        self.face_d = np.array([f["d"] for f in faces], dtype=np.int32).reshape(-1, 3)
        self.face_tgt = np.array([f["target"] for f in faces], dtype=np.int32)
        self.face_hodge = np.array([f["hodge"] for f in faces], dtype=np.float64)
        self.face_src = np.repeat(np.arange(self.V, dtype=np.int32), counts)
        self.vol = self._solve_volumes()

    def _solve_volumes(self):
        lookup = {}
        for f in range(self.nfaces):
            lookup[(int(self.face_src[f]), int(self.face_tgt[f]), tuple(self.face_d[f]))] = self.face_hodge[f]
        vol = np.full(self.V, np.nan)
        vol[0] = 1.0
        changed = True
        while changed:
            changed = False
            for (i, j, d), h in lookup.items():
                back = lookup.get((j, i, tuple(-x for x in d)))
                if back is None:
                    raise ValueError("%s table is not symmetric at face %d->%d %s" % (self.name, i, j, d))
                if np.isnan(vol[j]) and not np.isnan(vol[i]):
                    vol[j] = vol[i] * h / back
                    changed = True
        if np.isnan(vol).any():
            raise ValueError("%s table is not connected" % self.name)
        return vol * (float(np.prod(self.block_dims)) / vol.sum())

    def laplacian_symbol(self, k, block_scale=1.0):
        """Bloch matrix of -1/2 L for wave vector k (physical units).

        Returns the V x V matrix D^(1/2) M(k) D^(-1/2), D = diag(vol), where
        (-1/2 L) [u_v exp(i k.r)] = [M(k) u]_v exp(i k.r) on the infinite lattice; the
        similarity transform makes it Hermitian (vol_i w_ij = vol_j w_ji).
        """
        s = block_scale
        k = np.asarray(k, dtype=np.float64)
        M = np.zeros((self.V, self.V), dtype=np.complex128)
        for f in range(self.nfaces):
            i, j = self.face_src[f], self.face_tgt[f]
            w = -0.5 * self.face_hodge[f] / (s * s)
            dr = s * (self.face_d[f] * self.block_dims + self.pos[j] - self.pos[i])
            M[i, j] += w * np.exp(1j * np.dot(k, dr))
            M[i, i] -= w
        sq = np.sqrt(self.vol)
        return (sq[:, None] * M) / sq[None, :]

    def kinetic_bound(self, block_scale=1.0):
        """Gershgorin bound on the spectrum of -1/2 L: max over unknowns of sum_f hodge_f / s^2
        (diagonal sum_f hodge_f / 2s^2 plus as much off the diagonal).  Exact for BCC (6/s^2)."""
        per = np.zeros(self.V)
        np.add.at(per, self.face_src, self.face_hodge)
        return float(per.max()) / (block_scale * block_scale)

    def packed_faces(self):
        """Faces packed as int32: (dx+1) | (dy+1) << 2 | (dz+1) << 4 | target << 8."""
        d = self.face_d + 1
        return (d[:, 0] | (d[:, 1] << 2) | (d[:, 2] << 4) | (self.face_tgt << 8)).astype(np.int32)


_CACHE = {}


def get(name):
    """Cached :class:`Lattice` by name."""
    if name not in _CACHE:
        _CACHE[name] = Lattice(name)
    return _CACHE[name]
