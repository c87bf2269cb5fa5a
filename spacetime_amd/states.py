"""Initial states on the DEC unknowns (host side, FP64 numpy, ABI layout (C, nz, ny, nx, V)).

* Thomas-Fermi profile sqrt(max(mu - V, 0) / g) with mu fixed by the norm -- a closed form
  owned by this project (the reference loads its states from ground_state.dat files that
  it does not ship, GPUGrossPitaevskii_BCC.cu:589-590, Spin1GpeBCC.cu:1228-1237).
* Phase imprints: straight vortex line exp(i kappa atan2(y, x)) and vortex ring
  exp(i atan2(z, rho - R)) in the z = 0 plane.
* Seeded uniform-disk noise psi *= 1 + eps (GPBlocks.cpp:517-521 uses Random(54363)
  getUniformCircle scaled by its amplitude); spinor noise psi_m += amp sqrt(n) eps_m in the
  spirit of Spin1GpeBCC.cu:1257-1292.
* Spinor projections keeping n = |psi|^2: polar / ferromagnetic / mixed
  (Spin1GpeBCC.cu:463-551) and UN / BN / cyclic (Spin2GpeBCC.cu:417-558).
"""
import ctypes
import os

import numpy as np

from . import lattice as _lattice


class _VortexParams(ctypes.Structure):
    _fields_ = [("kappa", ctypes.c_int), ("g", ctypes.c_double), ("rmin", ctypes.c_double),
                ("rmax", ctypes.c_double), ("zmax", ctypes.c_double), ("rstep", ctypes.c_double),
                ("zstep", ctypes.c_double), ("ratio", ctypes.c_double), ("maxiters", ctypes.c_int),
                ("eps", ctypes.c_double)]


def _vortex_lib():
    from . import gpe
    lib = gpe.load_library()
    if not getattr(lib, "_vortex_ready", False):
        vp = ctypes.POINTER(_VortexParams)
        dp = ctypes.POINTER(ctypes.c_double)
        lib.gpe_vortex_size.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        lib.gpe_vortex_relax.argtypes = [vp, dp, dp]
        lib.gpe_vortex_fill.argtypes = [ctypes.c_void_p, vp, dp, dp]
        lib._vortex_ready = True
    return lib


def vortex_state(kappa, g, rmax, zmax, rstep, zstep, ratio=1.0, maxiters=20000, eps=1e-29, rmin=0.0):
    """Axisymmetric stationary vortex f(r, z) and mu by the reference's relaxation
    (VortexState.cpp:42-196), computed by libgpe's host code (gpe_vortex_relax)."""
    lib = _vortex_lib()
    vp = _VortexParams(kappa, g, rmin, rmax, zmax, rstep, zstep, ratio, maxiters, eps)
    rs, zs = ctypes.c_int(), ctypes.c_int()
    n = lib.gpe_vortex_size(ctypes.byref(vp), ctypes.byref(rs), ctypes.byref(zs))
    if n <= 0:
        raise ValueError("bad vortex profile parameters")
    f = np.zeros(n)
    mu = ctypes.c_double()
    rc = lib.gpe_vortex_relax(ctypes.byref(vp), f.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                              ctypes.byref(mu))
    if rc != 0:
        raise ValueError("gpe_vortex_relax failed (%d)" % rc)
    return dict(params=vp, f=f, mu=mu.value, rsize=rs.value, zsize=zs.value)


def vortex_fill(cfg, vs):
    """psi0 = f(r, z) ((x + i y) / r)^kappa on every unknown of cfg (component 0)."""
    lib = _vortex_lib()
    out = np.zeros(cfg.shape, dtype=np.complex128)
    c = cfg.to_c()
    rc = lib.gpe_vortex_fill(ctypes.byref(c), ctypes.byref(vs["params"]),
                             vs["f"].ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                             out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    if rc != 0:
        raise ValueError("gpe_vortex_fill failed (%d)" % rc)
    return out


def positions(cfg):
    """Physical positions, shape (nz, ny, nx, V, 3): p0 + s (i * block edge + pos_v), the
    association order of GPBlocks.cpp:487."""
    lat = _lattice.get(cfg.lattice)
    nx, ny, nz = cfg.dims
    s = cfg.block_scale
    w = lat.block_dims
    out = np.empty((nz, ny, nx, lat.V, 3))
    out[..., 0] = cfg.p0[0] + s * (np.arange(nx)[None, None, :, None] * w[0] + lat.pos[None, None, None, :, 0])
    out[..., 1] = cfg.p0[1] + s * (np.arange(ny)[None, :, None, None] * w[1] + lat.pos[None, None, None, :, 1])
    out[..., 2] = cfg.p0[2] + s * (np.arange(nz)[:, None, None, None] * w[2] + lat.pos[None, None, None, :, 2])
    return out


def dual_volumes(cfg):
    lat = _lattice.get(cfg.lattice)
    return cfg.block_scale ** 3 * lat.vol


def trap_potential(cfg, r):
    w = np.asarray(cfg.trap, dtype=np.float64)
    return 0.5 * ((w[0] * r[..., 0]) ** 2 + (w[1] * r[..., 1]) ** 2 + (w[2] * r[..., 2]) ** 2) + cfg.trap_offset


def norm(cfg, psi):
    dv = dual_volumes(cfg)
    return float(np.sum(np.abs(psi) ** 2 * dv[None, None, None, None, :]))


def thomas_fermi(cfg, g, target=1.0, r=None):
    """sqrt(max(mu - V, 0) / g) on every unknown with mu chosen so that the norm is `target`."""
    if r is None:
        r = positions(cfg)
    V = trap_potential(cfg, r)
    dv = np.broadcast_to(dual_volumes(cfg), V.shape)
    lo, hi = float(V.min()), float(V.min()) + 1.0
    mass = lambda mu: float(np.sum(np.maximum(mu - V, 0.0) * dv)) / g  # noqa: E731
    while mass(hi) < target:
        hi = lo + 2 * (hi - lo)
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if mass(mid) < target:
            lo = mid
        else:
            hi = mid
    mu = 0.5 * (lo + hi)
    return np.sqrt(np.maximum(mu - V, 0.0) / g), mu


def vortex_line_phase(r, kappa=1):
    return np.exp(1j * kappa * np.arctan2(r[..., 1], r[..., 0]))


def vortex_ring_phase(r, radius):
    rho = np.hypot(r[..., 0], r[..., 1])
    return np.exp(1j * np.arctan2(r[..., 2], rho - radius))


def uniform_disk(rng, shape):
    """Uniform samples in the unit disk as complex numbers (rejection sampling)."""
    n = int(np.prod(shape))
    out = np.empty(n, dtype=np.complex128)
    filled = 0
    while filled < n:
        m = int((n - filled) * 1.3) + 16
        xy = rng.uniform(-1.0, 1.0, size=(m, 2))
        ok = xy[(xy ** 2).sum(1) <= 1.0]
        k = min(len(ok), n - filled)
        out[filled:filled + k] = ok[:k, 0] + 1j * ok[:k, 1]
        filled += k
    return out.reshape(shape)


def multiplicative_noise(psi, amplitude, seed):
    """psi * (1 + amplitude * eps), eps uniform in the unit disk (seeded)."""
    rng = np.random.default_rng(seed)
    return psi * (1.0 + amplitude * uniform_disk(rng, psi.shape))


def spinor_noise(psi, amplitude, seed):
    """psi_m += amplitude * sqrt(n) * eps_m, eps uniform in the unit disk (seeded)."""
    rng = np.random.default_rng(seed)
    n = np.sum(np.abs(psi) ** 2, axis=0)
    return psi + amplitude * np.sqrt(n)[None] * uniform_disk(rng, psi.shape)


def project(psi, phase_name, mix=0.0, rel_phase=0.0):
    """Spinor phase with the same density n: spin-1 'polar', 'ferro', 'mixed' (mix = ferro
    fraction); spin-2 'un', 'bn_hori', 'bn_vert', 'cyclic' (rel_phase on the second entry)."""
    n = np.sum(np.abs(psi) ** 2, axis=0)
    out = np.zeros_like(psi)
    C = psi.shape[0]
    ph = np.exp(1j * rel_phase)
    if C == 3:
        if phase_name == "polar":
            out[1] = np.sqrt(n)
        elif phase_name == "ferro":
            out[0] = np.sqrt(n)
        elif phase_name == "mixed":
            out[0] = np.sqrt(n * mix)
            out[1] = np.sqrt(n * (1.0 - mix))
        else:
            raise ValueError(phase_name)
    elif C == 5:
        if phase_name == "un":
            out[2] = np.sqrt(n)
        elif phase_name == "bn_hori":
            out[0] = np.sqrt(n / 2)
            out[4] = ph * np.sqrt(n / 2)
        elif phase_name == "bn_vert":
            out[1] = np.sqrt(n / 2)
            out[3] = ph * np.sqrt(n / 2)
        elif phase_name == "cyclic":
            out[0] = np.sqrt(n / 3)
            out[3] = ph * np.sqrt(2 * n / 3)
        elif phase_name == "ferro":
            out[0] = np.sqrt(n)
        else:
            raise ValueError(phase_name)
    else:
        raise ValueError("projection needs 3 or 5 components")
    return out


def knot_ramp(knots, duration):
    """Piecewise-linear bias ramp Bz(t) through equally spaced knots over `duration`, as
    gpe_ramp bias segments (first knot constant at t = 0, then linear segments)."""
    k = len(knots)
    segs = [((0.0, 0.0, float(knots[0])), 0.0, 0)]
    segs += [((0.0, 0.0, float(v)), duration / (k - 1), 1) for v in knots[1:]]
    return segs


def read_reference_dat(path, cfg):
    """Load a state file of the reference solvers into the ABI layout cfg.shape.

    The files are raw dumps of the padded array of blocks written with ofstream::write
    (Spin1GpeBCC.cu:1464-1467,1690-1699; read back at :1228-1237): (nz+2)(ny+2)(nx+2) blocks,
    block (i, j, k) at index (k+1)(ny+2)(nx+2) + (j+1)(nx+2) + (i+1), each block holding V
    unknowns of ncomp complex doubles ordered m = F..-F.  Read by libgpe's host code
    (gpe_host_read_dat)."""
    from . import gpe
    lib = gpe.load_library()
    out = np.zeros(cfg.shape, dtype=np.complex128)
    rc = lib.gpe_host_read_dat(cfg.to_c(), os.fsencode(str(path)), out.ctypes.data)
    if rc != 0:
        raise ValueError("%s: not a state file of this grid (%d)" % (path, rc))
    return out


def write_reference_dat(path, cfg, psi):
    """Write an ABI-layout state as the reference's padded block file (zero halo), by
    libgpe's host code (gpe_host_write_dat)."""
    from . import gpe
    lib = gpe.load_library()
    psi = np.ascontiguousarray(psi, dtype=np.complex128)
    if psi.shape != cfg.shape:
        raise ValueError("state shape %s, expected %s" % (psi.shape, cfg.shape))
    rc = lib.gpe_host_write_dat(cfg.to_c(), psi.ctypes.data, os.fsencode(str(path)))
    if rc != 0:
        raise OSError("cannot write %s (%d)" % (path, rc))
