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
import numpy as np

from . import lattice as _lattice


def positions(cfg):
    """Physical positions, shape (nz, ny, nx, V, 3)."""
    lat = _lattice.get(cfg.lattice)
    nx, ny, nz = cfg.dims
    s = cfg.block_scale
    cell = s * lat.block_width
    out = np.empty((nz, ny, nx, lat.V, 3))
    out[..., 0] = cfg.p0[0] + cell * np.arange(nx)[None, None, :, None] + s * lat.pos[None, None, None, :, 0]
    out[..., 1] = cfg.p0[1] + cell * np.arange(ny)[None, :, None, None] + s * lat.pos[None, None, None, :, 1]
    out[..., 2] = cfg.p0[2] + cell * np.arange(nz)[:, None, None, None] + s * lat.pos[None, None, None, :, 2]
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
