"""Shared helpers for the GPU parity tests: mapping a product Config to oracle params,
random test states, and error norms."""
import os

import numpy as np


def gpu_available():
    """True when a HIP device is visible (without initialising HIP in this process)."""
    if os.environ.get("GPE_FORCE_NO_GPU"):
        return False
    try:
        import torch
        return torch.cuda.device_count() > 0
    except Exception:
        return os.path.exists("/dev/kfd")


def oracle_params(oracle, cfg):
    return oracle.params(cfg.ncomp, cfg.dims, cfg.block_scale, cfg.p0, trap=cfg.trap, trap_offset=cfg.trap_offset,
                         trap_off_time=cfg.trap_off_time, c0=cfg.c0, c1=cfg.c1, c2=cfg.c2, alpha=cfg.alpha,
                         quad_sign=cfg.to_c().quad_sign, q=cfg.q, center=cfg.center)


def with_field(p, bq=0.0, bb=(0, 0, 0), bq_quad=0.0, bb_quad=(0, 0, 0)):
    p.bq = bq
    p.bb[:] = list(bb)
    p.bq_quad = bq_quad
    p.bb_quad[:] = list(bb_quad)
    return p


def random_state(shape, seed, smooth=True):
    """Random complex state with a smooth envelope (keeps the stencil terms moderate)."""
    rng = np.random.default_rng(seed)
    psi = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    if smooth:
        nz, ny, nx = shape[1], shape[2], shape[3]
        z = np.sin(np.pi * (np.arange(nz) + 0.5) / nz)[:, None, None, None]
        y = np.sin(np.pi * (np.arange(ny) + 0.5) / ny)[None, :, None, None]
        x = np.sin(np.pi * (np.arange(nx) + 0.5) / nx)[None, None, :, None]
        psi = psi * 0.3 + 1.0
        psi = psi * (z * y * x)[None]
    return np.ascontiguousarray(psi.astype(np.complex128))


def normalized(psi, cfg, target=1.0):
    """psi scaled so that sum dv |psi|^2 = target (dv = s^3 vol of each unknown)."""
    from spacetime_amd import lattice
    dv = cfg.block_scale ** 3 * lattice.get(cfg.lattice).vol
    n = float(np.sum(np.abs(psi) ** 2 * dv[None, None, None, None, :]))
    return psi * np.sqrt(target / n)


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))
