"""The FP64 oracle adjoint (oracle/gpe_oracle.c orc_adjoint_sweep + tests/oracle_lib.py
driver) against central finite differences of the oracle's own forward model: this pins the
transposed-Jacobian derivation (density, spin-exchange, singlet, Zeeman, kinetic terms) and
the parameter sensitivities before any GPU code is compared with it."""
import numpy as np
import pytest

from helpers import normalized, random_state
from spacetime_amd import gpe, lattice, states


def _case(oracle, ncomp, static_field):
    cfg = gpe.Config.centred_box("bcc", (3, 3, 2), 0.7, ncomp=ncomp, dt=1e-3, trap=(1.0, 0.9, 1.2), c0=80.0,
                                 c1=12.0, c2=-9.0, q=0.2)
    lat = lattice.get("bcc")
    L = oracle.lattice(lat)
    from helpers import oracle_params
    p = oracle_params(oracle, cfg)
    p.bq = static_field[0]
    p.bb[:] = list(static_field[1])
    psi0 = normalized(random_state(cfg.shape, 3), cfg)
    target = normalized(random_state(cfg.shape, 4), cfg)
    return cfg, lat, L, p, psi0, target


def _J(oracle, L, p, lat, cfg, psi0, target, objective, N):
    return oracle.adjoint(L, p, lat, cfg.dt, N, psi0, objective, target)["J"]


@pytest.mark.parametrize("ncomp,objective", [(1, 0), (1, 1), (3, 0), (3, 1), (3, 2), (5, 0), (5, 1)])
def test_oracle_adjoint_matches_fd_static_field(oracle, ncomp, objective):
    N = 14
    cfg, lat, L, p, psi0, target = _case(oracle, ncomp, (0.07, (0.03, -0.02, 0.05)))
    res = oracle.adjoint(L, p, lat, cfg.dt, N, psi0, objective, target)
    h = 1e-5
    fd_trap = []
    for a in range(3):
        w = list(p.trap)
        p.trap[a] = w[a] + h
        jp = _J(oracle, L, p, lat, cfg, psi0, target, objective, N)
        p.trap[a] = w[a] - h
        jm = _J(oracle, L, p, lat, cfg, psi0, target, objective, N)
        p.trap[a] = w[a]
        fd_trap.append((jp - jm) / (2 * h))
    scale = max(np.abs(fd_trap).max(), 1e-30)
    np.testing.assert_allclose(res["grad_trap"], fd_trap, rtol=0, atol=1e-6 * scale + 1e-11)  # FD roundoff floor
    if ncomp > 1:
        fd_b = []
        for a in range(3):
            b0 = p.bb[a]
            p.bb[a] = b0 + h
            jp = _J(oracle, L, p, lat, cfg, psi0, target, objective, N)
            p.bb[a] = b0 - h
            jm = _J(oracle, L, p, lat, cfg, psi0, target, objective, N)
            p.bb[a] = b0
            fd_b.append((jp - jm) / (2 * h))
        b0 = p.bq
        p.bq = b0 + h
        jp = _J(oracle, L, p, lat, cfg, psi0, target, objective, N)
        p.bq = b0 - h
        jm = _J(oracle, L, p, lat, cfg, psi0, target, objective, N)
        p.bq = b0
        fd_q = (jp - jm) / (2 * h)
        scale = max(scale, np.abs(fd_b).max(), abs(fd_q))
        np.testing.assert_allclose(res["grad_bias"], fd_b, rtol=0, atol=1e-6 * scale + 1e-11)
        if objective != 1:  # the energy's explicit bq term is not part of the oracle driver
            assert res["grad_bq"] == pytest.approx(fd_q, abs=1e-6 * scale)


def test_oracle_adjoint_knot_ramp(oracle):
    N = 16
    cfg, lat, L, p, psi0, target = _case(oracle, 3, (0.0, (0, 0, 0)))
    knots = np.random.default_rng(1).uniform(-0.1, 0.1, 6)
    T = N * cfg.dt
    segs = states.knot_ramp(knots, T)

    def ramp_for(kn):
        return oracle.ramp(q_segments=[(0.05, 1e30, 0)], b_segments=states.knot_ramp(kn, T))

    def phb(t):
        out = []
        for j in range(len(knots)):
            e = np.zeros(len(knots))
            e[j] = 1.0
            out.append(oracle.ramp_signal(ramp_for(e), t)[1][2])
        return np.array(out)

    res = oracle.adjoint(L, p, lat, cfg.dt, N, psi0, 0, target, ramp=ramp_for(knots),
                         ramp_bases=(lambda t: 1.0, phb))
    h = 1e-5
    fd = []
    for j in range(len(knots)):
        kp, km = knots.copy(), knots.copy()
        kp[j] += h
        km[j] -= h
        jp = oracle.adjoint(L, p, lat, cfg.dt, N, psi0, 0, target, ramp=ramp_for(kp))["J"]
        jm = oracle.adjoint(L, p, lat, cfg.dt, N, psi0, 0, target, ramp=ramp_for(km))["J"]
        fd.append((jp - jm) / (2 * h))
    fd = np.array(fd)
    np.testing.assert_allclose(res["grad_ramp_b"][:, 2], fd, rtol=0, atol=1e-6 * np.abs(fd).max())
    assert segs[0][2] == 0
