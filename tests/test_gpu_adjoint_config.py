"""The discrete adjoint at the configured size: spin-1 BCC 64^3 blocks (box 24, c0 = 1e4, c1 =
+500, q = 0.1, quadrupole gradient 4e-3, 16-knot Bz ramp), 4000 sweeps, checkpoints every 200,
overlap with a target made by another seeded ramp, against central finite differences of FP64
forward runs (Richardson-combined over h and h/2: knots h = 1e-2, trap frequencies h = 1e-4)
for knots 0, 5, 10, 15 and the three trap frequencies.  Errors are relative to the largest
finite difference of each group (knot gradients ~1e-8, trap gradients ~1e-6).

* FP32 context, FP32 psi and mu: the knot gradients within 1e-3.
* FP32 context with fp64_trajectory: every checked gradient within 1e-3.
The initial state and the target are rounded to FP32 before the finite differences, so both
sides solve the problem the FP32 context holds.
* FP32 psi and mu do not resolve the trap gradients (a ~1e-7 residue of cancelling per-sweep
  terms): that check is kept, skipped, with the measured reason.
"""
import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

gpe = pytest.importorskip("spacetime_amd.gpe")

N, K = 4000, 200
KNOTS_CHECKED = [0, 5, 10, 15]


def _config(precision):
    return gpe.Config.centred_box("bcc", (64, 64, 64), 24.0 / 64, ncomp=3, precision=precision, dt=1e-4,
                                  trap=(1.0, 1.0, 166.0 / 126.0), c0=1.0e4, c1=500.0, q=0.1)


@pytest.fixture(scope="module")
def problem():
    from spacetime_amd import lattice
    from spacetime_amd.adjoint import RampTrapProblem
    cfg64 = _config(gpe.FP64)
    lam3 = cfg64.trap[0] * cfg64.trap[1] * cfg64.trap[2]
    mu = 0.5 * (15.0 * cfg64.c0 * lam3 / (4.0 * np.pi)) ** 0.4
    dt_itp = min(2e-4, 1.0 / (lattice.get("bcc").kinetic_bound(cfg64.block_scale) + mu))
    knots = np.random.default_rng(2024).uniform(-0.1, 0.1, 16)
    ref_knots = np.random.default_rng(4000).uniform(-0.1, 0.1, 16)
    duration = N * cfg64.dt
    with gpe.Stepper(cfg64) as st:
        st.fill_thomas_fermi(cfg64.c0, mu, np.array([0, 1, 0], dtype=complex))
        st.itp(1000, dt_itp)
        st.project("polar")
        st.add_noise(0.005, 54363)
        st.scale(1.0 / np.sqrt(st.observables(0).norm))
        # the finite differences start from the state an FP32 context holds: the trap
        # gradients move by ~0.5% when psi0 alone is rounded to FP32
        psi0 = st.get_state(0).astype(np.complex64).astype(np.complex128)
        st.set_knot_ramp(ref_knots, duration, bq=4e-3)
        st.set_state(psi0)
        st.start(gpe.START_EULER)
        st.step(N - 1)
        target = st.get_state(0).astype(np.complex64).astype(np.complex128)
    fd = RampTrapProblem(cfg64, psi0, N, K, duration, objective=gpe.OBJ_OVERLAP, target=target, bq=4e-3)
    trap = np.array(cfg64.trap, float)

    def central(j, h):
        kp, km, tp, tm = knots.copy(), knots.copy(), trap.copy(), trap.copy()
        if j < 16:
            kp[j] += h
            km[j] -= h
        else:
            tp[j - 16] += h
            tm[j - 16] -= h
        return (fd.value(kp, tp) - fd.value(km, tm)) / (2 * h)

    g = {}
    for j in KNOTS_CHECKED + [16, 17, 18]:
        h = 1e-2 if j < 16 else 1e-4
        g[j] = (4 * central(j, h / 2) - central(j, h)) / 3
    fd.close()
    return dict(psi0=psi0, target=target, knots=knots, duration=duration, fd=g)


def _adjoint(problem, fp64_trajectory):
    from spacetime_amd.adjoint import RampTrapProblem
    cfg = _config(gpe.FP32)
    prob = RampTrapProblem(cfg, problem["psi0"], N, K, problem["duration"], objective=gpe.OBJ_OVERLAP,
                           target=problem["target"], bq=4e-3, fp64_trajectory=fp64_trajectory)
    J, gk, gt = prob.value_and_grad(problem["knots"], cfg.trap)
    prob.close()
    return gk, gt


def _errors(problem, gk, gt):
    fd = problem["fd"]
    fk = np.array([fd[j] for j in KNOTS_CHECKED])
    ft = np.array([fd[16 + a] for a in range(3)])
    ek = np.abs(gk[KNOTS_CHECKED] - fk).max() / np.abs(fk).max()
    et = np.abs(gt - ft).max() / np.abs(ft).max()
    print("knots: adjoint %s fd %s\ntrap: adjoint %s fd %s\nerrors knots %.2e trap %.2e" % (
        gk[KNOTS_CHECKED], fk, gt, ft, ek, et))
    return ek, et


def test_fp32_adjoint_knot_gradients_at_config_size(problem):
    gk, gt = _adjoint(problem, False)
    ek, _ = _errors(problem, gk, gt)
    assert ek < 1e-3


def test_fp64_trajectory_adjoint_of_fp32_context_at_config_size(problem):
    gk, gt = _adjoint(problem, True)
    ek, et = _errors(problem, gk, gt)
    assert ek < 1e-3 and et < 1e-3


@pytest.mark.skip(reason="FP32 storage of psi and mu rounds both to 6e-8 every sweep; the trap gradients "
                         "(~6e-7, a residue of per-sweep terms ~1e-2) come out 14-16% off at this size, and "
                         "on the FP64 oracle rounding only psi moves them by 19-47%, only mu by up to 2.7%: "
                         "use fp64_trajectory for them")
def test_fp32_adjoint_trap_gradients_at_config_size(problem):
    gk, gt = _adjoint(problem, False)
    _, et = _errors(problem, gk, gt)
    assert et < 1e-3
