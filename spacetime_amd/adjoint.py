"""torch.autograd.Function over the C-ABI discrete adjoint (gpe_adjoint).

A :class:`RampTrapProblem` fixes a run (configuration, initial state, number of sweeps,
checkpoint interval, objective) whose parameters are the Bz values of a piecewise-linear
bias ramp (equally spaced knots) and the three trap frequencies.  ``GpeObjective.apply``
returns the objective as a torch scalar; its backward hands back the adjoint gradients.
The forward-only path (:meth:`RampTrapProblem.value`) is what finite differences use.
"""
import numpy as np

from . import gpe, states


class RampTrapProblem:
    def __init__(self, config, psi0, nsweeps, checkpoint_every, knot_duration, objective=gpe.OBJ_OVERLAP,
                 target=None, bq=0.0, fp64_trajectory=False):
        self.config = config
        self.psi0 = np.ascontiguousarray(psi0, dtype=np.complex128)
        self.nsweeps = int(nsweeps)
        self.checkpoint_every = int(checkpoint_every)
        self.knot_duration = float(knot_duration)
        self.objective = objective
        self.target = None if target is None else np.ascontiguousarray(target, dtype=np.complex128)
        self.bq = float(bq)
        self.fp64_trajectory = bool(fp64_trajectory)
        self.stepper = gpe.Stepper(config)
        self.last = None

    def close(self):
        self.stepper.close()

    def _prepare(self, knots, trap):
        st = self.stepper
        st.set_trap([float(t) for t in trap])
        st.set_ramp(q_segments=[(self.bq, 1e30, 0)], b_segments=states.knot_ramp(np.asarray(knots, float),
                                                                                    self.knot_duration))
        st.set_state(self.psi0)

    def value_and_grad(self, knots, trap):
        """J, dJ/dknots (Bz values), dJ/dtrap by one forward + adjoint pass."""
        self._prepare(knots, trap)
        res = self.stepper.adjoint(self.nsweeps, self.checkpoint_every, self.objective, self.target,
                                   fp64_trajectory=self.fp64_trajectory)
        self.last = res
        return res["J"], np.array(res["grad_ramp_b"][:, 2]), np.array(res["grad_trap"])

    def value(self, knots, trap):
        """J by the forward pass only (forward Euler + nsweeps-1 leapfrog sweeps)."""
        self._prepare(knots, trap)
        st = self.stepper
        st.start(gpe.START_EULER)
        st.step(self.nsweeps - 1)
        if self.objective == gpe.OBJ_OVERLAP:
            psi = st.get_state(0)
            dv = states.dual_volumes(self.config)[None, None, None, None, :]
            c = np.sum(dv * np.conj(self.target) * psi)
            return float(abs(c) ** 2)
        o = st.observables(0)
        return o.energy if self.objective == gpe.OBJ_ENERGY else float(o.magnetization[2])


try:
    import torch

    class GpeObjective(torch.autograd.Function):
        """J(knots, trap) with gradients from the discrete adjoint."""

        @staticmethod
        def forward(ctx, knots, trap, problem):
            J, gk, gt = problem.value_and_grad(knots.detach().cpu().double().numpy(),
                                               trap.detach().cpu().double().numpy())
            ctx.save_for_backward(torch.from_numpy(gk).to(knots.dtype), torch.from_numpy(gt).to(trap.dtype))
            return torch.tensor(J, dtype=torch.float64)

        @staticmethod
        def backward(ctx, grad_out):
            gk, gt = ctx.saved_tensors
            return grad_out.to(gk.dtype) * gk, grad_out.to(gt.dtype) * gt, None

except ImportError:  # torch is optional for the C ABI itself
    GpeObjective = None
