"""torch.autograd.Function over the library's adjoint (cup2d_adjoint_*).

DragObjective maps the Fourier coefficients of a forced, fixed disk's
surface velocity to the time-averaged drag over a run; its backward is the
hand-written discrete adjoint on the GPU (checkpoints every `every` steps,
each segment recomputed), not torch's autodiff.  torch is imported here
only; the rest of the package does not need it."""
import numpy as np
import torch


class DragObjective(torch.autograd.Function):
    """J(theta) = (1/nsteps) sum_n drag_n.  make_sim() returns a fresh
    Simulation (reset done) holding one forced, fixed disk; dts (optional)
    fixes the time steps (the adjoint holds them fixed; finite differences
    must use the same sequence)."""

    @staticmethod
    def forward(ctx, theta, make_sim, nsteps, every, dts):
        sim = make_sim()
        sim.set_surface_velocity(theta.detach().cpu().double().numpy())
        J, used = sim.adjoint_forward(nsteps, every, dts)
        ctx.sim = sim
        ctx.dts = used
        return theta.new_tensor(J, dtype=torch.float64)

    @staticmethod
    def backward(ctx, gJ):
        g = ctx.sim.adjoint_backward()
        ctx.sim.close()
        grad = gJ * torch.from_numpy(np.ascontiguousarray(g)).to(gJ.device)
        return grad, None, None, None, None


def drag_objective(theta, make_sim, nsteps, every=20, dts=None):
    return DragObjective.apply(theta, make_sim, nsteps, every, dts)
