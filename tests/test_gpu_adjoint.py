"""The hand-written adjoint of the whole step (cup2d_adjoint_*): gradient of
the time-averaged drag of a forced, fixed cylinder with respect to the 64
Fourier coefficients of its surface velocity, checked against central finite
differences of the same forward run (the time steps held at the base run's,
as the adjoint holds them).  One step exercises the drag and penalization
adjoints only (the flow before the first penalization does not depend on
the coefficients); two and more steps every operator; the full workload is
1024x512 for 200 steps with checkpoints every 20 steps."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
import bench  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[gpu case] {request.node.name}", flush=True)
    sys.stdout.flush()


def _fd_check(cup, make, theta, nsteps, every, ndirs, seed, eps=1e-6):
    g = make()
    g.set_surface_velocity(theta)
    J, dts = g.adjoint_forward(nsteps, every)
    grad = g.adjoint_backward()
    g.close()
    rng = np.random.default_rng(seed)
    worst = 0.0
    for _ in range(ndirs):
        d = rng.standard_normal(theta.size)
        Jp = Jm = None
        for sgn in (1, -1):
            h = make()
            h.set_surface_velocity(theta + sgn * eps * d)
            val, _ = h.adjoint_forward(nsteps, every, dts)
            h.close()
            if sgn > 0:
                Jp = val
            else:
                Jm = val
        fd = (Jp - Jm) / (2 * eps)
        ad = float(grad @ d)
        err = abs(fd - ad) / max(abs(fd), 1e-300)
        worst = max(worst, err)
        print(f"    dir: adjoint {ad:+.10e} fd {fd:+.10e} rel {err:.2e}", flush=True)
    return J, grad, worst


@pytest.mark.parametrize("nsteps", [1, 2, 5])
def test_adjoint_short_runs(cup, nsteps):
    """The same cylinder on 256x128 cells (8x4 blocks), one to five steps,
    checkpoints every two steps, 4 directions (coefficients 10x larger),
    central differences with eps = 1e-7, to 1e-8.

    J is piecewise smooth: the drag takes max(0, -F.u_hat) per surface point,
    the fade and the obstacle RHS correction rescale by sign, advection
    upwinds by the sign of the velocity.  A difference whose +-eps interval
    crosses one of those switches is not the derivative: at eps = 1e-6, 9 of
    24 directions over 3 to 5 steps differed from the adjoint by 1e-7 to
    2.7e-6, while at eps = 1e-7 all 24 agreed to 1.3e-9 or better and the
    forward and backward one-sided differences agreed with each other
    (measured on MI355X).  Smaller eps is dominated by round-off of J."""
    cfg = dict(bench.ADJOINT)
    theta = bench.adjoint_theta(cfg) * 10
    J, grad, worst = _fd_check(cup, lambda: bench.adjoint_sim(cup, cfg, 8, 4), theta, nsteps,
                               2, 4, 100 + nsteps, eps=1e-7)
    print(f"  {nsteps} steps: J {J:.6e}, |grad| {np.linalg.norm(grad):.3e}, worst rel {worst:.2e}",
          flush=True)
    assert worst < 1e-8


def test_adjoint_workload_1024x512(cup):
    """The adjoint workload itself: 1024x512, 200 steps, checkpoints every 20,
    8 random directions, central differences with eps = 1e-6."""
    cfg = bench.ADJOINT
    theta = bench.adjoint_theta(cfg)
    J, grad, worst = _fd_check(cup, lambda: bench.adjoint_sim(cup, cfg), theta, cfg["nsteps"],
                               cfg["every"], 8, 7)
    print(f"  J {J:.8e}, |grad| {np.linalg.norm(grad):.4e}, worst rel error over 8 "
          f"directions {worst:.2e}", flush=True)
    assert worst < 1e-6


def test_adjoint_autograd_harness(cup):
    """torch.autograd.Function over the C-ABI: backward() returns the adjoint
    gradient; the checkpoint interval does not change it."""
    torch = pytest.importorskip("torch")
    from cubismup_2d_amd.adjoint import drag_objective
    cfg = dict(bench.ADJOINT)
    theta0 = bench.adjoint_theta(cfg) * 10
    grads = []
    for every in (3, 7):
        th = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
        J = drag_objective(th, lambda: bench.adjoint_sim(cup, cfg, 8, 4), 10, every)
        J.backward()
        grads.append(th.grad.numpy().copy())
    e = np.linalg.norm(grads[0] - grads[1]) / np.linalg.norm(grads[1])
    print(f"  |grad| {np.linalg.norm(grads[1]):.4e}, checkpoint 3 vs 7: {e:.2e}", flush=True)
    assert e < 1e-12
