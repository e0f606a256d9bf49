"""Pins of the oracle's body and force routines by identities that hold
exactly (up to round-off) whatever the discretisation, so that a plausible
slip in the restatement (a dropped term, a flipped sign, a wrong operand)
fails one of them:

- obstacle part of the pressure RHS (PressureSingle.cpp:130-213): the
  source term plus the perimeter rescale leaves the total RHS unchanged, and
  cells deep inside the body end at zero;
- integrateMomenta + Shape::updateVelocity (PressureSingle.cpp:23-67,
  Shape.cpp:25-59): a rigid-body velocity field returns its own (U, V, w);
- preventCollidingObstacles (PressureSingle.cpp:243-344): a free disk hit by
  a disk forced into it bounces with twice the closing speed (harmonic mean
  mass of a free and an infinitely heavy body), momentum exchange along the
  line of centres;
- ComputeForces (ComputeForces.cpp:50-100): a rigid rotation and a uniform
  flow have zero viscous stress, thrust - drag equals the force along the
  body velocity (the per-point positive parts of :91-93), and the power
  output of a uniform flow is U . F.
"""
import numpy as np
import pytest

import numpy_ref as R


def cell_centres(o):
    h = o.scalars()["h"]
    x = (np.arange(o.NX) + 0.5) * h
    y = (np.arange(o.NY) + 0.5) * h
    return np.meshgrid(x, y)


def test_obstacle_rhs_preserves_total(oracle_mod):
    """Two disks (one free, one towed) on 16x8 blocks; random velocity.  The
    total of the RHS over the domain is what the plain divergence gives."""
    rng = np.random.default_rng(101)
    shapes = [oracle_mod.disk(0.07, 0.31, 0.27, xvel=0.2, bFixed=True, bForced=True),
              oracle_mod.disk(0.05, 0.62, 0.21, rhoS=2.0)]
    o = oracle_mod.Oracle(16, 8, extent=1.0, nu=1e-3, poisson="freespace", shapes=shapes)
    o.reset()
    vel = 0.1 * rng.standard_normal((2, o.NY, o.NX))
    o.set_field(0, vel)
    dt = 1e-3
    h = o.scalars()["h"]
    r0 = R.pressure_rhs(vel[0], vel[1], dt, h)
    o.pressure_rhs(dt)
    r1 = o.get_field(oracle_mod.TMP)[0]
    chi = o.get_field(oracle_mod.CHI)[0]
    changed = r1 != r0
    assert changed.any()
    # the correction touches only cells of the bodies' blocks
    blk = np.zeros((o.NY // 32, o.NX // 32), bool)
    blk[np.nonzero(chi.reshape(o.NY // 32, 32, o.NX // 32, 32).max(axis=(1, 3)) > 0)] = True
    assert not changed[~np.kron(blk, np.ones((32, 32), bool))].any()
    # total preserved (each shape's source is balanced on its perimeter)
    assert abs(r1.sum() - r0.sum()) <= 1e-12 * np.abs(r0).sum()
    # deep interior (chi = 1 with its four neighbours): r = 0
    inner = chi == 1
    inner[1:-1, 1:-1] &= (chi[2:, 1:-1] == 1) & (chi[:-2, 1:-1] == 1) & \
        (chi[1:-1, 2:] == 1) & (chi[1:-1, :-2] == 1)
    assert inner.sum() > 50 and np.all(r1[inner] == 0)
    # cells with 0 < chi < 1 off the perimeter do not exist for a disk, and
    # cells with chi = 0 and no chi gradient keep the divergence
    gx = np.zeros_like(chi)
    gy = np.zeros_like(chi)
    gx[:, 1:-1] = chi[:, 2:] - chi[:, :-2]
    gy[1:-1, :] = chi[2:, :] - chi[:-2, :]
    flat = (chi == 0) & (gx * gx + gy * gy == 0)
    assert np.array_equal(r1[flat], r0[flat])


@pytest.mark.parametrize("U,V,W", [(0.13, -0.07, 0.9), (0.0, 0.2, -1.7)])
def test_momenta_of_rigid_motion(oracle_mod, U, V, W):
    """A free disk in the velocity field of its own rigid motion about its
    centre of mass: the 3x3 solve of Shape::updateVelocity gives back
    (U, V, omega)."""
    o = oracle_mod.Oracle(8, 8, extent=1.0, nu=1e-3, poisson="freespace",
                          shapes=[oracle_mod.disk(0.09, 0.47, 0.53, rhoS=1.7)])
    o.reset()
    cx, cy = o.shape()["centerOfMass"]
    x, y = cell_centres(o)
    o.set_field(0, np.stack([U - W * (y - cy), V + W * (x - cx)]))
    o.body_update(1e-3)
    s = o.shape()
    assert abs(s["u"] - U) <= 1e-13 and abs(s["v"] - V) <= 1e-13
    assert abs(s["omega"] - W) <= 1e-12 * abs(W)
    # the fluid momenta are the rigid-body integrals: UM = U PM - w PY, ...
    assert abs(s["fluidMomX"] - (U * s["penalM"] - W * s["penalDY"])) <= 1e-13 * s["penalM"]
    assert abs(s["fluidMomY"] - (V * s["penalM"] + W * s["penalDX"])) <= 1e-13 * s["penalM"]


@pytest.mark.parametrize("speed", [0.3, 0.05])
def test_collision_free_disk_hit_by_forced_disk(oracle_mod, speed):
    """Disk i (free, fluid at rest: its velocity after the momentum solve is
    zero) overlaps disk j, forced to move into it at `speed`.  The forced body
    has infinite mass, so the harmonic mean mass is 2 M_i and disk i leaves
    with velocity -2 speed along the line of centres (here the x axis, about
    which the grid is mirror symmetric); disk j keeps its imposed velocity."""
    r = 0.06
    xi, y0 = 0.40, 0.25  # y0 = 64 h is a cell face: the pair is mirror symmetric in y
    shapes = [oracle_mod.disk(r, xi, y0, rhoS=1.0),
              oracle_mod.disk(r, xi + 1.6 * r, y0, xvel=speed, bForced=True)]
    o = oracle_mod.Oracle(8, 4, extent=1.0, nu=1e-3, poisson="freespace", shapes=shapes)
    o.reset()
    o.set_field(0, np.zeros((2, o.NY, o.NX)))
    o.body_update(1e-3)
    a, b = o.shape(0), o.shape(1)
    assert abs(a["u"] - (-2 * speed)) <= 1e-12 * speed, a["u"]
    assert abs(a["v"]) <= 1e-12 * speed and abs(a["omega"]) <= 1e-10 * speed
    assert b["u"] == -speed and b["v"] == 0.0  # forcedu = -xvel (Shape.cpp:342)
    # no overlap, no impulse: the same pair apart
    shapes[1] = oracle_mod.disk(r, xi + 2.2 * r, y0, xvel=speed, bForced=True)
    o = oracle_mod.Oracle(8, 4, extent=1.0, nu=1e-3, poisson="freespace", shapes=shapes)
    o.reset()
    o.set_field(0, np.zeros((2, o.NY, o.NX)))
    o.body_update(1e-3)
    assert abs(o.shape(0)["u"]) <= 1e-15


def test_collision_two_free_disks_conserve_momentum(oracle_mod):
    """Two free disks of different density closing on each other (each
    disk's side of the domain carries its own translation).  The same state
    is updated with and without preventCollidingObstacles: the impulse
    conserves M_i u_i + M_j u_j, and with the line of centres on the x axis
    (the grid is mirror symmetric about it) each disk's change is the
    harmonic mean mass 2 M_i M_j / (M_i + M_j) times the closing speed of the
    overlap's mean velocities, divided by its own mass.  Those mean
    velocities are the bodies' translations here (omega = 0 by symmetry, no
    deformation velocity), so the closing speed is u_i - u_j."""
    r = 0.06
    xi, y0 = 0.40, 0.25
    shapes = [oracle_mod.disk(r, xi, y0, rhoS=1.0), oracle_mod.disk(r, xi + 1.6 * r, y0, rhoS=3.0)]
    out = []
    for collide in (False, True):
        o = oracle_mod.Oracle(8, 4, extent=1.0, nu=1e-3, poisson="freespace", shapes=shapes)
        o.set_collisions(collide)
        o.reset()
        x, y = cell_centres(o)
        u = np.where(x < xi + 0.8 * r, 0.2, -0.1)
        o.set_field(0, np.stack([u, np.zeros_like(u)]))
        o.body_update(1e-3)
        out.append((o.shape(0), o.shape(1)))
    (a0, b0), (a1, b1) = out
    Mi, Mj = a1["M"], b1["M"]
    assert a0["M"] == Mi and b0["M"] == Mj
    ui0, uj0 = a0["u"], b0["u"]
    assert ui0 > uj0  # closing
    meanM = 2 / (1 / Mi + 1 / Mj)
    close = ui0 - uj0
    assert abs((Mi * a1["u"] + Mj * b1["u"]) - (Mi * ui0 + Mj * uj0)) <= 1e-14 * Mi
    assert abs((a1["u"] - ui0) - (-meanM * close / Mi)) <= 1e-12 * close
    assert abs((b1["u"] - uj0) - (meanM * close / Mj)) <= 1e-12 * close
    assert abs(a1["v"]) <= 1e-13 and abs(b1["v"]) <= 1e-13
    # apart: the two updates agree
    shapes[1] = oracle_mod.disk(r, xi + 2.3 * r, y0, rhoS=3.0)
    o = oracle_mod.Oracle(8, 4, extent=1.0, nu=1e-3, poisson="freespace", shapes=shapes)
    o.reset()
    o.set_field(0, np.stack([u, np.zeros_like(u)]))
    o.body_update(1e-3)
    o2 = oracle_mod.Oracle(8, 4, extent=1.0, nu=1e-3, poisson="freespace", shapes=shapes)
    o2.set_collisions(False)
    o2.reset()
    o2.set_field(0, np.stack([u, np.zeros_like(u)]))
    o2.body_update(1e-3)
    assert o.shape(0)["u"] == o2.shape(0)["u"] and o.shape(1)["u"] == o2.shape(1)["u"]


def _forced_disk(oracle_mod, xvel=0.2, yvel=-0.05):
    o = oracle_mod.Oracle(8, 8, extent=1.0, nu=3e-3, poisson="freespace",
                          shapes=[oracle_mod.disk(0.11, 0.46, 0.52, xvel=xvel, yvel=yvel,
                                                  bForced=True)])
    o.reset()
    return o


def test_forces_rigid_rotation_and_uniform_flow_have_no_viscous_stress(oracle_mod):
    rng = np.random.default_rng(103)
    o = _forced_disk(oracle_mod)
    cx, cy = o.shape()["centerOfMass"]
    x, y = cell_centres(o)
    p = rng.standard_normal((1, o.NY, o.NX))
    o.set_field(oracle_mod.PRES, p)
    for vel in (np.stack([-1.3 * (y - cy), 1.3 * (x - cx)]),
                np.stack([np.full_like(x, 0.4), np.full_like(x, -0.25)])):
        o.set_field(0, vel)
        o.compute_forces()
        s = o.shape()
        scale = abs(s["forcex_P"]) + abs(s["forcey_P"])
        for k in ["forcex_V", "forcey_V"]:
            assert abs(s[k]) <= 1e-13 * scale, (k, s[k])
        assert abs(s["torque_V"]) <= 1e-13 * scale
        assert s["forcex"] == s["forcex_P"] + s["forcex_V"] or \
            abs(s["forcex"] - s["forcex_P"]) <= 1e-13 * scale
    # a shear field has a viscous stress: the zero above is not vacuous
    o.set_field(0, np.stack([0.7 * y, np.zeros_like(y)]))
    o.compute_forces()
    assert abs(o.shape()["forcex_V"]) + abs(o.shape()["forcey_V"]) > 0


def test_forces_thrust_drag_and_power(oracle_mod):
    """thrust - drag is the total force along the body's unit velocity; with
    a uniform flow (U0, V0) the power output is U0 Fx + V0 Fy."""
    rng = np.random.default_rng(104)
    o = _forced_disk(oracle_mod, xvel=0.2, yvel=-0.05)
    s0 = o.shape()
    us, vs = s0["u"], s0["v"]
    nrm = np.hypot(us, vs)
    o.set_field(oracle_mod.PRES, rng.standard_normal((1, o.NY, o.NX)))
    o.set_field(0, 0.2 * rng.standard_normal((2, o.NY, o.NX)))
    o.compute_forces()
    s = o.shape()
    fpar = (s["forcex"] * us + s["forcey"] * vs) / nrm
    assert s["thrust"] >= 0 and s["drag"] >= 0
    assert abs((s["thrust"] - s["drag"]) - fpar) <= 1e-13 * (s["thrust"] + s["drag"])
    assert abs(s["Pthrust"] - s["thrust"] * nrm) <= 1e-15 * s["thrust"] * nrm
    U0, V0 = 0.31, -0.12
    o.set_field(0, np.stack([np.full((o.NY, o.NX), U0), np.full((o.NY, o.NX), V0)]))
    o.compute_forces()
    s = o.shape()
    assert abs(s["Pout"] - (U0 * s["forcex"] + V0 * s["forcey"])) <= \
        1e-13 * (abs(U0 * s["forcex"]) + abs(V0 * s["forcey"]))
    assert s["defPower"] == 0.0  # disks carry no deformation velocity
