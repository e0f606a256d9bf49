"""GPU parity of the branches the plain whole-step tests do not reach: the
non-finite check, a swimmer with the cosine solver (advDiff on the side
stream feeding the separate RHS pass), collisions, implicit penalization,
every step's deferred ComputeForces in the device loop, and the divergence
left after the projection.  Same bar as the other parity tests: velocity
and pressure within 1e-10 relative L2 of the oracle, forces within 1e-8."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[gpu case] {request.node.name}", flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def divergence(vel, h):
    """(u_E - u_W + v_N - v_S) / 2h with zero-gradient faces, as the
    pressure RHS forms it (PressureSingle.cpp:112-128)."""
    u, v = vel
    up = np.pad(u, 1, mode="edge")
    vp = np.pad(v, 1, mode="edge")
    return ((up[1:-1, 2:] - up[1:-1, :-2]) + (vp[2:, 1:-1] - vp[:-2, 1:-1])) / (2 * h)


def test_nonfinite_velocity_is_reported(cup):
    """A NaN or Inf cell reaches findMaxU: calcMaxTimestep and the device loop
    report CUP2D_ERR_NONFINITE (the reference's Checker aborts,
    Helpers.cpp:178-288)."""
    g = cup.Simulation(4, 2, extent=1.0, nu=1e-3,
                       shapes=[cup.disk(0.05, 0.3, 0.25, xvel=0.1, bFixed=True, bForced=True)])
    g.reset()
    vel = np.zeros((2, g.NY, g.NX))
    vel[0, 17, 40] = np.nan
    g.set_field(cup.VEL, vel)
    with pytest.raises(cup.Cup2dError) as e:
        g.calc_max_timestep()
    assert e.value.code == cup.ERR_NONFINITE
    g.reset()  # clears the sticky error
    g.calc_max_timestep()
    vel[0, 17, 40] = 0.0
    vel[1, 3, 100] = np.inf
    g.set_field(cup.VEL, vel)
    with pytest.raises(cup.Cup2dError) as e:
        g.simulate(2)
    assert e.value.code == cup.ERR_NONFINITE
    g.close()


@pytest.mark.parametrize("device_loop", [False, True])
def test_swimmer_with_cosine_solver(cup, oracle_mod, device_loop):
    """A swimmer (advDiff runs on the side stream beside the placement chain)
    with the cosine solver, whose RHS is a separate pass that must wait for
    advDiff's output."""
    fish = dict(xpos=0.5, ypos=0.25, L=0.2, T=1.0)
    kw = dict(extent=1.0, nu=1e-4, CFL=0.2, poisson="cosine")
    g = cup.Simulation(16, 8, shapes=[cup.carlingfish(**fish)], **kw)
    o = oracle_mod.Oracle(16, 8, shapes=[oracle_mod.carlingfish(**fish)], **kw)
    g.reset()
    o.reset()
    rng = np.random.default_rng(51)
    vel = 0.02 * rng.standard_normal((2, g.NY, g.NX))
    g.set_field(0, vel)
    o.set_field(0, vel)
    if device_loop:
        g.simulate(10)
        o.simulate(10)
    else:
        for _ in range(6):
            dg, do = g.calc_max_timestep(), o.calc_max_timestep()
            assert abs(dg - do) <= 1e-13 * do
            g.advance(dg)
            o.advance(do)
    ev, ep = rel(g.get_field(0), o.get_field(0)), rel(g.get_field(2), o.get_field(2))
    print(f"  rel L2 vel {ev:.3e} pres {ep:.3e}", flush=True)
    assert ev < 1e-10 and ep < 1e-10
    a, b = g.shape(), o.shape()
    assert abs(a["drag"] - b["drag"]) <= 1e-8 * max(abs(b["drag"]), abs(b["forcex"]))


def _collision_pair(mod, r, xi, y0, dx, speed):
    return [mod.disk(r, xi, y0, rhoS=1.0), mod.disk(r, xi + dx, y0, xvel=speed, bForced=True)]


def test_collision_body_update(cup, oracle_mod):
    """The body update on two overlapping disks, one forced into the other
    (the state the oracle's collision pin uses): the free disk bounces with
    twice the closing speed, on the device as in the oracle."""
    r, xi, y0, speed = 0.06, 0.40, 0.25, 0.3
    g = cup.Simulation(8, 4, extent=1.0, nu=1e-3, poisson="freespace",
                       shapes=_collision_pair(cup, r, xi, y0, 1.6 * r, speed))
    o = oracle_mod.Oracle(8, 4, extent=1.0, nu=1e-3, poisson="freespace",
                          shapes=_collision_pair(oracle_mod, r, xi, y0, 1.6 * r, speed))
    for sim in (g, o):
        sim.reset()
        sim.set_field(0, np.zeros((2, sim.NY, sim.NX)))
        sim.body_update(1e-3)
    a, b = g.shape(0), o.shape(0)
    assert abs(a["u"] - (-2 * speed)) <= 1e-12 * speed, a["u"]
    # omega is zero up to round-off; its scale is speed / r
    for key, scale in (("u", speed), ("v", speed), ("omega", speed / r)):
        assert abs(a[key] - b[key]) <= 1e-12 * scale, (key, a[key], b[key])


def test_collision_whole_steps(cup, oracle_mod):
    """A forced disk driven into a free one: the bodies come into contact
    within the run, the collision impulse fires (the free disk ends up moving
    away, in -x), on the device as in the oracle."""
    r, xi, y0, speed = 0.05, 0.40, 0.25, 0.5
    kw = dict(extent=1.0, nu=1e-3, CFL=0.2, poisson="freespace")
    g = cup.Simulation(8, 4, shapes=_collision_pair(cup, r, xi, y0, 2.08 * r, speed), **kw)
    o = oracle_mod.Oracle(8, 4, fs_pad2=True,
                          shapes=_collision_pair(oracle_mod, r, xi, y0, 2.08 * r, speed), **kw)
    g.reset()
    o.reset()
    for sim in (g, o):  # past the CFL ramp (Simulation.cpp:264-269): full-size steps
        sim.set_scalars(time=0.0, step=100)
    hit = False
    for step in range(40):
        dg, do = g.calc_max_timestep(), o.calc_max_timestep()
        assert abs(dg - do) <= 1e-12 * do
        g.advance(dg)
        o.advance(do)
        if o.shape(0)["u"] < -0.2:
            hit = True
            break
    assert hit, "the disks never collided"
    print(f"  collision at step {step}, u_free {o.shape(0)['u']:.4f}", flush=True)
    for _ in range(3):  # and a few steps after it
        dg, do = g.calc_max_timestep(), o.calc_max_timestep()
        g.advance(dg)
        o.advance(do)
    ev, ep = rel(g.get_field(0), o.get_field(0)), rel(g.get_field(2), o.get_field(2))
    print(f"  rel L2 vel {ev:.3e} pres {ep:.3e}", flush=True)
    assert ev < 1e-10 and ep < 1e-10
    for k in range(2):
        a, b = g.shape(k), o.shape(k)
        # omega is round-off for these head-on disks: its scale is speed / r
        for key, scale in (("u", speed), ("v", speed), ("omega", speed / r)):
            assert abs(a[key] - b[key]) <= 1e-10 * max(abs(b[key]), scale), (k, key, a[key], b[key])


@pytest.mark.parametrize("device_loop", [False, True])
def test_implicit_penalization(cup, oracle_mod, device_loop):
    """The implicit Brinkman factor alpha = 1/(1 + lambda dt chi) with
    lambda = 1e6 (PressureSingle.cpp:96-99 without EXPL_INTEGRATE_MOM, and
    the lambda-weighted momenta of :45-53), on the towed cylinder and a free
    disk."""
    shapes = [dict(radius=0.05, xpos=0.25, ypos=0.25, xvel=0.2, bFixed=True, bForced=True),
              dict(radius=0.04, xpos=0.6, ypos=0.3, rhoS=1.3)]
    kw = dict(extent=1.0, nu=2e-5, CFL=0.1, poisson="freespace", implicit_penal=True, lam=1e6)
    g = cup.Simulation(16, 8, shapes=[cup.disk(**s) for s in shapes], **kw)
    o = oracle_mod.Oracle(16, 8, fs_pad2=True, shapes=[oracle_mod.disk(**s) for s in shapes],
                          **kw)
    g.reset()
    o.reset()
    n = 30
    if device_loop:
        g.simulate(n)
        o.simulate(n)
    else:
        for _ in range(n):
            dg, do = g.calc_max_timestep(), o.calc_max_timestep()
            g.advance(dg)
            o.advance(do)
    ev, ep = rel(g.get_field(0), o.get_field(0)), rel(g.get_field(2), o.get_field(2))
    print(f"  rel L2 vel {ev:.3e} pres {ep:.3e}", flush=True)
    assert ev < 1e-10 and ep < 1e-10
    for k in range(2):
        a, b = g.shape(k), o.shape(k)
        assert abs(a["drag"] - b["drag"]) <= 1e-8 * max(abs(b["drag"]), abs(b["forcex"]), 1e-12)
        for key in ["u", "v", "omega"]:
            assert abs(a[key] - b[key]) <= 1e-10 * max(abs(b[key]), 1e-3), (k, key)


def _force_log_vs_oracle(g, o, nsteps, nshapes):
    """Every step's ComputeForces of the device loop (deferred behind the next
    step's list builder) against the oracle's forces of the same step."""
    g.force_log_enable(nsteps + 4)
    g.simulate(nsteps)
    ref = {k: [] for k in range(nshapes)}
    for _ in range(nsteps):
        t = o.scalars()["time"]
        o.advance(o.calc_max_timestep())
        for k in range(nshapes):
            s = o.shape(k)
            ref[k].append((t, s["forcex"], s["forcey"], s["drag"], s["thrust"], s["torque"],
                           s["Pout"]))
    cols = [cup_col("time"), cup_col("Fx"), cup_col("Fy"), cup_col("drag"),
            cup_col("thrust"), cup_col("tau"), cup_col("Pout")]
    worst = 0.0
    for k in range(nshapes):
        rec = g.force_log(k)
        assert rec.shape[0] == nsteps, rec.shape
        assert np.array_equal(rec[:, 1], np.arange(nsteps))  # step index of each record
        r = np.array(ref[k])
        got = rec[:, cols]
        assert np.allclose(got[:, 0], r[:, 0], rtol=1e-12, atol=0)
        scale = np.maximum(np.abs(r[:, 1:3]).max(axis=1), 1e-12)
        err = np.abs(got[:, 1:] - r[:, 1:]).max(axis=1) / scale
        worst = max(worst, err.max())
    print(f"  {nsteps} steps: worst force error {worst:.2e} (relative to |F|)", flush=True)
    assert worst <= 1e-8


def cup_col(name):
    import cubismup_2d_amd as cup
    return cup.FORCE_LOG_COLUMNS.index(name)


def test_force_log_every_step_disks(cup, oracle_mod):
    """Towed cylinder and a free disk (disk layout: ComputeForces deferred to
    the side stream behind the next step's placement chain)."""
    shapes = [dict(radius=0.05, xpos=0.25, ypos=0.25, xvel=0.2, bFixed=True, bForced=True),
              dict(radius=0.04, xpos=0.6, ypos=0.3, rhoS=1.3)]
    kw = dict(extent=1.0, nu=2e-5, CFL=0.1, poisson="freespace")
    g = cup.Simulation(16, 8, shapes=[cup.disk(**s) for s in shapes], **kw)
    o = oracle_mod.Oracle(16, 8, fs_pad2=True, shapes=[oracle_mod.disk(**s) for s in shapes],
                          **kw)
    g.reset()
    o.reset()
    _force_log_vs_oracle(g, o, 40, 2)


def test_force_log_every_step_swimmer(cup, oracle_mod):
    """A swimmer and a towed disk (swimmer layout: ComputeForces deferred to
    the midline stream)."""
    fish = dict(xpos=0.9, ypos=0.5, angle=8.0, L=0.2, T=1.0, envelope="quadratic",
                profile="naca", phi=0.25)
    disk = dict(radius=0.05, xpos=1.5, ypos=0.4, xvel=0.1, bFixed=True, bForced=True)
    kw = dict(extent=2.0, nu=1e-4, CFL=0.2, poisson="periodic")
    g = cup.Simulation(32, 16, shapes=[cup.carlingfish(**fish), cup.disk(**disk)], **kw)
    o = oracle_mod.Oracle(32, 16, shapes=[oracle_mod.carlingfish(**fish), oracle_mod.disk(**disk)],
                          **kw)
    g.reset()
    o.reset()
    _force_log_vs_oracle(g, o, 25, 2)


@pytest.mark.parametrize("poisson", ["periodic", "freespace", "cosine"])
def test_divergence_after_projection(cup, oracle_mod, poisson):
    """The divergence the projection leaves is the oracle's: same field to
    1e-9 relative L2.  Central differences of a spectrally projected field are
    not exactly divergence-free (the wide-stencil divergence of the wide
    gradient is not the spectral Laplacian, and the penalization acts after
    the projection), so the initial field is smooth; away from the body and
    the domain faces its divergence must drop by two orders of magnitude."""
    shapes = [dict(radius=0.05, xpos=0.3, ypos=0.25, xvel=0.2, bFixed=True, bForced=True)]
    kw = dict(extent=1.0, nu=1e-4, CFL=0.2, poisson=poisson)
    g = cup.Simulation(16, 8, shapes=[cup.disk(**s) for s in shapes], **kw)
    o = oracle_mod.Oracle(16, 8, fs_pad2=True, shapes=[oracle_mod.disk(**s) for s in shapes],
                          **kw)
    g.reset()
    o.reset()
    h = g.scalars()["h"]
    X, Y = np.meshgrid((np.arange(g.NX) + 0.5) * h, (np.arange(g.NY) + 0.5) * h)
    vel = np.stack([0.05 * np.sin(2 * np.pi * X) * np.cos(2 * np.pi * Y)
                    + 0.03 * np.cos(4 * np.pi * Y),
                    0.05 * np.cos(2 * np.pi * X) * np.sin(4 * np.pi * Y)])
    g.set_field(0, vel)
    o.set_field(0, vel)
    g.simulate(5)
    o.simulate(5)
    dg, do = divergence(g.get_field(0), h), divergence(o.get_field(0), h)
    d0 = divergence(vel, h)
    e = rel(dg, do)
    far = (X - 0.3) ** 2 + (Y - 0.25) ** 2 > 0.15 ** 2
    inner = np.zeros_like(far)
    inner[8:-8, 8:-8] = True
    far &= inner
    drop = np.linalg.norm(do[far]) / np.linalg.norm(d0[far])
    print(f"  |div| {np.linalg.norm(do):.4e} (initial {np.linalg.norm(d0):.4e}), "
          f"away from body and faces x{drop:.2e}, rel L2 vs oracle {e:.2e}", flush=True)
    assert e < 1e-9
    assert drop < 1e-2  # the projection did its work


@pytest.mark.parametrize("device_loop", [False, True])
def test_surface_velocity_vs_oracle(cup, oracle_mod, device_loop):
    """The Fourier-mode surface velocity (16 modes, normal and tangential,
    coefficients N(0, 1e-2)) on a towed disk and on a free disk: it enters
    the penalization target, the free disk's momentum integrals and the
    deformation power, on the device as in the oracle."""
    rng = np.random.default_rng(91)
    shapes = [dict(radius=0.05, xpos=0.25, ypos=0.25, xvel=0.2, bFixed=True, bForced=True),
              dict(radius=0.04, xpos=0.6, ypos=0.3, rhoS=1.3)]
    kw = dict(extent=1.0, nu=1e-4, CFL=0.2, poisson="freespace")
    g = cup.Simulation(16, 8, shapes=[cup.disk(**s) for s in shapes], **kw)
    o = oracle_mod.Oracle(16, 8, fs_pad2=True, shapes=[oracle_mod.disk(**s) for s in shapes],
                          **kw)
    g.reset()
    o.reset()
    # one shape carries a surface velocity at a time: the free disk, so the
    # momentum integrals see it
    coef = rng.normal(0.0, 1e-2, 64)
    g.set_surface_velocity(coef, shape=1)
    o.set_surface_velocity(coef, shape=1)
    n = 12
    if device_loop:
        g.simulate(n)
        o.simulate(n)
    else:
        for _ in range(n):
            dg, do = g.calc_max_timestep(), o.calc_max_timestep()
            g.advance(dg)
            o.advance(do)
    ev, ep = rel(g.get_field(0), o.get_field(0)), rel(g.get_field(2), o.get_field(2))
    a, b = g.shape(1), o.shape(1)
    print(f"  rel L2 vel {ev:.3e} pres {ep:.3e}; free disk u {a['u']:.6e} vs {b['u']:.6e}",
          flush=True)
    assert ev < 1e-10 and ep < 1e-10
    for key in ["u", "v", "omega"]:
        assert abs(a[key] - b[key]) <= 1e-10 * max(abs(b[key]), 1e-3), key
    for key in ["drag", "defPower"]:
        assert abs(a[key] - b[key]) <= 1e-8 * max(abs(b[key]), 1e-6), key
