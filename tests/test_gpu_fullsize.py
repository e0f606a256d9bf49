"""Parity at the configurations' own sizes and lengths: 4096x2048 cells
(128x64 blocks of 32x32, periodic, the workloads bench.py measures) for 200
steps, the 512x256 cylinder for 500 steps, the 8192x4096 scale-up grid for
100 steps.  The oracle runs these in seconds, so the whole runs are
compared directly (fields, body state, every step's forces), and the FFT row
and column lengths of these sizes are pinned by closed forms."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
import bench  # noqa: E402  (the workload definitions and initial condition)

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[gpu case] {request.node.name}", flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _numpy_periodic_solve(r):
    """FFTW_periodic.h:77-92 restated with numpy's FFT: c^ *= -1/(kx~^2+ky~^2),
    k~ = 2 pi min(i, N-i)/N (index units), DC = 0."""
    NY, NX = r.shape
    R = np.fft.rfft2(r)
    kx = 2 * np.pi * np.minimum(np.arange(NX // 2 + 1), NX - np.arange(NX // 2 + 1)) / NX
    ky = 2 * np.pi * np.minimum(np.arange(NY), NY - np.arange(NY)) / NY
    lam = ky[:, None] ** 2 + kx[None, :] ** 2
    lam[0, 0] = 1
    P = -R / lam
    P[0, 0] = 0
    return np.fft.irfft2(P, s=(NY, NX))


def test_fullsize_poisson_fourier_modes(cup):
    """cos modes are eigenvectors of the periodic solve (FFTW_periodic.h:
    77-92): p = -r / (kx~^2 + ky~^2) in index units.  The closed form holds to
    the round-off of r itself, amplified by lambda / lambda_min through the
    lowest modes (numpy's FFT on the same r shows the same); the GPU solve
    matches numpy's to the pressure tolerance, 1e-10 relative L2 (observed
    <= 1e-12)."""
    g = cup.Simulation(128, 64, extent=2.0, poisson="periodic")
    assert (g.NX, g.NY) == (4096, 2048)
    x, y = np.meshgrid(np.arange(g.NX), np.arange(g.NY))
    for kx, ky in [(1, 0), (0, 1), (g.NX // 2, g.NY // 2), (1234, 567), (g.NX - 3, 1000)]:
        # phase reduced exactly in integers, so r is a Fourier mode to round-off
        # (cos of a large unreduced argument carries a phase error that grows
        # with x and y, which the solve amplifies through the lowest modes)
        ph = ((kx * x) % g.NX) * g.NY + ((ky * y) % g.NY) * g.NX
        r = np.cos(2 * np.pi * ph / (g.NX * g.NY))
        kxt = 2 * np.pi * min(kx, g.NX - kx) / g.NX
        kyt = 2 * np.pi * min(ky, g.NY - ky) / g.NY
        lam = kxt * kxt + kyt * kyt
        p = g.poisson_dense(r)
        err = np.abs(p + r / lam).max() * lam
        dnp = rel(p, _numpy_periodic_solve(r))
        print(f"  mode ({kx},{ky}) closed form {err:.2e}, vs numpy {dnp:.2e}", flush=True)
        assert err < 1e-10
        assert dnp < 1e-10


def test_fullsize_advdiff_bitexact(cup, oracle_mod):
    """advDiff over the full grid (tile numbering, XCD remap, strip loads)
    is bit-identical to the oracle away from the flux-corrected boundary."""
    rng = np.random.default_rng(41)
    kw = dict(extent=2.0, nu=1e-5)
    g = cup.Simulation(128, 64, **kw)
    o = oracle_mod.Oracle(128, 64, **kw)
    vel = 0.3 * rng.standard_normal((2, g.NY, g.NX))
    for sim in (g, o):
        sim.set_field(0, vel)
        sim.set_scalars(uinf=(0.05, -0.02))
        sim.advdiff(2.5e-5)
    a, b = g.get_field(cup.VEL), o.get_field(oracle_mod.VEL)
    inner = (slice(None), slice(1, -1), slice(1, -1))
    assert np.array_equal(a[inner], b[inner])
    assert np.abs(a - b).max() < 1e-15


def _force_log_check(g, o_records, nshapes, tol=1e-8):
    """The device loop's per-step force records against the oracle's."""
    import cubismup_2d_amd as cup
    cols = [cup.FORCE_LOG_COLUMNS.index(c) for c in ("Fx", "Fy", "drag", "thrust", "tau")]
    worst = 0.0
    for k in range(nshapes):
        rec = g.force_log(k)
        ref = np.array(o_records[k])
        assert rec.shape[0] == ref.shape[0]
        scale = np.maximum(np.abs(ref[:, :2]).max(axis=1), 1e-12)
        worst = max(worst, (np.abs(rec[:, cols] - ref).max(axis=1) / scale).max())
    print(f"  worst per-step force error {worst:.2e} over {ref.shape[0]} steps", flush=True)
    assert worst <= tol


def _oracle_steps(o, nsteps, nshapes):
    recs = {k: [] for k in range(nshapes)}
    for _ in range(nsteps):
        o.advance(o.calc_max_timestep())
        for k in range(nshapes):
            s = o.shape(k)
            recs[k].append((s["forcex"], s["forcey"], s["drag"], s["thrust"], s["torque"]))
    return recs


@pytest.mark.parametrize("config", ["periodic4096", "periodic4096_disk"])
def test_fullsize_bench_workload_vs_oracle(cup, oracle_mod, config):
    """The benchmarked workload itself (same seeded initial condition and
    projection as bench.py) for its full 200 steps on the device loop against
    the oracle: velocity and pressure to 1e-10 relative L2, body velocity and
    position, and every step's forces (the device loop's deferred
    ComputeForces) to 1e-8.  The run crosses step 100, where the CFL ramp
    ends."""
    cfg = bench.CONFIGS[config]
    kw = dict(extent=cfg["extent"], nu=cfg["nu"], CFL=0.1, poisson=cfg["poisson"])
    g = cup.Simulation(cfg["bpdx"], cfg["bpdy"], shapes=[bench.make_shape(cup, cfg["body"])],
                       **kw)
    o = oracle_mod.Oracle(cfg["bpdx"], cfg["bpdy"],
                          shapes=[bench.make_shape(oracle_mod, cfg["body"])], **kw)
    vel = bench.random_vorticity_velocity(g.NX, g.NY, g.scalars()["h"])
    bench.setup(g, vel)
    bench.setup(o, vel)
    assert rel(g.get_field(0), o.get_field(0)) < 1e-12
    n = 200
    g.force_log_enable(n)
    g.simulate(n)
    recs = _oracle_steps(o, n, 1)
    sg, so = g.scalars(), o.scalars()
    assert sg["step"] == so["step"] == n
    assert abs(sg["time"] - so["time"]) <= 1e-12 * so["time"]
    assert abs(sg["dt"] - so["dt"]) <= 1e-12 * so["dt"]
    ev = rel(g.get_field(0), o.get_field(0))
    ep = rel(g.get_field(2), o.get_field(2))
    print(f"  {n} steps: rel L2 vel {ev:.3e} pres {ep:.3e}", flush=True)
    assert ev < 1e-10 and ep < 1e-10
    a, b = g.shape(0), o.shape(0)
    for key in ["u", "v", "omega"]:
        assert abs(a[key] - b[key]) <= 1e-10 * max(abs(b[key]), 1e-3), (key, a[key], b[key])
    assert np.allclose(a["centerOfMass"], b["centerOfMass"], rtol=1e-12, atol=1e-14)
    _force_log_check(g, recs, 1)


@pytest.mark.parametrize("implicit", [False, True])
def test_cylinder512_500_steps(cup, oracle_mod, implicit):
    """The impulsively started cylinder at 512x256 (free-space, r = 0.05 towed
    at 0.2, nu = 2e-5, CFL 0.1, zero initial velocity) for its full 500
    steps, explicit penalization and implicit with lambda = 1e6: fields,
    body state and every step's drag against the oracle."""
    cfg = bench.CONFIGS["cylinder512"]
    kw = dict(extent=cfg["extent"], nu=cfg["nu"], CFL=0.1, poisson=cfg["poisson"])
    if implicit:
        kw.update(implicit_penal=True, lam=1e6)
    g = cup.Simulation(cfg["bpdx"], cfg["bpdy"], shapes=[bench.make_shape(cup, cfg["body"])],
                       **kw)
    o = oracle_mod.Oracle(cfg["bpdx"], cfg["bpdy"], fs_pad2=True,
                          shapes=[bench.make_shape(oracle_mod, cfg["body"])], **kw)
    g.reset()
    o.reset()
    n = 500
    g.force_log_enable(n)
    g.simulate(n)
    recs = _oracle_steps(o, n, 1)
    ev = rel(g.get_field(0), o.get_field(0))
    ep = rel(g.get_field(2), o.get_field(2))
    print(f"  {n} steps: rel L2 vel {ev:.3e} pres {ep:.3e}, drag {o.shape()['drag']:.6e}",
          flush=True)
    assert ev < 1e-10 and ep < 1e-10
    _force_log_check(g, recs, 1)


def test_freespace_4096x2048_vs_oracle(cup, oracle_mod):
    """Free-space solve at 4096x2048 (padded rows of 8192, columns of 4096)
    against the oracle's 2N x 2M Hockney solve (FFTW_freespace.cpp:78-138)."""
    rng = np.random.default_rng(19)
    g = cup.Simulation(128, 64, extent=2.0, poisson="freespace")
    o = oracle_mod.Oracle(128, 64, extent=2.0, poisson="freespace", fs_pad2=True)
    r = rng.standard_normal((g.NY, g.NX))
    e = rel(g.poisson_dense(r), o.poisson_dense(r))
    print(f"  rel L2 {e:.3e}", flush=True)
    assert e < 1e-12


def _steps_vs_oracle(cup, oracle_mod, bp, shapes_of, nsteps, **kw):
    g = cup.Simulation(*bp, shapes=shapes_of(cup), **kw)
    o = oracle_mod.Oracle(*bp, shapes=shapes_of(oracle_mod), **kw)
    vel = bench.random_vorticity_velocity(g.NX, g.NY, g.scalars()["h"])
    bench.setup(g, vel)
    bench.setup(o, vel)
    del vel
    nshapes = len(shapes_of(cup))
    g.force_log_enable(nsteps)
    g.simulate(nsteps)
    recs = _oracle_steps(o, nsteps, nshapes)
    ev = rel(g.get_field(0), o.get_field(0))
    ep = rel(g.get_field(2), o.get_field(2))
    print(f"  {g.NX}x{g.NY}: rel L2 vel {ev:.3e} pres {ep:.3e}", flush=True)
    assert ev < 1e-10 and ep < 1e-10
    for k in range(len(shapes_of(cup))):
        a, b = g.shape(k), o.shape(k)
        for key in ["u", "v", "omega"]:
            assert abs(a[key] - b[key]) <= 1e-10 * max(abs(b[key]), 1e-3), (k, key, a[key], b[key])
        assert np.allclose(a["centerOfMass"], b["centerOfMass"], rtol=1e-12, atol=1e-14), k
    _force_log_check(g, recs, nshapes)


def test_8192x4096_disks_vs_oracle(cup, oracle_mod):
    """The scale-up grid (periodic 8192x4096, rows of 8192 and columns of
    4096 in the FFT) with the sixteen static disks, nu = 5e-6, seeded
    random-vorticity initial condition; ten device-loop steps."""
    _steps_vs_oracle(cup, oracle_mod, (256, 128), lambda m: bench.mixed_layout(m, 0, 16), 10,
                     extent=2.0, nu=5e-6, CFL=0.1, poisson="periodic")


def test_8192x4096_swimmers_and_disks_vs_oracle(cup, oracle_mod):
    """The mixed scale-up configuration in FP64: periodic 8192x4096, four
    swimmers with seeded phases and positions and the sixteen static disks,
    nu = 5e-6; its full 100 device-loop steps, with every step's forces of
    all twenty bodies.  The swimmers' bounds (disks of radius L + amplitude +
    half-width) cover more blocks than the list builder holds at once, so its
    candidates are taken in windows."""
    _steps_vs_oracle(cup, oracle_mod, (256, 128), lambda m: bench.mixed_layout(m, 4, 16), 100,
                     extent=2.0, nu=5e-6, CFL=0.1, poisson="periodic")
