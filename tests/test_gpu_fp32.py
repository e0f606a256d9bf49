"""FP32 precision (cup2d_cfg.precision = CUP2D_PRECISION_FP32): float velocity
and RHS planes, float advDiff / RHS / FFT Poisson / projection, FP64 pressure
with FP64 defect-correction sweeps, FP64 body integrals and forces -- the
reference's float build (Definitions.h:18-22, fftwf plans FFTW_periodic.h:
120-126) with the Poisson solve refined in FP64.

The bar differs from the FP64 path's by nature: a float field carries 2^-24
relative rounding per operation, so results are compared with the FP64
oracle (or the FP64 device run, which matches the oracle to 1e-10) with the
error stated per field and per body force, and bounded.  What FP32 cannot
blur is pinned: the refinement sweeps must take the FP32 solve of a float
RHS to the FP64 solution of that same RHS (the spectral operator's residual
is formed in FP64), to the FP64 solve's own accuracy (1e-10 relative L2 at
4096x2048) within two sweeps; the measured residual |b - A p| / |b| falls
below 1e-10 after three sweeps at 4096x2048 (two at 512x256)."""
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


@pytest.mark.parametrize("poisson,bp", [("periodic", (16, 8)), ("cosine", (16, 8)),
                                        ("periodic", (128, 64))])
def test_fp32_refinement_reaches_fp64_solution(cup, oracle_mod, poisson, bp):
    """FP32 FFT solve + k FP64 sweeps of r = b - A p, p += S32(r), against the
    oracle's FP64 solve of the same (float-rounded) RHS: no sweep leaves the
    FP32 transform's round-off (~1e-6), each sweep gains three or more
    digits, two reach the FP64 solve's own level."""
    rng = np.random.default_rng(61)
    bpdx, bpdy = bp
    o = oracle_mod.Oracle(bpdx, bpdy, extent=1.0, poisson=poisson)
    NX, NY = bpdx * 32, bpdy * 32
    X, Y = np.meshgrid(np.arange(NX) / NX, np.arange(NY) / NY)
    r = np.sin(2 * np.pi * X) * np.cos(6 * np.pi * Y) + 0.1 * rng.standard_normal((NY, NX))
    r32 = r.astype(np.float32).astype(np.float64)
    ref = o.poisson_dense(r32)
    errs = []
    for sweeps in range(4):
        g = cup.Simulation(bpdx, bpdy, extent=1.0, poisson=poisson, precision="fp32",
                           refine_sweeps=sweeps)
        errs.append(rel(g.poisson_dense(r), ref))
        res = g.refine_residuals()
        g.close()
    print(f"  {poisson} {NX}x{NY}: rel L2 vs FP64 solve, sweeps 0/1/2/3: "
          + " ".join(f"{e:.2e}" for e in errs)
          + "; residual |b - A p|/|b| before each of 3 sweeps: "
          + " ".join(f"{e:.2e}" for e in res), flush=True)
    assert 1e-9 < errs[0] < 1e-5  # the FP32 solve alone is FP32-accurate, not better
    # each sweep gains three or more digits until the FP64 solve's own
    # round-off (cond ~ (N/pi)^2 times 1e-16) is reached
    assert errs[1] < 2e-9 and errs[2] < 1e-10 and errs[3] < 1e-10
    # the residual |b - A p| without its mean (which the operator discards);
    # A ~ k^2 weighs the FP32 solve's high-frequency round-off, so the
    # residual lags the error: measured 4.6e-2, 9.9e-8, 7.3e-10 before sweeps
    # 1, 2, 3 at 4096x2048 (5e-4, 5e-10, 2e-11 at 512x256)
    assert res[0] > res[1] > res[2] and res[2] < 1e-9


def test_fp32_advdiff_one_step(cup, oracle_mod):
    """One FP32 advDiff (the stencil evaluated in float) against FP64."""
    rng = np.random.default_rng(62)
    kw = dict(extent=2.0, nu=1e-4)
    g = cup.Simulation(32, 16, precision="fp32", **kw)
    o = oracle_mod.Oracle(32, 16, **kw)
    vel = 0.3 * rng.standard_normal((2, g.NY, g.NX))
    vel32 = vel.astype(np.float32).astype(np.float64)
    for sim in (g, o):
        sim.reset()
        sim.set_field(0, vel32)
        sim.set_scalars(uinf=(0.05, -0.02))
        sim.advdiff(1e-3)
    e = rel(g.get_field(0), o.get_field(0))
    print(f"  advDiff FP32 vs FP64 rel L2 {e:.2e}", flush=True)
    assert e < 3e-7


def _run_fp32_vs_fp64(cup, cfg, n, make_shapes):
    """The same workload at FP32 (refined Poisson) and FP64 on the device; the
    FP64 run is the reference (it matches the oracle to 1e-10, see the
    full-size parity tests).  Returns per-field errors and per-body force
    errors relative to the body's force scale."""
    kw = dict(extent=cfg["extent"], nu=cfg["nu"], CFL=0.1, poisson=cfg["poisson"])
    runs = {}
    for prec in ("fp64", "fp32"):
        g = cup.Simulation(cfg["bpdx"], cfg["bpdy"], shapes=make_shapes(), precision=prec, **kw)
        vel = bench.random_vorticity_velocity(g.NX, g.NY, g.scalars()["h"])
        bench.setup(g, vel)
        g.force_log_enable(n)
        g.simulate(n)
        nsh = len(make_shapes())
        runs[prec] = dict(vel=g.get_field(0), pres=g.get_field(2),
                          forces=[g.force_log(k) for k in range(nsh)],
                          shapes=[g.shape(k) for k in range(nsh)], scal=g.scalars())
        g.close()
    a, b = runs["fp32"], runs["fp64"]
    ev, ep = rel(a["vel"], b["vel"]), rel(a["pres"], b["pres"])
    cols = [cup.FORCE_LOG_COLUMNS.index(c) for c in ("Fx", "Fy")]
    ferr = []
    for fa, fb in zip(a["forces"], b["forces"]):
        scale = np.abs(fb[:, cols]).max()
        ferr.append(np.abs(fa[:, cols] - fb[:, cols]).max() / max(scale, 1e-300))
    return ev, ep, ferr, a, b


def test_fp32_bench_workload_200_steps(cup):
    """The benchmarked 4096x2048 periodic swimmer workload, 200 steps at FP32
    with one FP64 refinement sweep per solve (the default), against the FP64
    run: velocity, pressure and every step's swimmer force."""
    cfg = bench.CONFIGS["periodic4096"]
    ev, ep, ferr, a, b = _run_fp32_vs_fp64(cup, cfg, 200,
                                           lambda: [bench.make_shape(cup, cfg["body"])])
    assert a["scal"]["step"] == b["scal"]["step"] == 200
    print(f"  FP32+refine vs FP64 after 200 steps: rel L2 vel {ev:.3e} pres {ep:.3e}, "
          f"max per-step force error {ferr[0]:.3e} (of the run's max |F|)", flush=True)
    # measured on MI355X: vel 6.6e-8, pres 1.3e-6, force 5.7e-7
    assert ev < 1e-6 and ep < 2e-5 and ferr[0] < 1e-5


def test_fp32_scaleup_100_steps(cup):
    """The mixed scale-up workload (8192x4096, four swimmers, sixteen static
    disks, nu = 5e-6), 100 steps at FP32 against the FP64 device run: error
    per field and for each body's force."""
    cfg = bench.CONFIGS["scaleup8192"]
    ev, ep, ferr, a, b = _run_fp32_vs_fp64(cup, cfg, 100, lambda: bench.make_shapes(cup, cfg))
    print(f"  FP32+refine vs FP64 after 100 steps: rel L2 vel {ev:.3e} pres {ep:.3e}", flush=True)
    print("  per-body max per-step force error: "
          + " ".join(f"{e:.1e}" for e in ferr), flush=True)
    # measured on MI355X: vel 9.2e-8, pres 2.2e-6, forces <= 1.4e-5
    assert ev < 1e-6 and ep < 2e-5 and max(ferr) < 2e-4


def test_fp32_swimmer_in_fade_band(cup):
    """A swimmer whose obstacle blocks reach into the fade band, at FP32 with
    one refinement sweep, against the FP64 device run (which matches the
    oracle, see the parity tests): the early float row pass is then redone
    with the final fade factor, and the RHS plane the refinement reads must
    hold every cell's final value."""
    from test_gpu_parity import FISH_VARIANTS
    fish = dict(xpos=1.0, ypos=0.08, angle=0.0, **FISH_VARIANTS["carling"])
    kw = dict(extent=2.0, nu=1e-4, CFL=0.2, poisson="periodic", fadeLen=0.05)
    runs = {}
    for prec in ("fp64", "fp32"):
        g = cup.Simulation(32, 16, shapes=[cup.carlingfish(**fish)], precision=prec, **kw)
        g.reset()
        rng = np.random.default_rng(63)
        g.set_field(0, 0.02 * rng.standard_normal((2, g.NY, g.NX)))
        g.simulate(8)
        runs[prec] = (g.get_field(0), g.get_field(2))
        g.close()
    ev = rel(runs["fp32"][0], runs["fp64"][0])
    ep = rel(runs["fp32"][1], runs["fp64"][1])
    print(f"  FP32+refine vs FP64, swimmer in the fade band, 8 steps: rel L2 vel {ev:.2e} "
          f"pres {ep:.2e}", flush=True)
    assert ev < 1e-6 and ep < 2e-5
