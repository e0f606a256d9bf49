"""Benchmark of the CubismUP_2D per-timestep hot path on one MI355X.

Default workload: periodic 2.0 x 1.0 domain at 4096 x 2048 cells (128 x 64
blocks of 32x32), nu = 1e-5, CFL 0.1, FP64, initial velocity from a seeded
Gaussian random vorticity (sigma 1e-3) made divergence-free by one projection,
one free carangiform swimmer (CarlingFish midline kinematics, length 0.2,
lateral amplitude envelope L (0.02 + 0.08 (s/L)^2), period 1.0, wavelength 1.0 L,
NACA 0008 half-width profile) as the immersed body.  A step is the reference's
calcMaxTimestep + advance: PutObjectsOnGrid, advDiff, PressureSingle (RHS,
FFT Poisson, projection, body update, penalization) and ComputeForces, all
enqueued on the device with dt computed on the device.

Prints one JSON line (metric, value = Mcell-steps/s over the K timed steps,
roofline of the dominant kernel measured live with HIP events, and a CPU
baseline: the FP64 OpenMP restatement timed on the host on a bounded sample).
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# the CPU baseline's OpenMP threads: one per core, packed (set before the
# oracle library starts its OpenMP runtime)
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")

PEAK_HBM_GBS = 8000.0
# algorithmic HBM bytes per cell and launch of the full-grid kernels of a step,
# in reals (x 8 bytes, FP64), periodic solver.  The forward row FFT forms the
# pressure RHS from the advDiff output itself: it reads u*, v* (2) and writes
# the half spectrum (1); there is no separate RHS pass.
BALG = {"advdiff": 4, "fft_rows_fwd": 3, "fft_cols": 2, "fft_rows_inv": 2,
        "pressure_correction": 5}
BALG_FREESPACE = dict(BALG, fft_rows_fwd=4, fft_cols=5, fft_rows_inv=3)
# the metric's per-step figure: each stage with its own pass (advDiff 4, RHS 3,
# FFT 2+2+2, projection 5 = 18 reals = 144 B/cell periodic; 23 = 184 B free-space)
STEP_BALG_METRIC = {"periodic": 18, "freespace": 23}
# FP32 precision, periodic, bytes per cell of one step's full-grid passes:
# advDiff 16 (R u,v; W u*,v* in float), forward rows 16 (R u*,v*; W the
# float RHS b for the refinement; W the float half spectrum), columns 8,
# inverse rows 12 (R float spectrum; W the double pressure), projection 24
# (R double p, float u*,v*; W float u,v) = 76; each FP64 refinement sweep adds
# 76: rows 16 (R p; W double spectrum), apply columns 16, the residual's
# inverse rows 16 (R spectrum, b; W the residual's FP32 half spectrum: its
# forward rows are fused in), FP32 columns 8 and inverse rows 20 (R float
# spectrum, R and W p).  Without the fusion (mirrored cosine rows) the
# residual plane is written and read back: 84.
STEP_BYTES_FP32 = {"base": 76, "sweep": 76, "sweep_unfused": 84}
FULL_GRID = list(BALG)

SWIMMER = dict(kind="carlingfish", xpos=1.0, ypos=0.5, L=0.2, T=1.0, envelope="quadratic",
               env_a0=0.02, env_a2=0.08, waveLength=1.0, profile="naca", naca_t=0.08)
CONFIGS = {
    # name: (bpdx, bpdy, extent, nu, poisson, body)
    "periodic4096": dict(bpdx=128, bpdy=64, extent=2.0, nu=1e-5, poisson="periodic",
                         body=SWIMMER),
    "periodic4096_disk": dict(bpdx=128, bpdy=64, extent=2.0, nu=1e-5, poisson="periodic",
                              body=dict(radius=0.04, xpos=1.0, ypos=0.5)),
    "cylinder512": dict(bpdx=16, bpdy=8, extent=1.0, nu=2e-5, poisson="freespace",
                        body=dict(radius=0.05, xpos=0.25, ypos=0.25, xvel=0.2, bFixed=True,
                                  bForced=True)),
    # the mixed scale-up configuration in FP64: four swimmers, sixteen static disks
    "scaleup8192": dict(bpdx=256, bpdy=128, extent=2.0, nu=5e-6, poisson="periodic",
                        bodies=(4, 16)),
}


def ensemble_members(n=64, seed=3):
    """The batched-ensemble workload: n independent towed cylinders on grids
    drawn from the seed between 256x128 and 1024x512 cells (NX a multiple of
    64, NY = NX/2, domain 1.0 x 0.5), radius U[0.03, 0.08], centre in the
    left half, towing speed U[0.1, 0.3], nu = 1e-4, free-space Poisson.
    Returns (bpdx, bpdy, disk keyword arguments) per member."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        bpdx = 2 * int(rng.integers(4, 17))
        shape = dict(radius=float(rng.uniform(0.03, 0.08)), xpos=float(rng.uniform(0.15, 0.5)),
                     ypos=float(rng.uniform(0.15, 0.35)), xvel=float(rng.uniform(0.1, 0.3)),
                     bFixed=True, bForced=True)
        out.append((bpdx, bpdx // 2, shape))
    return out


ENSEMBLE_KW = dict(extent=1.0, nu=1e-4, CFL=0.1, poisson="freespace")

# The adjoint workload: the impulsively started towed cylinder at 1024x512
# (free-space, r = 0.05 at (0.25, 0.25) towed at 0.2, nu = 2e-5, CFL 0.1),
# 200 steps, gradient of the mean drag with respect to 64 surface-velocity
# coefficients (16 modes, normal and tangential) drawn N(0, 1e-3) from the
# seed, checkpoints every 20 steps.
ADJOINT = dict(bpdx=32, bpdy=16, extent=1.0, nu=2e-5, CFL=0.1, poisson="freespace",
               body=dict(radius=0.05, xpos=0.25, ypos=0.25, xvel=0.2, bFixed=True, bForced=True),
               nsteps=200, every=20, nmodes=16, sigma=1e-3, seed=4)


def adjoint_theta(cfg=ADJOINT):
    return np.random.default_rng(cfg["seed"]).normal(0.0, cfg["sigma"], 4 * cfg["nmodes"])


def adjoint_sim(mod, cfg=ADJOINT, bpdx=None, bpdy=None):
    """A reset Simulation of the adjoint workload (library or oracle module)."""
    kw = {k: cfg[k] for k in ("extent", "nu", "CFL", "poisson")}
    sim = mod.Simulation(bpdx or cfg["bpdx"], bpdy or cfg["bpdy"],
                         shapes=[mod.disk(**cfg["body"])], **kw)
    sim.reset()
    return sim


def random_vorticity_velocity(NX, NY, h, sigma=1e-3, seed=12345):
    """Seeded Gaussian vorticity -> periodic streamfunction -> central-difference u."""
    rng = np.random.default_rng(seed)
    w = rng.normal(0.0, sigma, (NY, NX))
    W = np.fft.rfft2(w)
    kx = 2 * np.pi * np.fft.rfftfreq(NX)[None, :] / h
    ky = 2 * np.pi * np.fft.fftfreq(NY)[:, None] / h
    k2 = kx * kx + ky * ky
    k2[0, 0] = 1.0
    P = W / k2
    P[0, 0] = 0
    psi = np.fft.irfft2(P, s=(NY, NX))
    u = (np.roll(psi, -1, axis=0) - np.roll(psi, 1, axis=0)) / (2 * h)
    v = -(np.roll(psi, -1, axis=1) - np.roll(psi, 1, axis=1)) / (2 * h)
    return np.stack([u, v])


def setup(sim, vel):
    sim.reset()
    sim.set_field(0, vel)
    sim.pressure_rhs(1.0)  # one projection makes the IC discretely divergence-free
    sim.poisson_solve()
    sim.pressure_correction(1.0)


def make_shape(mod, body):
    """Shape descriptor of the workload's body for the library or the oracle."""
    b = dict(body)
    if b.pop("kind", "disk") == "carlingfish":
        return mod.carlingfish(**b)
    return mod.disk(**b)


def mixed_layout(mod, nfish, ndisks, seed=2024):
    """A seeded layout of the mixed scale-up configuration: swimmers of the
    benchmark's kind with seeded phases and positions, static disks of radius
    U[0.01, 0.03] that keep clear of the swimmers and of each other."""
    rng = np.random.default_rng(seed)
    fish = []
    while len(fish) < nfish:
        x, y = rng.uniform(0.3, 1.7), rng.uniform(0.25, 0.75)
        if all((x - a) ** 2 + (y - b) ** 2 > 0.3 ** 2 for a, b, _ in fish):
            fish.append((x, y, rng.uniform(0.0, 1.0)))
    disks = []
    while len(disks) < ndisks:
        r = rng.uniform(0.01, 0.03)
        x, y = rng.uniform(0.1, 1.9), rng.uniform(0.1, 0.9)
        if all((x - a) ** 2 + (y - b) ** 2 > (r + 0.16) ** 2 for a, b, _ in fish) and \
           all((x - a) ** 2 + (y - b) ** 2 > (r + q + 0.02) ** 2 for a, b, q in disks):
            disks.append((x, y, r))
    sw = {k: v for k, v in SWIMMER.items() if k not in ("kind", "xpos", "ypos")}
    shapes = [mod.carlingfish(x, y, phi=phi, **sw) for x, y, phi in fish]
    shapes += [mod.disk(radius=r, xpos=x, ypos=y, bForced=True) for x, y, r in disks]
    return shapes


def make_shapes(mod, cfg):
    """Shape descriptors of a workload for the library or the oracle."""
    if "bodies" in cfg:
        return mixed_layout(mod, *cfg["bodies"])
    return [make_shape(mod, cfg["body"])]


def bodies_text(cfg):
    if "bodies" in cfg:
        return (f"{cfg['bodies'][0]} free carlingfish of the default kind and "
                f"{cfg['bodies'][1]} static disks r=U[0.01,0.03], seeded layout")
    return body_text(cfg["body"])


def body_text(body):
    if body.get("kind") == "carlingfish":
        return (f"one free carlingfish L={body['L']} T={body['T']}, envelope "
                f"L({body['env_a0']}+{body['env_a2']}(s/L)^2), wavelength "
                f"{body['waveLength']}L, NACA t={body['naca_t']}")
    return f"one disk r={body['radius']}"


def host_cpu():
    """nproc, the CPU model (lscpu) and the OpenMP binding of this host."""
    model = ""
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
                break
    except (OSError, subprocess.SubprocessError):
        pass
    return dict(nproc=os.cpu_count(), model=model,
                omp=f"OMP_PROC_BIND={os.environ.get('OMP_PROC_BIND', '')} "
                    f"OMP_PLACES={os.environ.get('OMP_PLACES', '')} "
                    f"OMP_NUM_THREADS={os.environ.get('OMP_NUM_THREADS', '')}")


def cpu_step_times(cfg, vel, steps, checker):
    """Per-step wall times of calcMaxTimestep + advance of the FP64 OpenMP
    restatement on the same workload (checker: the reference's NaN/Inf scan
    before every operator, SimulationData.cpp:187-194)."""
    import oracle
    o = oracle.Oracle(cfg["bpdx"], cfg["bpdy"], extent=cfg["extent"], nu=cfg["nu"], CFL=0.1,
                      poisson=cfg["poisson"], fs_pad2=True, checker=checker,
                      shapes=make_shapes(oracle, cfg))
    setup(o, vel)
    t = []
    for _ in range(steps):
        t0 = time.perf_counter()
        o.simulate(1)
        t.append(time.perf_counter() - t0)
    total, fft = o.timing()
    return np.array(t), fft / max(total, 1e-300)


def cpu_baseline(cfg, vel, steps):
    """The CPU baseline: both modes of the oracle, median step time over steps
    6..N (the first 5 are warm-up).  value = the faster (lean) mode."""
    import oracle
    cells = cfg["bpdx"] * cfg["bpdy"] * 1024
    res = {}
    for mode, chk in (("lean", False), ("faithful", True)):
        t, fft_share = cpu_step_times(cfg, vel, steps, chk)
        med = float(np.median(t[5:]))
        res[mode] = dict(value=round(cells / med / 1e6, 2), median_step_ms=round(med * 1e3, 3),
                         fft_share=round(fft_share, 3))
    best = max(res, key=lambda k: res[k]["value"])
    return dict(value=res[best]["value"], unit="Mcell-steps/s", cores=oracle.num_threads(),
                kind="port", host=host_cpu(), modes=res,
                sample=f"{steps} steps of the same workload per mode, median step time over "
                       f"steps 6..{steps}; value is the {best} mode (FP64 OpenMP "
                       f"restatement, no NaN checker); faithful adds the reference's NaN/Inf "
                       f"Checker before every operator")


KERNEL_OF = {"k_advdiff": "advdiff", "k_rhs<": "pressure_rhs", "k_rows_fwd": "fft_rows_fwd",
             "k_cols": "fft_cols", "k_rows_inv": "fft_rows_inv", "k_project": "pressure_correction",
             "k_find_max_u": "find_max_u"}


def copy_calibration(mib):
    """A plain device copy of known size (dst <- src, mib MiB each way), run
    before the workload so one counter pass sees both."""
    import torch
    n = mib * (1 << 20) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    return n * 8


def traffic_from_counters(fetch_csv, write_csv, copy_bytes, config):
    """HBM bytes per launch of each full-grid kernel from two rocprofv3 --pmc
    passes (FETCH_SIZE, WRITE_SIZE in KiB per dispatch), each scaled by
    copy_bytes / (the same counter for the calibration copy)."""
    import csv

    def per_kernel(path, counter):
        acc = {}
        with open(path) as f:
            for row in csv.DictReader(f):
                if row["Counter_Name"] != counter:
                    continue
                acc.setdefault(row["Kernel_Name"], []).append(float(row["Counter_Value"]))
        return acc

    fetch_all = per_kernel(fetch_csv, "FETCH_SIZE")
    write_all = per_kernel(write_csv, "WRITE_SIZE")
    # the calibration copy: the kernel other than ours that both reads and
    # writes the most (a fill kernel writes as much but reads nothing)
    others = [k for k in write_all if "cup2d" not in k and k in fetch_all]
    if not others:
        raise SystemExit("no calibration copy kernel in the counter files")
    ck = max(others, key=lambda k: min(max(fetch_all[k]), max(write_all[k])))
    if max(write_all[ck]) * 1024 < 0.5 * copy_bytes:
        raise SystemExit("no dispatch of the calibration copy's size in the counter files")
    f_scale = copy_bytes / (max(fetch_all[ck]) * 1024)
    w_scale = copy_bytes / (max(write_all[ck]) * 1024)
    fetch = {k: sum(v) / len(v) for k, v in fetch_all.items()}
    write = {k: sum(v) / len(v) for k, v in write_all.items()}
    out = {"_calibration": {"copy_bytes": copy_bytes, "kernel": ck[:80],
                            "fetch_scale": round(f_scale, 4), "write_scale": round(w_scale, 4)}}
    # several template instances can match one name (e.g. the row FFT of the
    # set-up solve and the whole-step one): take the one dispatched most often
    best = {}
    for k in fetch:
        if re.search(r"k_rows_fwd<\d+, \d+, \d+, 3,", k):
            continue  # the obstacle-row patch pass, not the forward row pass
        for pat, name in KERNEL_OF.items():
            if pat in k and k in write and len(fetch_all[k]) > best.get(name, (0, ""))[0]:
                best[name] = (len(fetch_all[k]), k)
    for name, (_, k) in best.items():
        out[name] = {"fetch_bytes": round(fetch[k] * 1024 * f_scale),
                     "write_bytes": round(write[k] * 1024 * w_scale),
                     "traffic": round((fetch[k] * f_scale + write[k] * w_scale) * 1024),
                     "dispatches": len(fetch_all[k])}
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    allcfg = {}
    if os.path.exists(path):
        with open(path) as f:
            allcfg = json.load(f)
    allcfg[config] = out
    with open(path, "w") as f:
        json.dump(allcfg, f, indent=1, sort_keys=True)
    print(json.dumps(out, indent=1))


def fp32_leg(cup, cfg, vel, args, sweeps):
    """The same workload at FP32 precision (float stencils and FFT Poisson,
    FP64 pressure with `sweeps` FP64 defect-correction sweeps per solve),
    timed like the FP64 leg."""
    sim = cup.Simulation(cfg["bpdx"], cfg["bpdy"], extent=cfg["extent"], nu=cfg["nu"], CFL=0.1,
                         poisson=cfg["poisson"], shapes=make_shapes(cup, cfg), precision="fp32",
                         refine_sweeps=sweeps)
    setup(sim, vel)
    sim.simulate(args.warmup)
    sim.synchronize()
    t0 = time.perf_counter()
    sim.simulate(args.steps)
    sim.synchronize()
    dt = time.perf_counter() - t0
    res = sim.refine_residuals()  # of the last step's solve, before each sweep
    sim.close()
    cells = cfg["bpdx"] * cfg["bpdy"] * 1024
    sweep_bytes = STEP_BYTES_FP32["sweep" if cfg["poisson"] == "periodic" else "sweep_unfused"]
    step_bytes = STEP_BYTES_FP32["base"] + sweeps * sweep_bytes
    achieved = step_bytes * cells * args.steps / dt / 1e9
    return {"value": round(cells * args.steps / dt / 1e6, 2), "unit": "Mcell-steps/s",
            "ms_per_step": round(dt / args.steps * 1e3, 5), "refine_sweeps": sweeps,
            "residual_before_sweeps": [float(f"{r:.3e}") for r in res],
            "dtype": "f32 (FP64 pressure)",
            "roofline": {"bound": "hbm", "scope": "whole step", "achieved": round(achieved, 1),
                         "peak": PEAK_HBM_GBS, "unit": "GB/s",
                         "frac": round(achieved / PEAK_HBM_GBS, 4),
                         "alg_bytes_per_cell_step": step_bytes}}


def ensemble_leg(cup, steps, precision):
    """The batched ensemble (64 towed cylinders, grids 256x128 .. 1024x512,
    free-space): one captured graph advances every member two steps; the
    lane count is autotuned on the live run first (those steps are not
    timed).  value = total member cell-steps per second."""
    members = ensemble_members()
    sims = []
    for bpdx, bpdy, shape in members:
        g = cup.Simulation(bpdx, bpdy, shapes=[cup.disk(**shape)], precision=precision,
                           **ENSEMBLE_KW)
        g.reset()
        sims.append(g)
    ens = cup.Ensemble(sims)
    lanes, ms_tuned = ens.tune(launches=2)
    ens.simulate(2)  # the tuned graph captured and warm
    t0 = time.perf_counter()
    ens.simulate(steps)
    dt = time.perf_counter() - t0
    cells = sum(bpdx * bpdy * 1024 for bpdx, bpdy, _ in members)
    ens.close()
    for g in sims:
        g.close()
    return {"value": round(cells * steps / dt / 1e6, 2), "unit": "Mcell-steps/s",
            "members": len(members), "cells_total": cells, "steps": steps,
            "ms_per_step": round(dt / steps * 1e3, 4), "lanes": lanes,
            "ms_per_launch_tuned": round(ms_tuned, 4), "precision": precision,
            "launches_per_step": 0.5}


def adjoint_leg(cup):
    """adjoint/fwd on the adjoint workload: wall time of the backward sweep
    (every segment recomputed from its checkpoint, then the adjoint steps)
    over that of the forward run (the same steps, time steps fixed)."""
    cfg = ADJOINT
    sim = adjoint_sim(cup, cfg)
    sim.set_surface_velocity(adjoint_theta(cfg))
    _, dts = sim.adjoint_forward(cfg["nsteps"], cfg["every"])  # the CFL time steps
    sim.close()
    sim = adjoint_sim(cup, cfg)
    sim.set_surface_velocity(adjoint_theta(cfg))
    t0 = time.perf_counter()
    J, _ = sim.adjoint_forward(cfg["nsteps"], cfg["every"], dts)
    t1 = time.perf_counter()
    grad = sim.adjoint_backward()
    t2 = time.perf_counter()
    sim.close()
    cells = cfg["bpdx"] * cfg["bpdy"] * 1024
    return {"adjoint_over_forward": round((t2 - t1) / (t1 - t0), 3),
            "forward_ms": round((t1 - t0) * 1e3, 2), "backward_ms": round((t2 - t1) * 1e3, 2),
            "workload": f"towed cylinder {cells // 512}x512 free-space, {cfg['nsteps']} steps, "
                        f"checkpoint every {cfg['every']}, {4 * cfg['nmodes']} surface-velocity "
                        f"coefficients",
            "J": J, "grad_norm": float(np.linalg.norm(grad))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--config", default="periodic4096", choices=sorted(CONFIGS))
    ap.add_argument("--cpu-steps", type=int, default=120,
                    help="steps of each CPU-baseline mode (lean, faithful)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32-precision legs")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint/fwd leg")
    ap.add_argument("--ensemble-steps", type=int, default=100,
                    help="steps of the 64-member ensemble legs (0 skips them)")
    ap.add_argument("--timer-every", type=int, default=8, metavar="K",
                    help="time the dominant kernel on every K-th step of the timed region "
                         "(each timed launch costs the stream about 5 us)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="run a plain MIB-MiB device copy first (counter calibration)")
    ap.add_argument("--traffic", nargs=3, metavar=("FETCH_CSV", "WRITE_CSV", "COPY_MIB"),
                    help="write profiles/hbm_traffic.json from two rocprofv3 --pmc passes")
    args = ap.parse_args()
    if args.traffic:
        traffic_from_counters(args.traffic[0], args.traffic[1],
                              int(args.traffic[2]) * (1 << 20), args.config)
        return
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; --gpus must be 1 "
              "(multi-GPU sharding is a design note, not implemented)", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
    import cubismup_2d_amd as cup
    cfg = CONFIGS[args.config]
    sim = cup.Simulation(cfg["bpdx"], cfg["bpdy"], extent=cfg["extent"], nu=cfg["nu"], CFL=0.1,
                         poisson=cfg["poisson"], shapes=make_shapes(cup, cfg))
    NX, NY = sim.NX, sim.NY
    h = sim.scalars()["h"]
    vel = random_vorticity_velocity(NX, NY, h)
    setup(sim, vel)
    cells = NX * NY
    balg = dict(BALG_FREESPACE if cfg["poisson"] == "freespace" else BALG)

    # warm-up with every full-grid timer on: picks the dominant kernel
    sim.timer_enable("*", 0)
    for k in FULL_GRID:
        sim.timer_enable(k, max(2 * args.warmup, 16))
    sim.simulate(args.warmup)
    per = {k: sim.timer_read(k) for k in FULL_GRID}
    dominant = max(per, key=lambda k: per[k][0])
    sim.timer_enable("*", 0)
    # the dominant kernel is timed on every --timer-every-th step of the timed
    # region: each event pair costs the stream a few microseconds
    sim.timer_sample(dominant, args.timer_every)
    sim.timer_enable(dominant, args.steps // args.timer_every + 8)

    sim.synchronize()
    t0 = time.perf_counter()
    sim.simulate(args.steps)
    sim.synchronize()
    t1 = time.perf_counter()
    tot_ms, launches = sim.timer_read(dominant)
    dt = t1 - t0
    value = cells * args.steps / dt / 1e6
    avg_s = tot_ms / 1e3 / max(launches, 1)
    bytes_launch = balg[dominant] * 8 * cells
    achieved = bytes_launch / avg_s / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            t = json.load(f).get(args.config, {}).get(dominant)
        traffic = t["traffic"] if t else None
    step_balg = STEP_BALG_METRIC[cfg["poisson"]] * 8
    out = {
        "metric": "Mcell-steps/s at 4096x2048 (FP64 vs FP32+refine), % HBM roofline, adjoint/fwd",
        "value": round(value, 2),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded random-vorticity initial velocity",
        "config": {"workload": f"{args.config}: {NX}x{NY} cells ({cfg['bpdx']}x{cfg['bpdy']} "
                               f"blocks), {cfg['poisson']} FFT Poisson, nu={cfg['nu']}, CFL 0.1, "
                               f"{bodies_text(cfg)}",
                   "cells": cells, "precision": "fp64", "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": round(achieved, 1),
                     "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
                     "traffic": traffic, "avg_launch_us": round(avg_s * 1e6, 2),
                     "timed_launches": launches,
                     "alg_bytes_per_launch": bytes_launch},
        "step_hbm_frac": round(step_balg * cells * args.steps / dt / 1e9 / PEAK_HBM_GBS, 4),
        "kernel_us_warmup": {k: round(per[k][0] * 1e3 / max(per[k][1], 1), 2) for k in per},
    }
    sim.close()
    if cfg["poisson"] == "periodic" and not args.no_fp32:
        out["fp32"] = {f"sweeps{k}": fp32_leg(cup, cfg, vel, args, k) for k in (0, 1, 2, 3)}
    if not args.no_adjoint:
        out["adjoint"] = adjoint_leg(cup)
    if args.ensemble_steps > 0:
        out["ensemble64"] = {p: ensemble_leg(cup, args.ensemble_steps, p)
                             for p in ("fp64", "fp32")}
    if not args.no_cpu and args.cpu_steps > 0:
        out["cpu_baseline"] = cpu_baseline(cfg, vel, args.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
