"""Captured whole steps (cup2d_simulate_graph) and ensembles (one captured
graph advancing many independent simulations, cup2d_ensemble_*): the same
kernels in the same order as the device loop, so every result is
bit-identical to cup2d_simulate's, and the ensemble workload (64 towed
cylinders on grids from 256x128 to 1024x512, free-space) matches the oracle
member by member after its 100 steps."""
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


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


FISH = dict(xpos=0.5, ypos=0.25, L=0.2, T=1.0)
DISK = dict(radius=0.05, xpos=0.3, ypos=0.2, xvel=0.2, bFixed=True, bForced=True)
CASES = {
    "swimmer_periodic": (16, 8, dict(extent=1.0, nu=1e-4, CFL=0.2, poisson="periodic"), "fish"),
    "disk_freespace": (16, 8, dict(extent=1.0, nu=1e-4, CFL=0.1, poisson="freespace"), "disk"),
    "disk_cosine": (16, 8, dict(extent=1.0, nu=1e-4, CFL=0.1, poisson="cosine"), "disk"),
    "disk_freespace_320x160": (10, 5, dict(extent=1.0, nu=1e-4, CFL=0.1, poisson="freespace"),
                               "disk"),
}


def _make(cup, case, **extra):
    bpdx, bpdy, kw, body = CASES[case]
    shape = cup.carlingfish(**FISH) if body == "fish" else cup.disk(**DISK)
    g = cup.Simulation(bpdx, bpdy, shapes=[shape], **kw, **extra)
    g.reset()
    rng = np.random.default_rng(81)
    g.set_field(0, 0.02 * rng.standard_normal((2, g.NY, g.NX)))
    return g


def _state(g):
    return g.get_field(0), g.get_field(2), g.shape(0)


def _same(a, b):
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    for k in ("u", "v", "omega", "forcex", "forcey", "drag", "thrust"):
        assert a[2][k] == b[2][k], k


@pytest.mark.parametrize("case", sorted(CASES))
def test_simulate_graph_bit_identical(cup, case):
    """11 steps (five graph launches and one step on its own) equal 11 steps of
    the device loop bit for bit; then graph, loop and graph again (parity
    changes: the other cached graph) equal the loop alone."""
    a, b = _make(cup, case), _make(cup, case)
    a.simulate(11)
    b.simulate_graph(11)
    _same(_state(a), _state(b))
    a.simulate(9)
    b.simulate_graph(3)
    b.simulate(1)
    b.simulate_graph(5)
    _same(_state(a), _state(b))
    assert a.scalars()["step"] == b.scalars()["step"] == 20
    print(f"  {case}: 20 steps bit-identical", flush=True)


def test_simulate_graph_fp32(cup):
    """The captured steps at FP32 precision (refinement sweeps inside)."""
    a = _make(cup, "swimmer_periodic", precision="fp32")
    b = _make(cup, "swimmer_periodic", precision="fp32")
    a.simulate(8)
    b.simulate_graph(8)
    _same(_state(a), _state(b))


def test_ensemble_members_bit_identical(cup):
    """An ensemble of mixed members (swimmer periodic, disks free-space and
    cosine, a 320x160 grid, one FP32 member) advanced together equals each
    member advanced alone, for every lane count the tuner tries."""
    cases = sorted(CASES)
    alone = [_make(cup, c) for c in cases] + [_make(cup, "disk_freespace", precision="fp32")]
    mem = [_make(cup, c) for c in cases] + [_make(cup, "disk_freespace", precision="fp32")]
    ens = cup.Ensemble(mem)
    ens.simulate(7)
    for g in alone:
        g.simulate(7)
    for g, h in zip(alone, mem):
        _same(_state(g), _state(h))
    lanes, ms = ens.tune(launches=2)  # 6 steps per lane count tried (1, 2, 4, 5)
    ncand = 4
    print(f"  tuned lanes {lanes}, {ms:.3f} ms per two-step launch", flush=True)
    assert lanes in (1, 2, 4, 5)
    ens.simulate(5)
    for g in alone:
        g.simulate(2 * 3 * ncand + 5)
    for g, h in zip(alone, mem):
        _same(_state(g), _state(h))
        assert h.scalars()["step"] == 7 + 6 * ncand + 5
    ens.close()


def test_ensemble_workload_vs_oracle(cup, oracle_mod):
    """The 64-member ensemble workload for its 100 steps (one captured graph
    per two steps, autotuned lanes) in FP64: every member's velocity and
    pressure within 1e-10 relative L2 of the oracle (the reference's 2N - 1
    padding) and its drag within 1e-8."""
    members = bench.ensemble_members()
    sims = []
    for bpdx, bpdy, shape in members:
        g = cup.Simulation(bpdx, bpdy, shapes=[cup.disk(**shape)], **bench.ENSEMBLE_KW)
        g.reset()
        sims.append(g)
    ens = cup.Ensemble(sims)
    lanes, ms = ens.tune(launches=1)  # 2 + 2 steps per lane count
    done = 4 * len([1, 2, 4, 8, 16, 32, 64])
    ens.simulate(100 - done)
    worst_v = worst_p = worst_d = 0.0
    for (bpdx, bpdy, shape), g in zip(members, sims):
        o = oracle_mod.Oracle(bpdx, bpdy, shapes=[oracle_mod.disk(**shape)], **bench.ENSEMBLE_KW)
        o.reset()
        o.simulate(100)
        assert g.scalars()["step"] == 100
        worst_v = max(worst_v, rel(g.get_field(0), o.get_field(0)))
        worst_p = max(worst_p, rel(g.get_field(2), o.get_field(2)))
        a, b = g.shape(0), o.shape(0)
        worst_d = max(worst_d, abs(a["drag"] - b["drag"]) / abs(b["drag"]))
    print(f"  64 members, lanes {lanes} ({ms:.3f} ms per launch): worst rel L2 vel "
          f"{worst_v:.2e} pres {worst_p:.2e}, drag {worst_d:.2e}", flush=True)
    assert worst_v < 1e-10 and worst_p < 1e-10 and worst_d < 1e-8
    ens.close()
