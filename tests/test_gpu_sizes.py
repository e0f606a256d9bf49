"""Grids whose sizes are not powers of two (any number of 32-cell blocks):
the free-space solver pads rows and columns to the next power of two >= 2N
(any cyclic length >= 2N - 1 gives the same linear convolution on the N data
points; the reference pads to 2N - 1, FFTW_freespace.h:25-27), so these
grids run whole steps.  Checked against the brute-force direct sum and
against the oracle with the reference's own 2N - 1 padding."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
import numpy_ref  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[gpu case] {request.node.name}", flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("bp", [(3, 1), (3, 3), (5, 2)])
def test_freespace_non_pow2_direct_sum(cup, bp):
    """96x32, 96x96 and 160x64 cells (row lengths 256, 256, 512; columns 64,
    256, 128): the padded solve equals the direct sum."""
    rng = np.random.default_rng(72)
    g = cup.Simulation(*bp, extent=1.0, poisson="freespace")
    r = rng.standard_normal((g.NY, g.NX))
    p = g.poisson_dense(r)
    ref = numpy_ref.poisson_freespace_direct(r, g.scalars()["h"])
    e = rel(p, ref)
    print(f"  {g.NX}x{g.NY}: rel L2 vs direct sum {e:.2e}", flush=True)
    assert e < 1e-12


@pytest.mark.parametrize("bp", [(10, 5), (9, 4), (13, 7)])
def test_freespace_non_pow2_towed_cylinder(cup, oracle_mod, bp):
    """The towed cylinder on 320x160, 288x128 and 416x224 grids, 30 steps in
    the device loop, against the oracle with 2N - 1 padding."""
    shape = dict(radius=0.05, xpos=0.3, ypos=0.2, xvel=0.2, bFixed=True, bForced=True)
    kw = dict(extent=1.0, nu=1e-4, CFL=0.1, poisson="freespace")
    g = cup.Simulation(*bp, shapes=[cup.disk(**shape)], **kw)
    o = oracle_mod.Oracle(*bp, fs_pad2=False, shapes=[oracle_mod.disk(**shape)], **kw)
    g.reset()
    o.reset()
    g.simulate(30)
    o.simulate(30)
    ev, ep = rel(g.get_field(0), o.get_field(0)), rel(g.get_field(2), o.get_field(2))
    a, b = g.shape(), o.shape()
    ed = abs(a["drag"] - b["drag"]) / abs(b["drag"])
    print(f"  {g.NX}x{g.NY}: rel L2 vel {ev:.2e} pres {ep:.2e}, drag {ed:.2e}", flush=True)
    assert ev < 1e-10 and ep < 1e-10 and ed < 1e-8


def test_non_pow2_periodic_and_cosine_refused(cup):
    """Periodic and cosine lengths are the grid's own (N and the mirrored 2N):
    not powers of two, so the Poisson calls fail with an argument error while
    the stencil operators still run."""
    for poisson in ("periodic", "cosine"):
        g = cup.Simulation(10, 5, extent=1.0, poisson=poisson)
        with pytest.raises(cup.Cup2dError) as e:
            g.poisson_dense(np.zeros((g.NY, g.NX)))
        assert e.value.code == cup.ERR_ARG
        g.advdiff(1e-3)
        g.close()
