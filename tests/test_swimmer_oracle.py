"""Swimmer (CarlingFish) restatement in the oracle, pinned without the
reference: the sin/cos it uses against numpy, the FishData discretisation
and half-width laws against their formulas, the midline against the
closed-form lateral envelope, frame invariants (arclength, zero internal
momenta), the Towers chi area against the analytic body area, and the
deformation velocity's zero momenta after Shape::removeMoments.
"""
import math

import numpy as np
import pytest


def test_det_sincos_matches_numpy(oracle_mod):
    rng = np.random.default_rng(3)
    xs = np.concatenate([rng.uniform(-60, 60, 4000), rng.uniform(-1e4, 1e4, 500),
                         np.array([0.0, -0.0, 1e-300, math.pi / 4, math.pi / 2, math.pi,
                                   2 * math.pi, -math.pi, 1e5])])
    err_s = err_c = 0.0
    for x in xs:
        s, c = oracle_mod.det_sincos(float(x))
        err_s = max(err_s, abs(s - math.sin(x)) / max(abs(math.sin(x)), 1e-3))
        err_c = max(err_c, abs(c - math.cos(x)) / max(abs(math.cos(x)), 1e-3))
    # about one ulp (2.2e-16) relative away from the zeros; absolute near them
    assert err_s < 5e-16 and err_c < 5e-16, (err_s, err_c)


def test_lane_sum(oracle_mod):
    rng = np.random.default_rng(4)
    for n in [1, 7, 256, 257, 1000, 3000]:
        v = rng.standard_normal(n)
        assert abs(oracle_mod.lane_sum(v) - math.fsum(v)) <= 1e-14 * np.abs(v).sum()


def _fish_sim(oracle_mod, bp=(16, 8), extent=1.0, **fish):
    kw = dict(L=0.2, T=1.0)
    kw.update(fish)
    o = oracle_mod.Oracle(*bp, extent=extent, nu=1e-4, CFL=0.1, poisson="freespace",
                          shapes=[oracle_mod.carlingfish(0.5 * extent, 0.25 * extent, **kw)])
    return o


def _discretisation(L, h):
    """FishData's arclength grid (FishLibrary.h:49-60, FishLibrary.cpp:20-33)."""
    fracRefined = 0.1
    fracMid = 1 - 2 * fracRefined
    dSmid_tgt = h / math.sqrt(2)
    Nmid = int(math.ceil(L * fracMid / dSmid_tgt / 8)) * 8
    dSmid = L * fracMid / Nmid
    Nend = int(math.ceil(fracRefined * L * 2 / (dSmid + 0.125 * h) / 4)) * 4
    dSref = fracRefined * L * 2 / Nend - dSmid
    return Nmid, Nend, dSmid, dSref


@pytest.mark.parametrize("profile", ["carling", "naca"])
def test_discretisation_and_width(oracle_mod, profile):
    o = _fish_sim(oracle_mod, profile=profile)
    o.reset()
    fi = o.fish_info()
    h = o.scalars()["h"]
    L = 0.2
    Nmid, Nend, dSmid, dSref = _discretisation(L, h)
    assert fi["Nm"] == Nmid + 2 * Nend + 1
    assert (fi["Nm"] - 1) % 8 == 0
    ds = np.diff(fi["rS"])
    assert np.all(ds > 0)
    assert np.allclose(ds[Nend:Nend + Nmid], dSmid, rtol=1e-9)
    assert abs(ds[0] - dSref) < 1e-12 and abs(ds[-1] - dSref) < 1e-12
    assert abs(fi["rS"][-1] - L) < 1e-12 and fi["rS"][-1] <= L
    s, w = fi["rS"], fi["width"]
    if profile == "carling":  # CarlingFish.cpp:41-48
        sb, st, wt, wh = .04 * L, .95 * L, .01 * L, .04 * L
        ref = np.where(s < sb, np.sqrt(np.maximum(2 * wh * s - s * s, 0)),
                       np.where(s < st, wh - (wh - wt) * (s - sb) / (st - sb), wt * (L - s) / (L - st)))
    else:
        x = s / L
        ref = 5 * 0.08 * L * (0.2969 * np.sqrt(x) - 0.1260 * x - 0.3516 * x ** 2 + 0.2843 * x ** 3
                              - 0.1036 * x ** 4)
        ref = np.maximum(ref, 0)
        assert abs(w.max() - 0.04 * L) < 1e-4 * L  # NACA 0008: max half-thickness 4% at x=0.3
    assert np.allclose(w, ref, rtol=1e-13, atol=1e-16)
    assert w[0] == 0 and w[-1] < 1e-15 and np.all(w[1:-1] > 0)


@pytest.mark.parametrize("env", [("carling", 0.0), ("quadratic", 0.3)])
def test_midline_closed_form(oracle_mod, env):
    """After time T the ramp is 1; the raw midline (before the centre-of-mass
    frame) is y(s) = A(s) sin(2 pi (s/lambda - t/T + phi)), x(s) the arclength
    reconstruction, and vy = -A(s) (2 pi / T) cos(...)."""
    envelope, phi = env
    o = _fish_sim(oracle_mod, envelope=envelope, phi=phi, profile="naca")
    o.reset()
    t = 1.37
    o.set_scalars(time=t)
    o.put_objects(0.0)
    fi = o.fish_info()
    assert fi["theta_internal"] == 0.0  # dt = 0: no rotation of the frame
    L, T = 0.2, 1.0
    s = fi["rS"]
    if envelope == "carling":
        A = 4 / 33 * (s + 0.03125 * L)
        arg = 2 * np.pi * (s / L - t / T + phi)
    else:
        A = L * (0.02 + 0.08 * (s / L) ** 2)
        arg = 2 * np.pi * (s / L / 1.0 - t / T + phi)
    yraw = fi["rY"] + fi["CoM_internal_y"]
    assert np.allclose(yraw, A * np.sin(arg), rtol=0, atol=1e-15)
    ds = np.diff(s)
    dx = np.sqrt(ds * ds - np.diff(A * np.sin(arg)) ** 2)
    xraw = fi["rX"] + fi["CoM_internal_x"]
    assert np.allclose(xraw, np.concatenate([[0], np.cumsum(dx)]), rtol=0, atol=1e-14)
    om = fi["angvel_internal"]
    vyraw = fi["vY"] + fi["vCoM_internal_y"] + om * fi["rX"]
    assert np.allclose(vyraw, -A * (2 * np.pi / T) * np.cos(arg), rtol=0, atol=1e-14)


@pytest.mark.parametrize("t", [0.0, 0.4, 2.3])
def test_midline_frame_invariants(oracle_mod, t):
    o = _fish_sim(oracle_mod)
    o.reset()
    o.set_scalars(time=t)
    o.put_objects(0.0)
    fi = o.fish_info()
    rX, rY, vX, vY = fi["rX"], fi["rY"], fi["vX"], fi["vY"]
    seg = np.hypot(np.diff(rX), np.diff(rY))
    assert np.allclose(seg, np.diff(fi["rS"]), rtol=1e-11)
    # the normals are unit vectors orthogonal to the midline
    nX, nY = fi["norX"], fi["norY"]
    assert np.allclose(np.hypot(nX, nY), 1, rtol=1e-11)
    assert np.allclose(nX[:-1] * np.diff(rX) + nY[:-1] * np.diff(rY), 0, atol=1e-15)
    # zero internal linear and angular momentum of the line density 2w
    s, w = fi["rS"], fi["width"]
    m = 2 * w
    P = np.array([np.trapezoid(m * vX, s), np.trapezoid(m * vY, s)])
    Mang = np.trapezoid(m * (rX * vY - rY * vX), s)
    scale = np.trapezoid(m * np.hypot(vX, vY), s) + 1e-30
    # the reference integrates the section's second moments too (fac2, fac3);
    # those terms are O(w^2) of the leading ones
    assert np.abs(P).max() < 2e-2 * scale and abs(Mang) < 2e-2 * scale * 0.2
    assert abs(fi["area_internal"] - np.trapezoid(m, s)) < 1e-15


@pytest.mark.parametrize("profile", ["carling", "naca"])
def test_chi_area_converges(oracle_mod, profile):
    """(FishData requires dSref > 0, FishLibrary.cpp:21: L = 0.2 needs at least
    16x8 blocks on the unit domain.)"""
    errs = []
    for bp in [(16, 8), (32, 16), (64, 32)]:
        o = _fish_sim(oracle_mod, bp=bp, profile=profile)
        o.reset()
        chi = o.get_field(oracle_mod.CHI)[0]
        assert chi.min() >= 0 and chi.max() <= 1 + 1e-12
        h = o.scalars()["h"]
        fi = o.fish_info()
        area = np.trapezoid(2 * fi["width"], fi["rS"])
        errs.append(abs(chi.sum() * h * h - area) / area)
        assert abs(o.shape()["M"] - chi.sum() * h * h) < 1e-12 * area
    assert errs[-1] < 0.01 and errs[2] < errs[0], errs


def test_udef_momenta_removed(oracle_mod):
    o = _fish_sim(oracle_mod, envelope="quadratic", profile="naca")
    o.reset()
    o.set_scalars(time=0.61)
    o.put_objects(0.0)
    chi = o.get_field(oracle_mod.CHI)[0]
    ud = o.get_field(oracle_mod.UDEF)
    h = o.scalars()["h"]
    st = o.shape()
    NY, NX = chi.shape
    x = (np.arange(NX) + 0.5) * h - st["centerOfMass"][0]
    y = (np.arange(NY) + 0.5) * h - st["centerOfMass"][1]
    X, Y = np.meshgrid(x, y)
    scale = (chi * np.hypot(ud[0], ud[1])).sum()
    assert scale > 0
    assert abs((chi * ud[0]).sum()) < 1e-12 * scale
    assert abs((chi * ud[1]).sum()) < 1e-12 * scale
    assert abs((chi * (X * ud[1] - Y * ud[0])).sum()) < 1e-12 * scale * 0.2
    # the centre of mass is the chi-weighted centroid
    assert abs((chi * X).sum()) < 1e-12 * chi.sum() and abs((chi * Y).sum()) < 1e-12 * chi.sum()


def test_straight_fish_is_mirror_symmetric(oracle_mod):
    """At t = 0 the ramp is 0: a straight fish along x at a cell face is
    symmetric about its axis."""
    o = _fish_sim(oracle_mod, bp=(16, 8))
    o.reset()
    chi = o.get_field(oracle_mod.CHI)[0]
    NY = chi.shape[0]
    assert np.abs(chi - chi[::-1, :]).max() < 1e-12
    assert NY % 2 == 0


def test_swimmer_moves_forward(oracle_mod):
    """A free CarlingFish starts to swim head first (towards -x for this
    midline orientation) once it undulates."""
    o = _fish_sim(oracle_mod, bp=(16, 8))
    o.reset()
    for _ in range(60):
        o.advance(o.calc_max_timestep())
    st = o.shape()
    assert np.isfinite(o.get_field(0)).all()
    assert st["center"][0] != 0.5


def test_reference_arithmetic_gap(oracle_mod):
    """The swimmer path in the device library's arithmetic (det_sincos,
    fixed-order midline sums and integrals, fixed-point hat P2M; what the
    device reproduces bit for bit) against the reference's own (libm
    sine/cosine, index-order sums, sequential P2M): the gap is rounding,
    measured here on the placement fields and after 30 whole steps."""
    def run(ref):
        oracle_mod.set_reference_arithmetic(ref)
        try:
            o = _fish_sim(oracle_mod)
            o.reset()
            chi = o.get_field(oracle_mod.CHI)[0].copy()
            ud = o.get_field(oracle_mod.UDEF).copy()
            for _ in range(30):
                o.advance(o.calc_max_timestep())
            vel = o.get_field(oracle_mod.VEL).copy()
            pres = o.get_field(oracle_mod.PRES)[0].copy()
            drag = o.shape(0)["drag"]
            del o
        finally:
            oracle_mod.set_reference_arithmetic(False)
        return chi, ud, vel, pres, drag

    a, b = run(False), run(True)
    rel = [np.linalg.norm(x - y) / np.linalg.norm(y) for x, y in zip(a[:4], b[:4])]
    dd = abs(a[4] - b[4]) / abs(b[4])
    print(f"  device vs reference arithmetic: chi {rel[0]:.1e}, udef {rel[1]:.1e} after "
          f"placement; after 30 steps vel {rel[2]:.1e}, pres {rel[3]:.1e}, drag {dd:.1e}")
    # measured: chi 7e-15, udef 2e-15; after 30 steps vel 1.5e-15, pres 8.7e-14,
    # drag 1.3e-14 (512x256 cells, free-space)
    assert rel[0] < 1e-13 and rel[1] < 1e-13
    assert rel[2] < 1e-12 and rel[3] < 1e-11 and dd < 1e-11
