"""GPU parity: libcup2d_hip.so (through its C-ABI) against the oracle.

Bar: stencil cell values bit-exact (same IEEE operation order, no
contraction); FFT-based results and anything downstream of a reduction
within relative L2 1e-10 (velocity, pressure) after N steps and 1e-8 for
drag, the tolerances stated for the FP64 path.
"""
import os
import sys

import numpy as np
import pytest

import numpy_ref as R

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[gpu case] {request.node.name}", flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def test_layout_roundtrip(cup):
    rng = np.random.default_rng(10)
    g = cup.Simulation(3, 2)
    d = rng.standard_normal((2, g.NY, g.NX))
    g.set_field(cup.VEL, d)
    assert np.array_equal(g.get_field(cup.VEL), d)
    blocks = g.get_field(cup.VEL, dense=False)
    import oracle
    assert np.array_equal(oracle.blocks_to_dense(blocks, 3, 2, 2), d)
    g.set_field(cup.PRES, oracle.dense_to_blocks(d[:1], 3, 2), dense=False)
    assert np.array_equal(g.get_field(cup.PRES)[0], d[0])


# strip widths 32 (half a wave), 64, 128 and 256 threads: the x neighbours
# come from adjacent lanes, the wave's edge lanes load theirs
@pytest.mark.parametrize("bp", [(3, 2), (10, 3), (4, 2), (8, 4)])
def test_advdiff_bitexact(cup, oracle_mod, bp):
    rng = np.random.default_rng(11)
    kw = dict(extent=1.0, nu=3e-3)
    g = cup.Simulation(*bp, **kw)
    o = oracle_mod.Oracle(*bp, **kw)
    vel = 0.3 * rng.standard_normal((2, g.NY, g.NX))
    for sim in (g, o):
        sim.set_field(0, vel)
        sim.set_scalars(uinf=(0.17, -0.08))
        sim.advdiff(1.3e-3)
    a, b = g.get_field(cup.VEL), o.get_field(oracle_mod.VEL)
    inner = (slice(None), slice(1, -1), slice(1, -1))
    assert np.array_equal(a[inner], b[inner])
    assert np.abs(a - b).max() < 1e-15


def test_advdiff_golden(cup):
    gd = np.load(os.path.join(GOLD, "advdiff_96x64.npz"))
    NY, NX = gd["u"].shape
    g = cup.Simulation(NX // 32, NY // 32, extent=1.0, nu=float(gd["nu"]))
    g.set_field(cup.VEL, np.stack([gd["u"], gd["v"]]))
    g.set_scalars(uinf=tuple(gd["uinf"]))
    g.advdiff(float(gd["dt"]))
    out = g.get_field(cup.VEL)
    assert np.array_equal(out[0][1:-1, 1:-1], gd["u1"][1:-1, 1:-1])
    assert np.abs(out[1] - gd["v1"]).max() < 1e-15


@pytest.mark.parametrize("bp", [(3, 2), (10, 3), (4, 2), (8, 4)])
def test_rhs_projection_bitexact(cup, bp):
    rng = np.random.default_rng(12)
    g = cup.Simulation(*bp, extent=1.0)
    h = g.scalars()["h"]
    u, v = rng.standard_normal((2, g.NY, g.NX))
    p = rng.standard_normal((g.NY, g.NX))
    dt = 2e-3
    g.set_field(cup.VEL, np.stack([u, v]))
    g.pressure_rhs(dt)
    assert np.array_equal(g.get_field(cup.TMP)[0], R.pressure_rhs(u, v, dt, h))
    g.set_field(cup.PRES, p[None])
    g.pressure_correction(dt)
    u1, v1 = R.pressure_correction(u, v, p, dt, h)
    out = g.get_field(cup.VEL)
    assert np.array_equal(out[0], u1) and np.array_equal(out[1], v1)


def test_find_max_u(cup, oracle_mod):
    rng = np.random.default_rng(13)
    g = cup.Simulation(4, 4)
    o = oracle_mod.Oracle(4, 4)
    vel = rng.standard_normal((2, g.NY, g.NX))
    for sim in (g, o):
        sim.set_field(0, vel)
        sim.set_scalars(uinf=(0.4, -0.2))
    assert g.find_max_u() == o.find_max_u()


@pytest.mark.parametrize("bp", [(4, 2), (8, 8), (16, 4), (64, 32)])
def test_poisson_periodic(cup, oracle_mod, bp):
    rng = np.random.default_rng(14)
    g = cup.Simulation(*bp, poisson="periodic")
    r = rng.standard_normal((g.NY, g.NX))
    pg = g.poisson_dense(r)
    pn = R.poisson_periodic(r)
    assert rel(pg, pn) < 1e-13
    if bp[0] * bp[1] <= 64:
        o = oracle_mod.Oracle(*bp, poisson="periodic")
        assert rel(pg, o.poisson_dense(r)) < 1e-13


def test_poisson_periodic_golden_and_mode(cup):
    gd = np.load(os.path.join(GOLD, "periodic_poisson_128x64.npz"))
    g = cup.Simulation(4, 2, poisson="periodic")
    assert rel(g.poisson_dense(gd["rhs"]), gd["p"]) < 1e-13
    x, y = np.meshgrid(np.arange(g.NX), np.arange(g.NY))
    for kx, ky in [(0, 1), (g.NX // 2, 0), (g.NX // 2, g.NY // 2), (7, 5)]:
        r = np.cos(2 * np.pi * (kx * x / g.NX + ky * y / g.NY))
        lam = (2 * np.pi * kx / g.NX) ** 2 + (2 * np.pi * ky / g.NY) ** 2
        assert np.abs(g.poisson_dense(r) + r / lam).max() < 1e-12 / lam


def test_poisson_freespace_direct(cup):
    gd = np.load(os.path.join(GOLD, "freespace_direct_64x32.npz"))
    g = cup.Simulation(2, 1, extent=1.0, poisson="freespace")
    assert rel(g.poisson_dense(gd["rhs"]), gd["p"]) < 1e-13


@pytest.mark.parametrize("bp", [(4, 4), (16, 8), (32, 16)])
def test_poisson_freespace_vs_oracle(cup, oracle_mod, bp):
    rng = np.random.default_rng(15)
    g = cup.Simulation(*bp, extent=1.0, poisson="freespace")
    o = oracle_mod.Oracle(*bp, extent=1.0, poisson="freespace", fs_pad2=False)
    r = rng.standard_normal((g.NY, g.NX))
    assert rel(g.poisson_dense(r), o.poisson_dense(r)) < 1e-12


@pytest.mark.parametrize("poisson", ["periodic", "freespace"])
def test_poisson_solve_with_fade(cup, oracle_mod, poisson):
    rng = np.random.default_rng(16)
    kw = dict(extent=1.0, poisson=poisson, fadeLen=0.03)
    g = cup.Simulation(8, 4, **kw)
    o = oracle_mod.Oracle(8, 4, fs_pad2=True, **kw)
    r = rng.standard_normal((1, g.NY, g.NX))
    for sim in (g, o):
        sim.set_field(3, r)
        sim.poisson_solve()
    assert np.abs(g.get_field(cup.TMP) - o.get_field(oracle_mod.TMP)).max() < 1e-14
    assert rel(g.get_field(cup.PRES), o.get_field(oracle_mod.PRES)) < 1e-12


def _disk_pair(cup, oracle_mod, bp, shapes, **kw):
    g = cup.Simulation(*bp, shapes=[cup.disk(**s) for s in shapes], **kw)
    o = oracle_mod.Oracle(*bp, fs_pad2=True, shapes=[oracle_mod.disk(**s) for s in shapes], **kw)
    return g, o


def test_put_objects_reset(cup, oracle_mod):
    shapes = [dict(radius=0.07, xpos=0.31, ypos=0.27, xvel=0.2, bFixed=True, bForced=True)]
    g, o = _disk_pair(cup, oracle_mod, (8, 4), shapes, extent=1.0, nu=1e-3, poisson="freespace")
    g.reset()
    o.reset()
    assert np.array_equal(g.get_field(cup.CHI), o.get_field(oracle_mod.CHI))
    assert np.array_equal(g.get_field(cup.UDEF), o.get_field(oracle_mod.UDEF))
    assert np.abs(g.get_field(cup.INVRHO) - o.get_field(oracle_mod.INVRHO)).max() == 0
    assert np.array_equal(g.get_field(cup.VEL), o.get_field(oracle_mod.VEL))
    sg, so = g.shape(), o.shape()
    for k in ["M", "J"]:
        assert abs(sg[k] - so[k]) <= 1e-14 * abs(so[k])
    for k in ["centerOfMass", "center"]:
        assert np.abs(np.array(sg[k]) - np.array(so[k])).max() < 1e-15
    assert sg["n_obstacle_blocks"] == so["n_obstacle_blocks"]
    assert g.scalars()["uinf"] == o.scalars()["uinf"]


def _run_and_compare(g, o, nsteps, nshapes=1, device_loop=False, tol=1e-10, drag_tol=1e-8):
    g.reset()
    o.reset()
    if device_loop:
        g.simulate(nsteps)
        o.simulate(nsteps)
    else:
        for _ in range(nsteps):
            dtg, dto = g.calc_max_timestep(), o.calc_max_timestep()
            assert abs(dtg - dto) <= 1e-13 * dto
            g.advance(dtg)
            o.advance(dto)
    sg, so = g.scalars(), o.scalars()
    assert sg["step"] == so["step"] == nsteps
    assert abs(sg["time"] - so["time"]) <= 1e-12 * so["time"]
    ev = rel(g.get_field(0), o.get_field(0))
    ep = rel(g.get_field(2), o.get_field(2))
    print(f"  rel L2 vel {ev:.3e} pres {ep:.3e}", flush=True)
    assert ev < tol and ep < tol
    for k in range(nshapes):
        a, b = g.shape(k), o.shape(k)
        for key in ["drag", "forcex", "forcey", "thrust", "perimeter"]:
            scale = max(abs(b[key]), abs(b["forcex"]), 1e-12)
            assert abs(a[key] - b[key]) <= drag_tol * scale, (key, a[key], b[key])
        for key in ["u", "v", "omega"]:
            assert abs(a[key] - b[key]) <= 1e-10 * max(abs(b[key]), 1e-3), (key, a[key], b[key])


def test_towed_cylinder_freespace_steps(cup, oracle_mod):
    shapes = [dict(radius=0.05, xpos=0.25, ypos=0.25, xvel=0.2, bFixed=True, bForced=True)]
    g, o = _disk_pair(cup, oracle_mod, (16, 8), shapes, extent=1.0, nu=2e-5, CFL=0.1,
                      poisson="freespace")
    _run_and_compare(g, o, 20)


def test_towed_cylinder_device_loop(cup, oracle_mod):
    shapes = [dict(radius=0.05, xpos=0.25, ypos=0.25, xvel=0.2, bFixed=True, bForced=True)]
    g, o = _disk_pair(cup, oracle_mod, (16, 8), shapes, extent=1.0, nu=2e-5, CFL=0.1,
                      poisson="freespace")
    _run_and_compare(g, o, 20, device_loop=True)


def test_free_disks_shared_blocks(cup, oracle_mod):
    """Two free, heavier disks sharing obstacle blocks (shape-ordered grid writes,
    momentum solve, collision sums)."""
    shapes = [dict(radius=0.06, xpos=0.40, ypos=0.25, rhoS=1.5),
              dict(radius=0.05, xpos=0.53, ypos=0.27, rhoS=2.0, xvel=0.1, bForcedx=True,
                   bForcedy=False, bBlockAng=False)]
    g, o = _disk_pair(cup, oracle_mod, (8, 4), shapes, extent=1.0, nu=1e-3, CFL=0.2,
                      poisson="freespace")
    rng = np.random.default_rng(17)
    g.reset()
    o.reset()
    vel = 0.05 * rng.standard_normal((2, g.NY, g.NX))
    for sim in (g, o):
        sim.set_field(0, vel)
    for _ in range(8):
        dtg, dto = g.calc_max_timestep(), o.calc_max_timestep()
        g.advance(dtg)
        o.advance(dto)
    assert rel(g.get_field(0), o.get_field(0)) < 1e-10
    assert rel(g.get_field(2), o.get_field(2)) < 1e-10
    for k in range(2):
        a, b = g.shape(k), o.shape(k)
        for key in ["u", "v", "omega", "centerOfMass"]:
            assert np.allclose(a[key], b[key], rtol=1e-9, atol=1e-13), (k, key, a[key], b[key])


def test_periodic_random_field_with_disk(cup, oracle_mod):
    shapes = [dict(radius=0.04, xpos=0.6, ypos=0.25, xvel=0.1, bFixed=True, bForced=True)]
    g, o = _disk_pair(cup, oracle_mod, (8, 4), shapes, extent=1.0, nu=1e-4, CFL=0.2,
                      poisson="periodic")
    rng = np.random.default_rng(18)
    g.reset()
    o.reset()
    vel = 0.05 * rng.standard_normal((2, g.NY, g.NX))
    for sim in (g, o):
        sim.set_field(0, vel)
    for _ in range(10):
        dtg, dto = g.calc_max_timestep(), o.calc_max_timestep()
        g.advance(dtg)
        o.advance(dto)
    assert rel(g.get_field(0), o.get_field(0)) < 1e-10
    assert rel(g.get_field(2), o.get_field(2)) < 1e-10


def test_body_out_of_domain_error(cup):
    g = cup.Simulation(2, 2, extent=1.0,
                       shapes=[cup.disk(0.05, 0.05, 0.5, xvel=5.0, bForced=True)])
    g.reset()
    with pytest.raises(cup.Cup2dError) as e:
        g.put_objects(1.0)
    assert e.value.code == cup.ERR_BODY_OUT


def test_bad_arguments(cup):
    with pytest.raises(cup.Cup2dError):
        cup.Simulation(0, 2)
    g = cup.Simulation(3, 2, poisson="periodic")  # 96 cells: no FFT plan
    with pytest.raises(cup.Cup2dError) as e:
        g.poisson_dense(np.zeros((g.NY, g.NX)))
    assert e.value.code == cup.ERR_ARG
    with pytest.raises(cup.Cup2dError):
        g.get_field(17)


@pytest.mark.parametrize("poisson", ["freespace", "periodic"])
def test_disk_on_boundary_blocks(cup, oracle_mod, poisson):
    """A forced disk whose obstacle blocks touch the domain faces and the
    cub2rhs fade band: the obstacle RHS, the deferred advDiff flux correction
    and the fade meet in the same blocks."""
    shapes = [dict(radius=0.05, xpos=0.055, ypos=0.06, xvel=0.15, yvel=0.05, bFixed=True,
                   bForced=True)]
    g, o = _disk_pair(cup, oracle_mod, (8, 4), shapes, extent=1.0, nu=1e-3, CFL=0.2,
                      poisson=poisson, fadeLen=0.05)
    rng = np.random.default_rng(19)
    g.reset()
    o.reset()
    vel = 0.05 * rng.standard_normal((2, g.NY, g.NX))
    for sim in (g, o):
        sim.set_field(0, vel)
    for _ in range(6):
        dtg, dto = g.calc_max_timestep(), o.calc_max_timestep()
        g.advance(dtg)
        o.advance(dto)
    ev = rel(g.get_field(0), o.get_field(0))
    ep = rel(g.get_field(2), o.get_field(2))
    print(f"  rel L2 vel {ev:.3e} pres {ep:.3e}", flush=True)
    assert ev < 1e-10 and ep < 1e-10
    a, b = g.shape(), o.shape()
    assert abs(a["drag"] - b["drag"]) <= 1e-8 * max(abs(b["drag"]), abs(b["forcex"]))


def test_pressure_iterator_towed_cylinder(cup, oracle_mod):
    """-iterativePenalization 1 (PressureIterator_unif): whole steps against the
    oracle; the iteration count of every step and the adaptive tolerance
    match, and so do the fields."""
    shapes = [dict(radius=0.08, xpos=0.3, ypos=0.25, xvel=0.2, bFixed=True, bForced=True)]
    g, o = _disk_pair(cup, oracle_mod, (8, 4), shapes, extent=1.0, nu=1e-3, CFL=0.1,
                      poisson="freespace", iterative_penal=True)
    g.reset()
    o.reset()
    for step in range(4):
        dtg, dto = g.calc_max_timestep(), o.calc_max_timestep()
        assert abs(dtg - dto) <= 1e-13 * dto
        g.advance(dtg)
        o.advance(dto)
    ev = rel(g.get_field(0), o.get_field(0))
    ep = rel(g.get_field(2), o.get_field(2))
    print(f"  rel L2 vel {ev:.3e} pres {ep:.3e}", flush=True)
    assert ev < 1e-10 and ep < 1e-10
    a, b = g.shape(), o.shape()
    assert abs(a["drag"] - b["drag"]) <= 1e-8 * max(abs(b["drag"]), abs(b["forcex"]))


def test_pressure_iterator_entry_point(cup, oracle_mod):
    """cup2d_pressure_iterator on the same state as the oracle's operator."""
    shapes = [dict(radius=0.06, xpos=0.4, ypos=0.25, rhoS=1.5)]
    g, o = _disk_pair(cup, oracle_mod, (8, 4), shapes, extent=1.0, nu=1e-3, CFL=0.2,
                      poisson="freespace")
    rng = np.random.default_rng(24)
    g.reset()
    o.reset()
    vel = 0.05 * rng.standard_normal((2, g.NY, g.NX))
    for sim in (g, o):
        sim.set_field(0, vel)
    n_g, r_g = g.pressure_iterator(1e-3)
    n_o = o.pressure_iterator(1e-3)
    st = o.iter_stats()
    assert n_g == n_o
    assert abs(r_g - st["relDF"]) <= 1e-9 * st["relDF"]
    assert rel(g.get_field(0), o.get_field(0)) < 1e-10
    assert rel(g.get_field(2), o.get_field(2)) < 1e-10
    assert rel(g.get_field(cup.VFLUID), o.get_field(oracle_mod.VFLUID)) < 1e-10


@pytest.mark.parametrize("bp", [(4, 2), (16, 8), (32, 16), (8, 8)])
def test_poisson_cosine_vs_oracle(cup, oracle_mod, bp):
    """Cosine (DCT) solver: the device runs the mirrored-grid FFT, the oracle the
    direct DCT sums pinned against scipy.fft.dct."""
    rng = np.random.default_rng(26)
    g = cup.Simulation(*bp, extent=1.0, poisson="cosine")
    o = oracle_mod.Oracle(*bp, extent=1.0, poisson="cosine")
    r = rng.standard_normal((g.NY, g.NX))
    pg, po = g.poisson_dense(r), o.poisson_dense(r)
    assert rel(pg, po) < 1e-12


def test_towed_cylinder_cosine(cup, oracle_mod):
    """Whole steps with the reference's default solver (poissonType "")."""
    shapes = [dict(radius=0.05, xpos=0.25, ypos=0.25, xvel=0.2, bFixed=True, bForced=True)]
    g, o = _disk_pair(cup, oracle_mod, (16, 8), shapes, extent=1.0, nu=2e-5, CFL=0.1,
                      poisson="cosine")
    _run_and_compare(g, o, 12)


# ------------------------------------------------------------------ swimmers
FISH_VARIANTS = {
    "carling": dict(L=0.2, T=1.0),
    "quadratic_naca": dict(L=0.2, T=1.0, envelope="quadratic", profile="naca", phi=0.25),
}


def _fish_pair(cup, oracle_mod, bp, fish, extent=1.0, poisson="freespace", **kw):
    kw = dict(extent=extent, nu=kw.pop("nu", 1e-4), CFL=kw.pop("CFL", 0.1), poisson=poisson, **kw)
    g = cup.Simulation(*bp, shapes=[cup.carlingfish(**f) for f in fish], **kw)
    o = oracle_mod.Oracle(*bp, fs_pad2=True, shapes=[oracle_mod.carlingfish(**f) for f in fish],
                          **kw)
    return g, o


@pytest.mark.parametrize("variant", sorted(FISH_VARIANTS))
@pytest.mark.parametrize("time", [0.0, 1.37])
def test_fish_midline_bitexact(cup, oracle_mod, variant, time):
    """Midline, normals, internal frame and momenta of the swimmer kernel equal
    the oracle's bit for bit (same operation sequence, same sin/cos, same
    summation order)."""
    fish = [dict(xpos=0.5, ypos=0.25, angle=17.0, **FISH_VARIANTS[variant])]
    g, o = _fish_pair(cup, oracle_mod, (16, 8), fish)
    g.reset()
    o.reset()
    if time:
        for sim in (g, o):
            sim.set_scalars(time=time)
            sim.put_objects(0.0)
    a, b = g.fish_midline(0), o.fish_info(0)
    assert a["Nm"] == b["Nm"]
    for k in cup.FISH_ARRAYS + cup.FISH_SCALARS:
        assert np.array_equal(a[k], b[k]), (k, np.max(np.abs(np.asarray(a[k]) - np.asarray(b[k]))))


@pytest.mark.parametrize("variant", sorted(FISH_VARIANTS))
def test_fish_put_objects(cup, oracle_mod, variant):
    """PutObjectsOnGrid with a swimmer (Fish::create + PutFishOnBlocks gathered
    per cell): chi and invRho bit-exact, udef after removeMoments within
    round-off of the momentum sums."""
    fish = [dict(xpos=0.47, ypos=0.26, angle=-23.0, **FISH_VARIANTS[variant])]
    g, o = _fish_pair(cup, oracle_mod, (16, 8), fish)
    g.reset()
    o.reset()
    for sim in (g, o):
        sim.set_scalars(time=0.83)
        sim.put_objects(0.0)
    cg, co = g.get_field(cup.CHI), o.get_field(oracle_mod.CHI)
    assert np.array_equal(cg, co), np.abs(cg - co).max()
    assert np.array_equal(g.get_field(cup.INVRHO), o.get_field(oracle_mod.INVRHO))
    ug, uo = g.get_field(cup.UDEF), o.get_field(oracle_mod.UDEF)
    assert np.abs(ug - uo).max() <= 1e-13 * np.abs(uo).max()
    sg, so = g.shape(), o.shape()
    for k in ["M", "J"]:
        assert abs(sg[k] - so[k]) <= 1e-13 * abs(so[k]), k
    for k in ["centerOfMass", "center"]:
        assert np.abs(np.array(sg[k]) - np.array(so[k])).max() < 1e-14
    assert sg["n_obstacle_blocks"] == so["n_obstacle_blocks"]
    assert sg["n_surface_points"] == so["n_surface_points"]


def test_free_swimmer_steps(cup, oracle_mod):
    fish = [dict(xpos=0.5, ypos=0.25, **FISH_VARIANTS["carling"])]
    g, o = _fish_pair(cup, oracle_mod, (16, 8), fish, nu=1e-4, CFL=0.2)
    _run_and_compare(g, o, 15)
    a, b = g.fish_midline(0), o.fish_info(0)
    assert abs(a["theta_internal"] - b["theta_internal"]) <= 1e-12


def test_swimmer_many_bodies_device_loop(cup, oracle_mod):
    """A swimmer and four towed disks whose obstacle blocks lie in more than
    a quarter of the block rows: the early row pass of the pressure solve is
    switched off on the device and the patch launch does the whole pass."""
    fish = dict(xpos=0.9, ypos=0.5, angle=0.0, **FISH_VARIANTS["quadratic_naca"])
    disks = [dict(radius=0.05, xpos=x, ypos=y, xvel=0.1, bFixed=True, bForced=True)
             for x, y in ((1.5, 0.15), (1.6, 0.35), (1.5, 0.65), (1.6, 0.85))]
    kw = dict(extent=2.0, nu=1e-4, CFL=0.2, poisson="periodic")
    g = cup.Simulation(32, 16, shapes=[cup.carlingfish(**fish)] + [cup.disk(**d) for d in disks],
                       **kw)
    o = oracle_mod.Oracle(32, 16, shapes=[oracle_mod.carlingfish(**fish)]
                          + [oracle_mod.disk(**d) for d in disks], **kw)
    _run_and_compare(g, o, 10, nshapes=5, device_loop=True)


def test_free_swimmer_device_loop_periodic(cup, oracle_mod):
    """The C2-style swimmer (quadratic envelope, NACA profile) in a periodic
    box next to a towed disk, advanced by the device loop."""
    fish = [dict(xpos=0.9, ypos=0.5, angle=8.0, **FISH_VARIANTS["quadratic_naca"])]
    g = cup.Simulation(32, 16, extent=2.0, nu=1e-4, CFL=0.2, poisson="periodic",
                       shapes=[cup.carlingfish(**fish[0]),
                               cup.disk(radius=0.05, xpos=1.5, ypos=0.4, xvel=0.1, bFixed=True,
                                        bForced=True)])
    o = oracle_mod.Oracle(32, 16, extent=2.0, nu=1e-4, CFL=0.2, poisson="periodic",
                          shapes=[oracle_mod.carlingfish(**fish[0]),
                                  oracle_mod.disk(radius=0.05, xpos=1.5, ypos=0.4, xvel=0.1,
                                                  bFixed=True, bForced=True)])
    _run_and_compare(g, o, 10, nshapes=2, device_loop=True)


@pytest.mark.parametrize("poisson", ["periodic", "freespace"])
def test_swimmer_in_fade_band_device_loop(cup, oracle_mod, poisson):
    """A swimmer whose obstacle blocks reach into the cub2rhs fade band
    (PoissonSolver.cpp:71-98: the outermost block row here), advanced by the
    device loop: the fade sums of band blocks are formed from the advDiff
    output beside the placement chain, and those of obstacle blocks again
    from their final RHS, which must give the oracle's solve."""
    fish = [dict(xpos=1.0, ypos=0.08, angle=0.0, **FISH_VARIANTS["carling"])]
    kw = dict(extent=2.0, nu=1e-4, CFL=0.2, poisson=poisson, fadeLen=0.05)
    g = cup.Simulation(32, 16, shapes=[cup.carlingfish(**fish[0])], **kw)
    o = oracle_mod.Oracle(32, 16, fs_pad2=True, shapes=[oracle_mod.carlingfish(**fish[0])], **kw)
    _run_and_compare(g, o, 8, device_loop=True)
