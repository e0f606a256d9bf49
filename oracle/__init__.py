"""Oracle: TEST INFRASTRUCTURE ONLY.

ctypes wrapper over ``libcup2d_oracle.so``, the FP64 C/OpenMP restatement of
the CubismUP_2D per-timestep hot path (see ``cup2d_oracle.c``).  Only the
tests, ``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py``
use it, and only as the checker / the timed CPU reference -- never as the
product path.

Fields are exchanged in the reference's block layout (blockID = bx + bpdx*by,
data[iy][ix][comp]); ``blocks_to_dense`` / ``dense_to_blocks`` convert to a
(comp, NY, NX) array.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libcup2d_oracle.so")
BS = 32
VEL, CHI, PRES, TMP, TMPV, UDEF, INVRHO, VFLUID = range(8)
NCOMP = [2, 1, 1, 1, 2, 2, 1, 2]


class Cfg(C.Structure):
    _fields_ = [("bpdx", C.c_int), ("bpdy", C.c_int), ("extent", C.c_double),
                ("nu", C.c_double), ("CFL", C.c_double), ("lam", C.c_double),
                ("fadeLen", C.c_double), ("poisson", C.c_int), ("fs_pad2", C.c_int),
                ("implicit_penal", C.c_int), ("checker", C.c_int),
                ("iterative_penal", C.c_int)]


class ShapeDesc(C.Structure):
    _fields_ = [("kind", C.c_int), ("radius", C.c_double), ("xpos", C.c_double),
                ("ypos", C.c_double), ("angle_deg", C.c_double), ("rhoS", C.c_double),
                ("bFixedx", C.c_int), ("bFixedy", C.c_int), ("bForcedx", C.c_int),
                ("bForcedy", C.c_int), ("bBlockAng", C.c_int), ("xvel", C.c_double),
                ("yvel", C.c_double), ("angvel", C.c_double), ("timeForced", C.c_double),
                ("tAccel", C.c_double), ("length", C.c_double), ("Tperiod", C.c_double),
                ("phaseShift", C.c_double), ("amplitudeFactor", C.c_double),
                ("envelope", C.c_int), ("env_a0", C.c_double), ("env_a2", C.c_double),
                ("waveLength", C.c_double), ("profile", C.c_int), ("naca_t", C.c_double)]


SHAPE_FIELDS = ["center", "centerOfMass", "labCenterOfMass", "d_gm", "orientation",
                "M", "J", "u", "v", "omega",
                "fluidAngMom", "fluidMomX", "fluidMomY", "penalDX", "penalDY", "penalM", "penalJ",
                "perimeter", "forcex", "forcey", "forcex_P", "forcey_P", "forcex_V", "forcey_V",
                "torque", "torque_P", "torque_V", "drag", "thrust", "circulation",
                "Pout", "PoutBnd", "defPower", "defPowerBnd", "Pthrust", "Pdrag", "EffPDef",
                "EffPDefBnd", "n_obstacle_blocks", "n_surface_points"]


class ShapeState(C.Structure):
    _fields_ = [(n, C.c_double * 2) if n in ("center", "centerOfMass", "labCenterOfMass", "d_gm")
                else (n, C.c_double) for n in SHAPE_FIELDS]

    def as_dict(self):
        out = {}
        for n in SHAPE_FIELDS:
            v = getattr(self, n)
            out[n] = tuple(v) if not isinstance(v, float) else v
        return out


def build(force=False):
    if force or not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", _HERE])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        P = C.c_void_p
        D = C.c_double
        dp = C.POINTER(C.c_double)
        L.orc_create.restype = P
        L.orc_create.argtypes = [C.POINTER(Cfg)]
        L.orc_destroy.argtypes = [P]
        L.orc_add_shape.argtypes = [P, C.POINTER(ShapeDesc)]
        L.orc_set_field.argtypes = [P, C.c_int, dp]
        L.orc_get_field.argtypes = [P, C.c_int, dp]
        L.orc_get_scalars.argtypes = [P, dp, dp, C.POINTER(C.c_int), dp, dp]
        L.orc_set_scalars.argtypes = [P, D, C.c_int, D, D]
        L.orc_get_shape.argtypes = [P, C.c_int, C.POINTER(ShapeState)]
        for n in ["orc_reset", "orc_apply_obj_vel", "orc_cub2rhs", "orc_poisson_solve",
                  "orc_compute_forces"]:
            getattr(L, n).argtypes = [P]
        for n in ["orc_advdiff", "orc_pressure_rhs", "orc_pressure_correction",
                  "orc_body_update", "orc_penalize", "orc_pressure_single"]:
            getattr(L, n).argtypes = [P, D]
        L.orc_put_objects.argtypes = [P, D]
        L.orc_advance.argtypes = [P, D]
        L.orc_simulate.argtypes = [P, C.c_int]
        L.orc_find_max_u.argtypes = [P]
        L.orc_find_max_u.restype = D
        L.orc_calc_max_timestep.argtypes = [P]
        L.orc_calc_max_timestep.restype = D
        L.orc_poisson_dense.argtypes = [P, dp, dp]
        L.orc_pressure_iterator.argtypes = [P, D]
        L.orc_pressure_iterator.restype = C.c_int
        L.orc_iter_stats.argtypes = [P, C.POINTER(C.c_int), dp, dp]
        L.orc_set_collisions.argtypes = [P, C.c_int]
        L.orc_set_surface_velocity.argtypes = [P, C.c_int, C.c_int, C.POINTER(C.c_double)]
        L.orc_set_surface_velocity.restype = C.c_int
        L.orc_fft1d.argtypes = [dp, C.c_int, C.c_int]
        L.orc_get_timing.argtypes = [P, dp, dp]
        L.orc_num_threads.restype = C.c_int
        L.orc_fish_info.argtypes = [P, C.c_int, dp]
        L.orc_fish_info.restype = C.c_int
        L.orc_det_sincos.argtypes = [D, dp, dp]
        L.orc_lane_sum.argtypes = [dp, C.c_int]
        L.orc_lane_sum.restype = D
        L.orc_set_reference_arithmetic.argtypes = [C.c_int]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def blocks_to_dense(a, bpdx, bpdy, nc):
    """(nblocks*32*32*nc) block layout -> (nc, NY, NX)."""
    a = np.asarray(a).reshape(bpdy, bpdx, BS, BS, nc)
    return np.ascontiguousarray(a.transpose(4, 0, 2, 1, 3).reshape(nc, bpdy * BS, bpdx * BS))


def dense_to_blocks(d, bpdx, bpdy):
    """(nc, NY, NX) -> block layout flat array."""
    d = np.asarray(d, dtype=np.float64)
    nc = d.shape[0]
    a = d.reshape(nc, bpdy, BS, bpdx, BS).transpose(1, 3, 2, 4, 0)
    return np.ascontiguousarray(a).reshape(-1)


def disk(radius, xpos, ypos, xvel=0.0, yvel=0.0, angvel=0.0, rhoS=1.0, angle=0.0,
         bFixed=False, bForced=False, bFixedx=None, bFixedy=None, bForcedx=None,
         bForcedy=None, bBlockAng=None, timeForced=1.7976931348623157e308, tAccel=-1.0):
    """Shape descriptor with the reference's flag defaults (Shape.cpp:331-345)."""
    bFixedx = bFixed if bFixedx is None else bFixedx
    bFixedy = bFixed if bFixedy is None else bFixedy
    bForcedx = bForced if bForcedx is None else bForcedx
    bForcedy = bForced if bForcedy is None else bForcedy
    bBlockAng = (bForcedx or bForcedy) if bBlockAng is None else bBlockAng
    return dict(kind=0, radius=radius, xpos=xpos, ypos=ypos, angle_deg=angle, rhoS=rhoS,
                bFixedx=int(bFixedx), bFixedy=int(bFixedy), bForcedx=int(bForcedx),
                bForcedy=int(bForcedy), bBlockAng=int(bBlockAng), xvel=xvel, yvel=yvel,
                angvel=angvel, timeForced=timeForced, tAccel=tAccel)


def carlingfish(xpos, ypos, L=0.1, T=1.0, phi=0.0, angle=0.0, amplitudeFactor=1.0,
                envelope="carling", env_a0=0.02, env_a2=0.08, waveLength=1.0,
                profile="carling", naca_t=0.08, rhoS=1.0, bFixed=False, bForced=False,
                bFixedx=None, bFixedy=None, bForcedx=None, bForcedy=None, bBlockAng=None,
                xvel=0.0, yvel=0.0, angvel=0.0, timeForced=1.7976931348623157e308):
    """CarlingFish descriptor (Fish.h:27-29, CarlingFish.cpp:90-95) with the
    reference's flag defaults (Shape.cpp:331-345).  envelope "carling" is
    4/33 (s + L/32); "quadratic" is L (env_a0 + env_a2 (s/L)^2) with wavelength
    waveLength*L.  profile "carling" is AmplitudeFish::_width, "naca" a NACA
    4-digit half-thickness of ratio naca_t."""
    d = disk(0.0, xpos, ypos, xvel=xvel, yvel=yvel, angvel=angvel, rhoS=rhoS, angle=angle,
             bFixed=bFixed, bForced=bForced, bFixedx=bFixedx, bFixedy=bFixedy,
             bForcedx=bForcedx, bForcedy=bForcedy, bBlockAng=bBlockAng,
             timeForced=timeForced)
    d.update(kind=1, length=L, Tperiod=T, phaseShift=phi, amplitudeFactor=amplitudeFactor,
             envelope={"carling": 0, "quadratic": 1}[envelope], env_a0=env_a0, env_a2=env_a2,
             waveLength=waveLength, profile={"carling": 0, "naca": 1}[profile], naca_t=naca_t)
    return d


FISH_ARRAYS = ["rS", "rX", "rY", "vX", "vY", "norX", "norY", "vNorX", "vNorY", "width"]
FISH_SCALARS = ["area_internal", "J_internal", "CoM_internal_x", "CoM_internal_y",
                "vCoM_internal_x", "vCoM_internal_y", "theta_internal", "angvel_internal"]


def det_sincos(x):
    a, b = C.c_double(), C.c_double()
    lib().orc_det_sincos(x, C.byref(a), C.byref(b))
    return a.value, b.value


def lane_sum(v):
    v = np.ascontiguousarray(v, dtype=np.float64)
    return lib().orc_lane_sum(_dp(v), len(v))


def set_reference_arithmetic(on):
    """Process-wide: True runs the swimmer path in the reference's own
    arithmetic (libm sine/cosine, index-order sums, sequential hat P2M);
    False (the default) in the device library's, bit for bit."""
    lib().orc_set_reference_arithmetic(1 if on else 0)


class Oracle:
    """One simulation in the oracle, same operator sequence as the reference."""

    def __init__(self, bpdx, bpdy, extent=1.0, nu=1e-2, CFL=0.1, lam=None, fadeLen=0.01,
                 poisson="periodic", fs_pad2=False, implicit_penal=False, checker=False,
                 iterative_penal=False, shapes=()):
        self.L = lib()
        lam = 1e3 / CFL if lam is None else lam
        cfg = Cfg(bpdx, bpdy, extent, nu, CFL, lam, fadeLen,
                  {"periodic": 0, "freespace": 1, "cosine": 2}[poisson], int(fs_pad2),
                  int(implicit_penal),
                  int(checker), int(iterative_penal))
        self.h = self.L.orc_create(C.byref(cfg))
        if not self.h:
            raise ValueError("orc_create failed")
        self.bpdx, self.bpdy = bpdx, bpdy
        self.NX, self.NY = bpdx * BS, bpdy * BS
        for sd in shapes:
            self.add_shape(sd)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_destroy(self.h)
            self.h = None

    def add_shape(self, d):
        r = self.L.orc_add_shape(self.h, C.byref(ShapeDesc(**d)))
        if r < 0:
            raise ValueError("unsupported shape")
        return r

    def set_field(self, f, arr, dense=True):
        nc = NCOMP[f]
        a = dense_to_blocks(np.asarray(arr).reshape(nc, self.NY, self.NX), self.bpdx, self.bpdy) \
            if dense else np.ascontiguousarray(arr, dtype=np.float64).reshape(-1)
        self.L.orc_set_field(self.h, f, _dp(a))

    def get_field(self, f, dense=True):
        nc = NCOMP[f]
        a = np.empty(self.bpdx * self.bpdy * BS * BS * nc)
        self.L.orc_get_field(self.h, f, _dp(a))
        return blocks_to_dense(a, self.bpdx, self.bpdy, nc) if dense else a

    def scalars(self):
        t, dt, hh = C.c_double(), C.c_double(), C.c_double()
        st = C.c_int()
        u = (C.c_double * 2)()
        self.L.orc_get_scalars(self.h, C.byref(t), C.byref(dt), C.byref(st), u, C.byref(hh))
        return dict(time=t.value, dt=dt.value, step=st.value, uinf=(u[0], u[1]), h=hh.value)

    def set_scalars(self, time=0.0, step=0, uinf=(0.0, 0.0)):
        self.L.orc_set_scalars(self.h, time, step, uinf[0], uinf[1])

    def fish_info(self, k=0):
        """Midline arrays and internal-frame scalars of swimmer k."""
        Nm = self.L.orc_fish_info(self.h, k, None)
        if Nm < 0:
            raise ValueError("shape %d is not a swimmer" % k)
        out = np.empty(10 * Nm + 8)
        self.L.orc_fish_info(self.h, k, _dp(out))
        d = {n: out[i * Nm:(i + 1) * Nm].copy() for i, n in enumerate(FISH_ARRAYS)}
        d.update({n: float(out[10 * Nm + i]) for i, n in enumerate(FISH_SCALARS)})
        d["Nm"] = Nm
        return d

    def shape(self, i=0):
        s = ShapeState()
        self.L.orc_get_shape(self.h, i, C.byref(s))
        return s.as_dict()

    def reset(self):
        self.L.orc_reset(self.h)

    def find_max_u(self):
        return self.L.orc_find_max_u(self.h)

    def calc_max_timestep(self):
        return self.L.orc_calc_max_timestep(self.h)

    def put_objects(self, dt):
        return self.L.orc_put_objects(self.h, dt)

    def advdiff(self, dt):
        self.L.orc_advdiff(self.h, dt)

    def pressure_rhs(self, dt):
        self.L.orc_pressure_rhs(self.h, dt)

    def cub2rhs(self):
        self.L.orc_cub2rhs(self.h)

    def poisson_solve(self):
        self.L.orc_poisson_solve(self.h)

    def pressure_correction(self, dt):
        self.L.orc_pressure_correction(self.h, dt)

    def body_update(self, dt):
        self.L.orc_body_update(self.h, dt)

    def penalize(self, dt):
        self.L.orc_penalize(self.h, dt)

    def pressure_single(self, dt):
        self.L.orc_pressure_single(self.h, dt)

    def pressure_iterator(self, dt):
        return self.L.orc_pressure_iterator(self.h, dt)

    def iter_stats(self):
        n, r, t = C.c_int(), C.c_double(), C.c_double()
        self.L.orc_iter_stats(self.h, C.byref(n), C.byref(r), C.byref(t))
        return dict(iters=n.value, relDF=r.value, target=t.value)

    def compute_forces(self):
        self.L.orc_compute_forces(self.h)

    def set_collisions(self, enabled):
        """Test hook: with enabled=False the body update skips
        preventCollidingObstacles."""
        self.L.orc_set_collisions(self.h, int(enabled))

    def set_surface_velocity(self, coef, shape=0):
        """Fourier-mode surface velocity of a disk (the library's extension;
        coef = 4M amplitudes: normal cos, normal sin, tangential cos, sin)."""
        c = np.ascontiguousarray(np.asarray(coef, dtype=np.float64).ravel())
        assert self.L.orc_set_surface_velocity(self.h, shape, c.size // 4, _dp(c)) == 0

    def advance(self, dt):
        return self.L.orc_advance(self.h, dt)

    def simulate(self, nsteps):
        return self.L.orc_simulate(self.h, nsteps)

    def apply_obj_vel(self):
        self.L.orc_apply_obj_vel(self.h)

    def poisson_dense(self, rhs):
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        out = np.empty_like(rhs)
        self.L.orc_poisson_dense(self.h, _dp(rhs), _dp(out))
        return out

    def timing(self):
        a, b = C.c_double(), C.c_double()
        self.L.orc_get_timing(self.h, C.byref(a), C.byref(b))
        return a.value, b.value


def fft1d(x, sign=-1):
    x = np.ascontiguousarray(np.asarray(x, dtype=np.complex128))
    buf = x.view(np.float64).copy()
    lib().orc_fft1d(_dp(buf), x.size, sign)
    return buf.view(np.complex128)


def num_threads():
    return lib().orc_num_threads()
