"""cubismup_2d_amd -- MI355X (gfx950) implementation of CubismUP_2D's per-step hot path.

The product is ``libcup2d_hip.so`` (HIP kernels behind the C-ABI declared in
``include/cup2d.h``).  This module is a thin ctypes binding of that C-ABI; it
mirrors the reference's operator sequence (Simulation::reset /
calcMaxTimestep / advance and the operators of the PressureSingle pipeline)
under the same names.  There is no CPU fallback: if the shared library is
missing or no HIP device is present the calls raise.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcup2d_hip.so")
BS = 32

VEL, CHI, PRES, TMP, TMPV, UDEF, INVRHO, VFLUID = range(8)
NCOMP = [2, 1, 1, 1, 2, 2, 1, 2]
LAYOUT_BLOCKS, LAYOUT_DENSE = 0, 1
POISSON = {"periodic": 0, "freespace": 1, "cosine": 2}
PRECISION = {"fp64": 0, "fp32": 1}

OK = 0
ERR_ARG, ERR_HIP, ERR_STATE, ERR_BODY_OUT, ERR_CAPACITY, ERR_NONFINITE = -1, -2, -3, -4, -5, -6


class Cup2dError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"cup2d error {code}: {msg}")
        self.code = code


class Cfg(C.Structure):
    _fields_ = [("bpdx", C.c_int), ("bpdy", C.c_int), ("extent", C.c_double),
                ("nu", C.c_double), ("CFL", C.c_double), ("lam", C.c_double),
                ("fadeLen", C.c_double), ("poisson", C.c_int), ("implicit_penal", C.c_int),
                ("iterative_penal", C.c_int), ("stream", C.c_void_p),
                ("precision", C.c_int), ("refine_sweeps", C.c_int)]


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


FISH_ARRAYS = ["rS", "rX", "rY", "vX", "vY", "norX", "norY", "vNorX", "vNorY", "width"]
FISH_SCALARS = ["area_internal", "J_internal", "CoM_internal_x", "CoM_internal_y",
                "vCoM_internal_x", "vCoM_internal_y", "theta_internal", "angvel_internal"]

SHAPE_FIELDS = ["center", "centerOfMass", "labCenterOfMass", "d_gm", "orientation",
                "M", "J", "u", "v", "omega",
                "fluidAngMom", "fluidMomX", "fluidMomY", "penalDX", "penalDY", "penalM", "penalJ",
                "perimeter", "forcex", "forcey", "forcex_P", "forcey_P", "forcex_V", "forcey_V",
                "torque", "torque_P", "torque_V", "drag", "thrust", "circulation",
                "Pout", "PoutBnd", "defPower", "defPowerBnd", "Pthrust", "Pdrag", "EffPDef",
                "EffPDefBnd", "n_obstacle_blocks", "n_surface_points"]
_VEC2 = ("center", "centerOfMass", "labCenterOfMass", "d_gm")


class ShapeState(C.Structure):
    _fields_ = [(n, C.c_double * 2) if n in _VEC2 else (n, C.c_double) for n in SHAPE_FIELDS]

    def as_dict(self):
        return {n: (tuple(getattr(self, n)) if n in _VEC2 else getattr(self, n))
                for n in SHAPE_FIELDS}


EXPORTS = [
    "cup2d_create", "cup2d_destroy", "cup2d_last_error", "cup2d_add_shape",
    "cup2d_upload_field", "cup2d_download_field", "cup2d_field_device_ptr",
    "cup2d_get_scalars", "cup2d_set_scalars", "cup2d_get_shape", "cup2d_fish_midline",
    "cup2d_reset",
    "cup2d_find_max_u", "cup2d_calc_max_timestep", "cup2d_put_objects", "cup2d_advdiff",
    "cup2d_pressure_rhs", "cup2d_poisson_solve", "cup2d_pressure_correction",
    "cup2d_body_update", "cup2d_penalize", "cup2d_pressure_single", "cup2d_pressure_iterator",
    "cup2d_compute_forces",
    "cup2d_advance", "cup2d_simulate", "cup2d_synchronize", "cup2d_poisson_dense",
    "cup2d_grid_size", "cup2d_timer_enable", "cup2d_timer_read", "cup2d_timer_sample",
    "cup2d_force_log_enable", "cup2d_force_log_read",
    "cup2d_find_max_u_noalloc", "cup2d_simulate_graph", "cup2d_refine_residuals",
    "cup2d_set_surface_velocity", "cup2d_adjoint_forward", "cup2d_adjoint_backward",
    "cup2d_ensemble_create", "cup2d_ensemble_destroy", "cup2d_ensemble_last_error",
    "cup2d_ensemble_set_lanes", "cup2d_ensemble_simulate", "cup2d_ensemble_synchronize",
    "cup2d_ensemble_tune",
]
FORCE_LOG_COLUMNS = ["time", "step", "Fx", "Fy", "FxPres", "FyPres", "FxVisc", "FyVisc", "tau",
                     "tauPres", "tauVisc", "drag", "thrust", "perimeter", "circulation",
                     "Pthrust", "Pdrag", "PoutBnd", "Pout", "defPowerBnd", "defPower",
                     "EffPDefBnd", "EffPDef", "n_surface_points"]
TIMERS = ["find_max_u", "dt", "advdiff", "flux_fix", "pressure_rhs", "rhs_obstacles", "fade",
          "fft_rows_fwd", "fft_cols", "fft_rows_inv", "pressure_correction", "put_objects",
          "body_update", "penalize", "forces", "refine"]

_lib = None


def lib():
    """Load libcup2d_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build() first "
                          "(there is no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    P, D, I = C.c_void_p, C.c_double, C.c_int
    dp = C.POINTER(C.c_double)
    L.cup2d_create.argtypes = [C.POINTER(Cfg), C.POINTER(C.c_void_p)]
    L.cup2d_destroy.argtypes = [P]
    L.cup2d_last_error.argtypes = [P]
    L.cup2d_last_error.restype = C.c_char_p
    L.cup2d_add_shape.argtypes = [P, C.POINTER(ShapeDesc), C.POINTER(I)]
    L.cup2d_upload_field.argtypes = [P, I, I, dp]
    L.cup2d_download_field.argtypes = [P, I, I, dp]
    L.cup2d_field_device_ptr.argtypes = [P, I, I, C.POINTER(C.c_void_p)]
    L.cup2d_get_scalars.argtypes = [P, dp, dp, C.POINTER(I), dp, dp]
    L.cup2d_set_scalars.argtypes = [P, D, I, D, D]
    L.cup2d_get_shape.argtypes = [P, I, C.POINTER(ShapeState)]
    L.cup2d_fish_midline.argtypes = [P, I, C.POINTER(I), dp]
    for n in ["cup2d_reset", "cup2d_poisson_solve", "cup2d_compute_forces",
              "cup2d_synchronize"]:
        getattr(L, n).argtypes = [P]
    for n in ["cup2d_put_objects", "cup2d_advdiff", "cup2d_pressure_rhs",
              "cup2d_pressure_correction", "cup2d_body_update", "cup2d_penalize",
              "cup2d_pressure_single", "cup2d_advance"]:
        getattr(L, n).argtypes = [P, D]
    L.cup2d_pressure_iterator.argtypes = [P, D, C.POINTER(I), dp]
    L.cup2d_find_max_u.argtypes = [P, dp]
    L.cup2d_calc_max_timestep.argtypes = [P, dp]
    L.cup2d_simulate.argtypes = [P, I]
    L.cup2d_poisson_dense.argtypes = [P, dp, dp]
    L.cup2d_grid_size.argtypes = [P, C.POINTER(I), C.POINTER(I)]
    L.cup2d_timer_enable.argtypes = [P, C.c_char_p, I]
    L.cup2d_timer_read.argtypes = [P, C.c_char_p, dp, C.POINTER(I)]
    L.cup2d_timer_sample.argtypes = [P, C.c_char_p, I]
    L.cup2d_force_log_enable.argtypes = [P, I]
    L.cup2d_force_log_read.argtypes = [P, I, I, dp, C.POINTER(I)]
    L.cup2d_find_max_u_noalloc.argtypes = [P, P, C.c_longlong, D, D, P, P]
    L.cup2d_simulate_graph.argtypes = [P, I]
    L.cup2d_refine_residuals.argtypes = [P, dp, C.POINTER(I)]
    L.cup2d_set_surface_velocity.argtypes = [P, I, I, dp]
    L.cup2d_adjoint_forward.argtypes = [P, I, I, dp, dp, dp]
    L.cup2d_adjoint_backward.argtypes = [P, dp]
    L.cup2d_ensemble_create.argtypes = [C.POINTER(C.c_void_p), I, C.POINTER(C.c_void_p)]
    L.cup2d_ensemble_destroy.argtypes = [P]
    L.cup2d_ensemble_last_error.argtypes = [P]
    L.cup2d_ensemble_last_error.restype = C.c_char_p
    L.cup2d_ensemble_set_lanes.argtypes = [P, I]
    L.cup2d_ensemble_simulate.argtypes = [P, I]
    L.cup2d_ensemble_synchronize.argtypes = [P]
    L.cup2d_ensemble_tune.argtypes = [P, I, C.POINTER(I), dp]
    for n in EXPORTS:
        if n not in ("cup2d_last_error", "cup2d_ensemble_last_error"):
            getattr(L, n).restype = C.c_int
    _lib = L
    return L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def disk(radius, xpos, ypos, xvel=0.0, yvel=0.0, angvel=0.0, rhoS=1.0, angle=0.0,
         bFixed=False, bForced=False, bFixedx=None, bFixedy=None, bForcedx=None,
         bForcedy=None, bBlockAng=None, timeForced=1.7976931348623157e308, tAccel=-1.0):
    """Disk descriptor with the reference's flag defaults (Shape.cpp:331-345)."""
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
    """CarlingFish descriptor (-L -T -phi -amplitudeFactor, Fish.h:27-29,
    CarlingFish.cpp:90-95) with the reference's flag defaults
    (Shape.cpp:331-345).  envelope "carling" is the reference's lateral
    amplitude 4/33 (s + L/32) with wavelength L; "quadratic" is
    L (env_a0 + env_a2 (s/L)^2) with wavelength waveLength*L.  profile
    "carling" is AmplitudeFish::_width; "naca" is a NACA 4-digit half-thickness
    of thickness ratio naca_t."""
    d = disk(0.0, xpos, ypos, xvel=xvel, yvel=yvel, angvel=angvel, rhoS=rhoS, angle=angle,
             bFixed=bFixed, bForced=bForced, bFixedx=bFixedx, bFixedy=bFixedy,
             bForcedx=bForcedx, bForcedy=bForcedy, bBlockAng=bBlockAng, timeForced=timeForced)
    d.update(kind=1, length=L, Tperiod=T, phaseShift=phi, amplitudeFactor=amplitudeFactor,
             envelope={"carling": 0, "quadratic": 1}[envelope], env_a0=env_a0, env_a2=env_a2,
             waveLength=waveLength, profile={"carling": 0, "naca": 1}[profile], naca_t=naca_t)
    return d


class Simulation:
    """One simulation on the GPU; methods mirror the reference's operators."""

    def __init__(self, bpdx, bpdy, extent=1.0, nu=1e-2, CFL=0.1, lam=None, fadeLen=0.01,
                 poisson="periodic", implicit_penal=False, iterative_penal=False, shapes=(),
                 stream=None, precision="fp64", refine_sweeps=None):
        self.L = lib()
        lam = 1e3 / CFL if lam is None else lam
        prec = PRECISION[precision]
        if refine_sweeps is None:  # FP32: one FP64 sweep where the solver takes one
            refine_sweeps = 1 if prec == PRECISION["fp32"] and poisson != "freespace" else 0
        cfg = Cfg(bpdx, bpdy, extent, nu, CFL, lam, fadeLen, POISSON[poisson],
                  int(implicit_penal), int(iterative_penal), stream, prec, int(refine_sweeps))
        self.precision = precision
        self._surf_modes = 0
        self._adj_modes = 0
        h = C.c_void_p()
        rc = self.L.cup2d_create(C.byref(cfg), C.byref(h))
        if rc != OK:
            raise Cup2dError(rc, self.L.cup2d_last_error(None).decode())
        self.h = h
        self.bpdx, self.bpdy = bpdx, bpdy
        self.NX, self.NY = bpdx * BS, bpdy * BS
        for sd in shapes:
            self.add_shape(sd)

    def close(self):
        if getattr(self, "h", None):
            self.L.cup2d_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _chk(self, rc):
        if rc != OK:
            raise Cup2dError(rc, self.L.cup2d_last_error(self.h).decode())
        return rc

    def add_shape(self, d):
        i = C.c_int()
        self._chk(self.L.cup2d_add_shape(self.h, C.byref(ShapeDesc(**d)), C.byref(i)))
        return i.value

    # -- data
    def set_field(self, f, arr, dense=True):
        nc = NCOMP[f] if 0 <= f < len(NCOMP) else 1  # the C-ABI validates f
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(-1)
        assert a.size == nc * self.NX * self.NY
        self._chk(self.L.cup2d_upload_field(self.h, f, LAYOUT_DENSE if dense else LAYOUT_BLOCKS,
                                            _dp(a)))

    def get_field(self, f, dense=True):
        nc = NCOMP[f] if 0 <= f < len(NCOMP) else 1  # the C-ABI validates f
        a = np.empty(nc * self.NX * self.NY)
        self._chk(self.L.cup2d_download_field(self.h, f,
                                              LAYOUT_DENSE if dense else LAYOUT_BLOCKS, _dp(a)))
        return a.reshape(nc, self.NY, self.NX) if dense else a

    def device_ptr(self, f, comp=0):
        p = C.c_void_p()
        self._chk(self.L.cup2d_field_device_ptr(self.h, f, comp, C.byref(p)))
        return p.value

    def scalars(self):
        t, dt, hh = C.c_double(), C.c_double(), C.c_double()
        st = C.c_int()
        u = (C.c_double * 2)()
        self._chk(self.L.cup2d_get_scalars(self.h, C.byref(t), C.byref(dt), C.byref(st), u,
                                           C.byref(hh)))
        return dict(time=t.value, dt=dt.value, step=st.value, uinf=(u[0], u[1]), h=hh.value)

    def set_scalars(self, time=0.0, step=0, uinf=(0.0, 0.0)):
        self._chk(self.L.cup2d_set_scalars(self.h, time, step, uinf[0], uinf[1]))

    def fish_midline(self, i=0):
        """Midline arrays and internal-frame scalars of swimmer i (same keys as
        the oracle's fish_info)."""
        n = C.c_int()
        self._chk(self.L.cup2d_fish_midline(self.h, i, C.byref(n), None))
        Nm = n.value
        out = np.empty(10 * Nm + 8)
        self._chk(self.L.cup2d_fish_midline(self.h, i, C.byref(n),
                                            out.ctypes.data_as(C.POINTER(C.c_double))))
        d = {k: out[j * Nm:(j + 1) * Nm].copy() for j, k in enumerate(FISH_ARRAYS)}
        d.update({k: float(out[10 * Nm + j]) for j, k in enumerate(FISH_SCALARS)})
        d["Nm"] = Nm
        return d

    def shape(self, i=0):
        s = ShapeState()
        self._chk(self.L.cup2d_get_shape(self.h, i, C.byref(s)))
        return s.as_dict()

    # -- operators (same names as the oracle wrapper)
    def reset(self):
        self._chk(self.L.cup2d_reset(self.h))

    def find_max_u(self):
        m = C.c_double()
        self._chk(self.L.cup2d_find_max_u(self.h, C.byref(m)))
        return m.value

    def calc_max_timestep(self):
        m = C.c_double()
        self._chk(self.L.cup2d_calc_max_timestep(self.h, C.byref(m)))
        return m.value

    def put_objects(self, dt):
        return self._chk(self.L.cup2d_put_objects(self.h, dt))

    def advdiff(self, dt):
        self._chk(self.L.cup2d_advdiff(self.h, dt))

    def pressure_rhs(self, dt):
        self._chk(self.L.cup2d_pressure_rhs(self.h, dt))

    def poisson_solve(self):
        self._chk(self.L.cup2d_poisson_solve(self.h))

    def pressure_correction(self, dt):
        self._chk(self.L.cup2d_pressure_correction(self.h, dt))

    def body_update(self, dt):
        self._chk(self.L.cup2d_body_update(self.h, dt))

    def penalize(self, dt):
        self._chk(self.L.cup2d_penalize(self.h, dt))

    def pressure_single(self, dt):
        self._chk(self.L.cup2d_pressure_single(self.h, dt))

    def pressure_iterator(self, dt):
        """PressureIterator_unif::operator(); returns (iterations, relDF)."""
        n, r = C.c_int(), C.c_double()
        self._chk(self.L.cup2d_pressure_iterator(self.h, dt, C.byref(n), C.byref(r)))
        return n.value, r.value

    def compute_forces(self):
        self._chk(self.L.cup2d_compute_forces(self.h))

    def advance(self, dt):
        return self._chk(self.L.cup2d_advance(self.h, dt))

    def simulate(self, nsteps, sync=True):
        self._chk(self.L.cup2d_simulate(self.h, nsteps))
        if sync:
            self.synchronize()
        return 0

    def set_surface_velocity(self, coef, shape=0):
        """Fourier-mode surface velocity of a disk: coef = (a, b, c, d) rows of
        M normal-cos, normal-sin, tangential-cos, tangential-sin amplitudes
        (or the flat 4M vector); None removes it."""
        if coef is None:
            self._chk(self.L.cup2d_set_surface_velocity(self.h, shape, 0, None))
            self._surf_modes = 0
            return
        c = np.ascontiguousarray(np.asarray(coef, dtype=np.float64).ravel())
        assert c.size % 4 == 0
        self._chk(self.L.cup2d_set_surface_velocity(self.h, shape, c.size // 4, _dp(c)))
        self._surf_modes = c.size // 4

    def adjoint_forward(self, nsteps, every=20, dts=None):
        """cup2d_adjoint_forward: J = mean drag over nsteps steps (dt from the
        CFL rule, or the given sequence); returns (J, the dts used)."""
        out = np.zeros(nsteps)
        J = C.c_double()
        din = None
        if dts is not None:
            din = np.ascontiguousarray(np.asarray(dts, dtype=np.float64))
            assert din.size == nsteps
        self._chk(self.L.cup2d_adjoint_forward(self.h, nsteps, every,
                                               _dp(din) if din is not None else None,
                                               _dp(out), C.byref(J)))
        self._adj_modes = self._surf_modes
        return J.value, out

    def adjoint_backward(self):
        """cup2d_adjoint_backward: dJ/dcoef of the last adjoint_forward."""
        g = np.zeros(4 * self._adj_modes)
        self._chk(self.L.cup2d_adjoint_backward(self.h, _dp(g)))
        return g

    def refine_residuals(self):
        """FP32: relative residual at the start of each refinement sweep of
        the last Poisson solve."""
        out = np.zeros(4)
        n = C.c_int()
        self._chk(self.L.cup2d_refine_residuals(self.h, _dp(out), C.byref(n)))
        return out[:n.value].copy()

    def simulate_graph(self, nsteps, sync=True):
        """cup2d_simulate_graph: the steps replayed from a captured HIP graph."""
        self._chk(self.L.cup2d_simulate_graph(self.h, nsteps))
        if sync:
            self.synchronize()

    def synchronize(self):
        self._chk(self.L.cup2d_synchronize(self.h))

    def timer_enable(self, kernel="*", reserve=1024):
        self._chk(self.L.cup2d_timer_enable(self.h, kernel.encode(), reserve))

    def timer_sample(self, kernel="*", every=1):
        """Time only every n-th launch of the kernel."""
        self._chk(self.L.cup2d_timer_sample(self.h, kernel.encode(), every))

    def timer_read(self, kernel):
        """(total_ms, launches) since the last read of this timer."""
        t, n = C.c_double(), C.c_int()
        self._chk(self.L.cup2d_timer_read(self.h, kernel.encode(), C.byref(t), C.byref(n)))
        return t.value, n.value

    def force_log_enable(self, capacity):
        """Keep the last `capacity` per-step force records of every shape on
        the device (0 disables)."""
        self._chk(self.L.cup2d_force_log_enable(self.h, capacity))
        self._flog_cap = capacity

    def force_log(self, i=0, max_records=None):
        """The recorded forces of shape i, oldest first: an (n, 24) array whose
        columns are FORCE_LOG_COLUMNS."""
        n = C.c_int()
        cap = getattr(self, "_flog_cap", 0) if max_records is None else max_records
        out = np.empty((cap, len(FORCE_LOG_COLUMNS)))
        self._chk(self.L.cup2d_force_log_read(self.h, i, cap, _dp(out), C.byref(n)))
        return out[:n.value].copy()

    def poisson_dense(self, rhs):
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        out = np.empty_like(rhs)
        self._chk(self.L.cup2d_poisson_dense(self.h, _dp(rhs), _dp(out)))
        return out


def find_max_u_noalloc(u_ptr, v_ptr, ncells, uinf, out_ptr, stream=None):
    """cup2d_find_max_u_noalloc: findMaxU over device planes u_ptr, v_ptr
    (ncells doubles each) into the device double at out_ptr, enqueued on
    `stream` (a raw hipStream_t; None = the null stream).  No allocation."""
    L = lib()
    rc = L.cup2d_find_max_u_noalloc(C.c_void_p(u_ptr), C.c_void_p(v_ptr), int(ncells),
                                    float(uinf[0]), float(uinf[1]), C.c_void_p(out_ptr),
                                    C.c_void_p(stream) if stream else None)
    if rc != 0:
        raise Cup2dError(rc, L.cup2d_last_error(None).decode())


class Ensemble:
    """Independent simulations advanced by one captured graph per two steps
    (cup2d_ensemble_*).  Members are Simulation objects (reset done)."""

    def __init__(self, sims, lanes=0):
        self.L = lib()
        self.sims = list(sims)
        arr = (C.c_void_p * len(self.sims))(*[s.h.value for s in self.sims])
        h = C.c_void_p()
        rc = self.L.cup2d_ensemble_create(arr, len(self.sims), C.byref(h))
        if rc != OK:
            raise Cup2dError(rc, self.L.cup2d_ensemble_last_error(None).decode())
        self.h = h
        if lanes:
            self._chk(self.L.cup2d_ensemble_set_lanes(self.h, lanes))

    def _chk(self, rc):
        if rc != OK:
            raise Cup2dError(rc, self.L.cup2d_ensemble_last_error(self.h).decode())
        return rc

    def simulate(self, nsteps, sync=True):
        self._chk(self.L.cup2d_ensemble_simulate(self.h, nsteps))
        if sync:
            self.synchronize()

    def synchronize(self):
        self._chk(self.L.cup2d_ensemble_synchronize(self.h))

    def tune(self, launches=3):
        """Autotune the lane count on the live run (2 * launches steps per
        candidate, real steps); returns (lanes, ms per two-step launch)."""
        lanes, ms = C.c_int(), C.c_double()
        self._chk(self.L.cup2d_ensemble_tune(self.h, launches, C.byref(lanes), C.byref(ms)))
        return lanes.value, ms.value

    def set_lanes(self, lanes):
        self._chk(self.L.cup2d_ensemble_set_lanes(self.h, lanes))

    def close(self):
        if getattr(self, "h", None):
            self.L.cup2d_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()
