/*
 * cup2d.h -- C-ABI of libcup2d_hip.so, the MI355X (gfx950) implementation of
 * CubismUP_2D's per-timestep hot path on the uniform block grid.
 *
 * Every entry point returns an int status (CUP2D_OK == 0, negative on error;
 * the message is in cup2d_last_error()).  No torch types, plain pointers and
 * sizes.  Each entry point names the reference interface it replaces.
 *
 * Data contract: the host exchanges fields either in the reference's block
 * layout (Cubism GridBlock<Element>::data[iy][ix], blocks in blockID order
 * blockID = bx + bpdx*by, vector components interleaved; Definitions.h:51-110)
 * or as dense row-major (comp, NY, NX) arrays.  On the device every field is
 * a dense row-major plane per component.
 *
 * Operators are enqueued on the context's stream; the device holds dt, time,
 * step, U_inf and every body's state, so a whole step runs without a host
 * round trip.  Entry points that return values to the host synchronize.
 */
#ifndef CUP2D_H
#define CUP2D_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CUP2D_BS 32

enum cup2d_status {
  CUP2D_OK = 0,
  CUP2D_ERR_ARG = -1,       /* invalid argument / size / layout            */
  CUP2D_ERR_HIP = -2,       /* HIP runtime or launch failure              */
  CUP2D_ERR_STATE = -3,     /* call not valid in the current state        */
  CUP2D_ERR_BODY_OUT = -4,  /* a body left the domain (PutObjectsOnGrid.cpp:213-218) */
  CUP2D_ERR_CAPACITY = -5,  /* obstacle-block pool exhausted              */
  CUP2D_ERR_NONFINITE = -6  /* NaN/Inf detected (Helpers.cpp:178-288)    */
};

enum cup2d_poisson {
  CUP2D_POISSON_PERIODIC = 0,   /* FFTW_periodic (FFTW_periodic.h)                   */
  CUP2D_POISSON_FREESPACE = 1,  /* FFTW_freespace (FFTW_freespace.cpp)               */
  CUP2D_POISSON_COSINE = 2      /* FFTW_dirichlet: DCT-II/III, the reference default  */
};
enum cup2d_field {
  CUP2D_VEL = 0, CUP2D_CHI, CUP2D_PRES, CUP2D_TMP, CUP2D_TMPV, CUP2D_UDEF, CUP2D_INVRHO,
  CUP2D_VFLUID,          /* sim.vFluid (SimulationData.h:31), iterative penalization only */
  CUP2D_NFIELDS
};
enum cup2d_layout { CUP2D_LAYOUT_BLOCKS = 0, CUP2D_LAYOUT_DENSE = 1 };

/* Run configuration: the SimulationData parameters parsed in
 * Simulation::parseRuntime (Simulation.cpp:66-104). */
typedef struct {
  int bpdx, bpdy;        /* -bpdx -bpdy                                      */
  double extent;         /* -extent                                          */
  double nu;             /* -nu                                              */
  double CFL;            /* -CFL                                             */
  double lambda;         /* -lambda (implicit penalization only)             */
  double fadeLen;        /* -fadeLen (fraction of extent)                    */
  int poisson;           /* -poissonType periodic|freespace|cosine            */
  int implicit_penal;    /* 0: EXPL_INTEGRATE_MOM (PressureSingle.cpp:20)    */
  int iterative_penal;   /* -iterativePenalization: 1 = PressureIterator_unif
                            replaces PressureSingle (Simulation.cpp:215-216) */
  void* stream;          /* hipStream_t to enqueue on; NULL = own stream     */
  /* Real type of the hot path (the reference's build switch
   * _FLOAT_PRECISION_, Definitions.h:18-22; fftwf plans FFTW_periodic.h:
   * 120-126).  CUP2D_PRECISION_FP32: velocity and Poisson-RHS planes in
   * float, advDiff / RHS / FFT Poisson / projection in float; the pressure,
   * obstacle fields, body integrals and forces stay FP64, and refine_sweeps
   * FP64 defect-correction sweeps (FP64 residual of the spectral operator,
   * FP32 solve of it, FP64 update) follow each FP32 Poisson solve (periodic
   * and cosine solvers; 0 for free-space).  Iterative penalization needs
   * FP64.  Fields still cross the C-ABI as double. */
  int precision;
  int refine_sweeps;
} cup2d_cfg;

enum cup2d_precision { CUP2D_PRECISION_FP64 = 0, CUP2D_PRECISION_FP32 = 1 };

/* Shape description: the keys parsed by Shape::Shape, Disk::Disk and
 * Fish / CarlingFish (Shape.cpp:331-345, ShapesSimple.h:19-21, Fish.h:27-29,
 * CarlingFish.cpp:90-95), defaults already resolved. */
typedef struct {
  int kind;              /* 0 = disk, 1 = carlingfish                        */
  double radius, xpos, ypos, angle_deg, rhoS;
  int bFixedx, bFixedy, bForcedx, bForcedy, bBlockAng;
  double xvel, yvel, angvel, timeForced, tAccel;
  /* carlingfish: -L -T -phi -amplitudeFactor and the midline / half-width
   * laws.  envelope 0 is AmplitudeFish's 4/33 (s + L/32) lateral amplitude
   * with wavelength L (CarlingFish.cpp:19-27); envelope 1 is
   * L (env_a0 + env_a2 (s/L)^2) with wavelength waveLength*L.  profile 0 is
   * AmplitudeFish::_width (CarlingFish.cpp:41-48); profile 1 is the NACA
   * 4-digit half-thickness with closed trailing edge, ratio naca_t. */
  double length, Tperiod, phaseShift, amplitudeFactor;
  int envelope;
  double env_a0, env_a2, waveLength;
  int profile;
  double naca_t;
} cup2d_shape_desc;

/* Body state after a step (Shape.h:22-62). */
typedef struct {
  double center[2], centerOfMass[2], labCenterOfMass[2], d_gm[2], orientation;
  double M, J, u, v, omega;
  double fluidAngMom, fluidMomX, fluidMomY, penalDX, penalDY, penalM, penalJ;
  double perimeter, forcex, forcey, forcex_P, forcey_P, forcex_V, forcey_V;
  double torque, torque_P, torque_V, drag, thrust, circulation;
  double Pout, PoutBnd, defPower, defPowerBnd, Pthrust, Pdrag, EffPDef, EffPDefBnd;
  double n_obstacle_blocks, n_surface_points;
} cup2d_shape_state;

typedef struct cup2d_ctx cup2d_ctx;

/* --- lifetime ------------------------------------------------------------
 * Replaces SimulationData::allocateGrid (SimulationData.cpp:17-47) and the
 * solver constructors FFTW_periodic::FFTW_periodic (FFTW_periodic.h:99-130),
 * FFTW_freespace::FFTW_freespace (FFTW_freespace.cpp:78-138). */
int cup2d_create(const cup2d_cfg* cfg, cup2d_ctx** out);
/* Replaces SimulationData::~SimulationData (SimulationData.cpp:150-170). */
int cup2d_destroy(cup2d_ctx* ctx);
/* Message of the last error on this context ("" if none). */
const char* cup2d_last_error(const cup2d_ctx* ctx);
/* Replaces Simulation::createShapes for "disk" and "carlingfish"
 * (Simulation.cpp:106-173). */
int cup2d_add_shape(cup2d_ctx* ctx, const cup2d_shape_desc* desc, int* shape_id);

/* --- data exchange ---------------------------------------------------------
 * Host <-> device copy of one field (all components).  BLOCKS layout is the
 * reference's grid memory reached through BlockInfo::ptrBlock; DENSE is
 * (comp, NY, NX).  These are the data movements of PoissonSolver::cub2rhs /
 * sol2cub (PoissonSolver.cpp:125-168) when the caller keeps host blocks. */
int cup2d_upload_field(cup2d_ctx* ctx, int field, int layout, const double* host);
int cup2d_download_field(cup2d_ctx* ctx, int field, int layout, double* host);
/* Device pointer of one component plane (dense, row-major, pitch NX). */
int cup2d_field_device_ptr(cup2d_ctx* ctx, int field, int comp, void** dptr);
/* time, dt, step, U_inf (SimulationData.h:41-45) and grid spacing h. */
int cup2d_get_scalars(cup2d_ctx* ctx, double* time, double* dt, int* step, double uinf[2],
                      double* h);
int cup2d_set_scalars(cup2d_ctx* ctx, double time, int step, double uinfx, double uinfy);
int cup2d_get_shape(cup2d_ctx* ctx, int shape_id, cup2d_shape_state* out);
/* Midline of swimmer shape_id after the last PutObjectsOnGrid: the data
 * FishData::writeMidline2File dumps (FishLibrary.cpp:53-64) plus normals and
 * the internal frame.  *Nm receives the number of midline points; out (if not
 * NULL, 10*Nm + 8 doubles) receives the arrays rS rX rY vX vY norX norY vNorX
 * vNorY width (Nm each; centre-of-mass frame rotated by theta_internal) and
 * area J CoM_internal[2] vCoM_internal[2] theta_internal angvel_internal
 * (Fish.h:21-23). */
int cup2d_fish_midline(cup2d_ctx* ctx, int shape_id, int* Nm, double* out);

/* --- operators (Operator::operator()(dt), Operator.h:24-26) --------------- */
/* Simulation::reset: IC + PutObjectsOnGrid(0) + ApplyObjVel (Simulation.cpp:232-241). */
int cup2d_reset(cup2d_ctx* ctx);
/* findMaxU::run (Helpers.cpp:136-176). */
int cup2d_find_max_u(cup2d_ctx* ctx, double* maxU);
/* Simulation::calcMaxTimestep (Simulation.cpp:254-277); dt also stays on device. */
int cup2d_calc_max_timestep(cup2d_ctx* ctx, double* dt);
/* PutObjectsOnGrid::operator() (PutObjectsOnGrid.cpp:186-241). */
int cup2d_put_objects(cup2d_ctx* ctx, double dt);
/* advDiff::operator() (advDiff.cpp:36-118). */
int cup2d_advdiff(cup2d_ctx* ctx, double dt);
/* PressureSingle::updatePressureRHS (PressureSingle.cpp:110-214). */
int cup2d_pressure_rhs(cup2d_ctx* ctx, double dt);
/* PoissonSolver::solve(BSRC=tmp, BDST=pres): cub2rhs fade + FFT solve +
 * sol2cub (PoissonSolver.cpp:61-168, FFTW_periodic.h:135-166,
 * FFTW_freespace.cpp:14-56, CUDA_all.cu:41-128). */
int cup2d_poisson_solve(cup2d_ctx* ctx);
/* PressureSingle::pressureCorrection (PressureSingle.cpp:216-241). */
int cup2d_pressure_correction(cup2d_ctx* ctx, double dt);
/* integrateMomenta + Shape::updateVelocity + preventCollidingObstacles
 * (PressureSingle.cpp:23-67,243-344,362-367; Shape.cpp:25-59). */
int cup2d_body_update(cup2d_ctx* ctx, double dt);
/* PressureSingle::penalize (PressureSingle.cpp:69-108). */
int cup2d_penalize(cup2d_ctx* ctx, double dt);
/* PressureSingle::operator() (PressureSingle.cpp:346-373). */
int cup2d_pressure_single(cup2d_ctx* ctx, double dt);
/* PressureIterator_unif::operator() (PressureIterator_unif.cpp:216-283):
 * projection / body update / penalization fixed-point iteration with an
 * adaptive tolerance.  iters and relDF (may be NULL) receive the iterations
 * done and the last relative momentum change. */
int cup2d_pressure_iterator(cup2d_ctx* ctx, double dt, int* iters, double* relDF);
/* ComputeForces::operator() + Shape::computeForces (ComputeForces.cpp:16-111,
 * Shape.cpp:252-281). */
int cup2d_compute_forces(cup2d_ctx* ctx);
/* Simulation::advance (Simulation.cpp:279-305) with the given dt. */
int cup2d_advance(cup2d_ctx* ctx, double dt);
/* nsteps x (calcMaxTimestep + advance), dt computed on the device, no host
 * synchronization inside (Simulation::simulate, Simulation.cpp:243-252). */
int cup2d_simulate(cup2d_ctx* ctx, int nsteps);
/* Imposed surface velocity of a disk, the deformation velocity the reference
 * keeps in ObstacleBlock::udef (zero for its Disk, ShapeLibrary.cpp:21-34):
 * with phi the angle about the disk's centre of mass, the velocity along
 * the outward normal is sum_m a_m cos(m phi) + b_m sin(m phi) and along the
 * tangent sum_m c_m cos(m phi) + d_m sin(m phi), m = 1..nmodes (<= 32); coef
 * holds a_1..a_M b_1..b_M c_1..c_M d_1..d_M.  It acts wherever the
 * reference reads udef: integrateMomenta (PressureSingle.cpp:23-67), the
 * penalization target (:69-108) and the deformation power of ComputeForces
 * (ComputeForces.cpp:74-99); not the reset's ApplyObjVel.  nmodes = 0
 * removes it. */
int cup2d_set_surface_velocity(cup2d_ctx* ctx, int shape_id, int nmodes, const double* coef);

/* --- adjoint --------------------------------------------------------------
 * Gradient of the time-averaged drag J = (1/nsteps) sum_n drag_n of a forced,
 * fixed disk (ComputeForces.cpp:91-93) with respect to the coefficients of
 * its surface velocity (cup2d_set_surface_velocity), by a hand-written
 * discrete adjoint of the whole step.  No counterpart in the reference.
 * cup2d_adjoint_forward runs nsteps steps from the current state (dt from
 * the CFL rule, or dts_in[k] if given; the dts used go to dts_out when not
 * NULL), checkpoints the velocity every `every` steps and returns J;
 * cup2d_adjoint_backward recomputes each segment from its checkpoint,
 * sweeps the adjoint backwards and writes dJ/dcoef (4 nmodes doubles, in
 * the coefficient layout), then restores the state after the forward run.
 * The configuration: one disk, forced and fixed, explicit penalization,
 * FP64, its blocks outside the boundary fade band. */
int cup2d_adjoint_forward(cup2d_ctx* ctx, int nsteps, int every, const double* dts_in,
                          double* dts_out, double* J);
int cup2d_adjoint_backward(cup2d_ctx* ctx, double* grad);

/* FP32 precision: the relative residual |b - A p| / |b| of the FP64
 * spectral operator at the start of each refinement sweep of the last
 * Poisson solve (the residual of the FP32 solve, then of each corrected
 * solution but the last); *nsweeps receives the sweep count (0 at FP64 or
 * with the free-space solver).  res holds at least 4 doubles. */
int cup2d_refine_residuals(cup2d_ctx* ctx, double* res, int* nsweeps);

/* cup2d_simulate with the steps replayed from a captured HIP graph: two
 * whole steps (PutObjectsOnGrid .. ComputeForces, dt on the device) per
 * graph launch, captured on first use and cached per step parity.  Each
 * graph starts and ends with nothing pending on the context's side streams
 * (the last step's ComputeForces runs inside it), so it is launch-for-launch
 * the same work as cup2d_simulate.  Needs PressureSingle (no iterative
 * penalization: its convergence test waits on the host) and the kernel
 * timers off.  The reference's loop is Simulation::simulate
 * (Simulation.cpp:243-252). */
int cup2d_simulate_graph(cup2d_ctx* ctx, int nsteps);

/* Wait for the context's stream; reports device-side errors (body out of
 * domain) raised by enqueued steps. */
int cup2d_synchronize(cup2d_ctx* ctx);

/* Dense Poisson solve without the boundary fade: rhs (NY*NX, row-major,
 * host) -> sol (host).  Same transform as cup2d_poisson_solve. */
int cup2d_poisson_dense(cup2d_ctx* ctx, const double* rhs, double* sol);

/* Cells per dimension of the grid. */
int cup2d_grid_size(const cup2d_ctx* ctx, int* NX, int* NY);

/* Per-step force log: the rows ComputeForces appends to forceValues_<id>.dat
 * and powerValues_<id>.dat at every step (Shape.cpp:305-328), kept on the
 * device so a device loop (cup2d_simulate) records every step.
 * cup2d_force_log_enable keeps the last `capacity` records per shape
 * (0 disables; call it after the shapes are added).  cup2d_force_log_read
 * copies the min(max_records, held) most recent records of one shape, oldest
 * first, CUP2D_FORCE_LOG_REC doubles each: time step Fx Fy FxPres FyPres FxVisc
 * FyVisc tau tauPres tauVisc drag thrust perimeter circulation Pthrust Pdrag
 * PoutBnd Pout defPowerBnd defPower EffPDefBnd EffPDef n_surface_points. */
#define CUP2D_FORCE_LOG_REC 24
int cup2d_force_log_enable(cup2d_ctx* ctx, int capacity);
int cup2d_force_log_read(cup2d_ctx* ctx, int shape_id, int max_records, double* out,
                         int* nrecords);

/* Per-kernel timers: the role of SimulationData::startProfiler/stopProfiler
 * (SimulationData.cpp:187-205) without the Checker.  kernel is one of
 * find_max_u dt advdiff flux_fix pressure_rhs rhs_obstacles fade fft_rows_fwd
 * fft_cols fft_rows_inv pressure_correction put_objects body_update penalize
 * forces, or "*".  reserve_launches > 0 enables (pre-creating that many HIP
 * event pairs), 0 disables.  Each launch is bracketed by two events on the
 * context's stream.  timer_read returns and clears the summed duration. */
int cup2d_timer_enable(cup2d_ctx* ctx, const char* kernel, int reserve_launches);
int cup2d_timer_read(cup2d_ctx* ctx, const char* kernel, double* total_ms, int* launches);
/* Time only every n-th launch of the kernel ("*": all timers; n >= 1, default
 * 1), so a long device loop can be timed with few event pairs on its stream. */
int cup2d_timer_sample(cup2d_ctx* ctx, const char* kernel, int every);

/* --- ensembles ------------------------------------------------------------
 * Independent simulations advanced together: the reference runs one
 * Simulation per process under an RL driver (main_RL_smartCyl.cpp:102-205);
 * here captured graphs hold two whole steps of every member.  Members are
 * spread over `lanes` lanes (default 1; 0 = one lane per member): each lane
 * is one captured graph, its members one after another, launched on its own
 * stream, the lanes side by side, so two steps of the whole ensemble cost
 * `lanes` graph launches.  Members are existing contexts on one device (any
 * grid sizes, solvers and precisions, reset done, same step parity); each
 * keeps its own API (fields, shapes, forces) between ensemble calls, and
 * cup2d_ensemble_simulate orders itself after and before their streams.
 * cup2d_ensemble_tune times every lane count 1, 2, 4, .. n over `launches`
 * launches each on the live run (those steps are real steps) and keeps the
 * fastest.  Errors are read with cup2d_ensemble_last_error. */
typedef struct cup2d_ensemble cup2d_ensemble;
int cup2d_ensemble_create(cup2d_ctx* const* ctxs, int n, cup2d_ensemble** out);
int cup2d_ensemble_destroy(cup2d_ensemble* ens);
const char* cup2d_ensemble_last_error(const cup2d_ensemble* ens);
int cup2d_ensemble_set_lanes(cup2d_ensemble* ens, int lanes);
int cup2d_ensemble_simulate(cup2d_ensemble* ens, int nsteps);
int cup2d_ensemble_synchronize(cup2d_ensemble* ens);
int cup2d_ensemble_tune(cup2d_ensemble* ens, int launches, int* lanes, double* ms_per_launch);

/* findMaxU over caller-owned device planes, allocating nothing.  No
 * counterpart in the reference: it is for callers that cannot allocate device
 * memory (code under stream capture, tight serving loops).  u and v are device
 * planes of ncells doubles (ncells <= 2^28); *maxU (device memory) receives
 * max over cells of |u+uinfx|, |v+uinfy|, |u|, |v| (NaN kept), the value
 * cup2d_find_max_u computes.  Enqueued on `stream` (a hipStream_t; NULL = the
 * null stream), no synchronization.  The partial maxima live in one
 * library-held device array per GPU (32 KiB), shared by every call on that
 * GPU: calls that overlap on one GPU race; calls ordered on one stream do
 * not.  Context-free: errors are read with cup2d_last_error(NULL). */
int cup2d_find_max_u_noalloc(const double* u, const double* v, long long ncells, double uinfx,
                             double uinfy, double* maxU, void* stream);

#ifdef __cplusplus
}
#endif
#endif
