/*
 * cup2d_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * FP64 C/OpenMP restatement of CubismUP_2D's per-timestep hot path on the
 * uniform block grid (constant density, PressureSingle pipeline).  It is the
 * checker for the HIP library and the timed CPU baseline; it is never linked
 * into, called by, or shipped with the product library.
 *
 * Fields are stored exactly like the reference: blocks of 32x32 elements in
 * blockID order (blockID = bx + bpdx*by), elements row-major data[iy][ix],
 * vector components interleaved (Definitions.h:51-56,84).
 */
#ifndef CUP2D_ORACLE_H
#define CUP2D_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_BS 32

typedef struct {
  int bpdx, bpdy;        /* blocks per dimension (Simulation.cpp:73-74)      */
  double extent;         /* -extent (Simulation.cpp:76)                      */
  double nu;             /* -nu                                              */
  double CFL;            /* -CFL                                             */
  double lambda;         /* -lambda (read only by the implicit penalization) */
  double fadeLen;        /* -fadeLen, fraction of extent (Simulation.cpp:97) */
  int poisson;           /* 0 = periodic (FFTW_periodic), 1 = freespace,     */
                         /* 2 = cosine (FFTW_dirichlet, DCT-II / DCT-III)    */
  int fs_pad2;           /* freespace: 0 = (2N-1) cyclic like the reference, */
                         /*            1 = 2N cyclic (same math)             */
  int implicit_penal;    /* 0 = EXPL_INTEGRATE_MOM (reference), 1 = implicit */
  int checker;           /* 1 = run the NaN/Inf Checker 17x per step        */
  int iterative_penal;   /* 1 = PressureIterator_unif instead of PressureSingle
                            (-iterativePenalization 1, Simulation.cpp:215-216) */
} orc_cfg;

typedef struct {
  int kind;              /* 0 = disk, 1 = carlingfish (CarlingFish.cpp)      */
  double radius, xpos, ypos, angle_deg, rhoS;
  int bFixedx, bFixedy, bForcedx, bForcedy, bBlockAng;
  double xvel, yvel, angvel, timeForced, tAccel;
  /* carlingfish: -L -T -phi -amplitudeFactor (Fish.h:27-29,
   * CarlingFish.cpp:90-95) and the midline / half-width laws */
  double length, Tperiod, phaseShift, amplitudeFactor;
  int envelope;          /* 0 = 4/33 (s + L/32) (CarlingFish.cpp:19-27),
                            1 = L (env_a0 + env_a2 (s/L)^2), wavelength
                            waveLength*L                                     */
  double env_a0, env_a2, waveLength;
  int profile;           /* 0 = AmplitudeFish::_width (CarlingFish.cpp:41-48),
                            1 = NACA 4-digit half-thickness, ratio naca_t    */
  double naca_t;
} orc_shape_desc;

/* Per-shape scalar state; same field order as cup2d_shape_state. */
typedef struct {
  double center[2], centerOfMass[2], labCenterOfMass[2], d_gm[2], orientation;
  double M, J, u, v, omega;
  double fluidAngMom, fluidMomX, fluidMomY, penalDX, penalDY, penalM, penalJ;
  double perimeter, forcex, forcey, forcex_P, forcey_P, forcex_V, forcey_V;
  double torque, torque_P, torque_V, drag, thrust, circulation;
  double Pout, PoutBnd, defPower, defPowerBnd, Pthrust, Pdrag, EffPDef, EffPDefBnd;
  double n_obstacle_blocks, n_surface_points;
} orc_shape_state;

enum { ORC_VEL = 0, ORC_CHI, ORC_PRES, ORC_TMP, ORC_TMPV, ORC_UDEF, ORC_INVRHO,
       ORC_VFLUID, ORC_NFIELDS };

typedef struct orc_sim orc_sim;

orc_sim* orc_create(const orc_cfg* cfg);
void orc_destroy(orc_sim* s);
int  orc_add_shape(orc_sim* s, const orc_shape_desc* d);
int  orc_ncomp(int field);
/* block-layout host arrays: nblocks*32*32*ncomp doubles */
int  orc_set_field(orc_sim* s, int field, const double* src);
int  orc_get_field(const orc_sim* s, int field, double* dst);
void orc_get_scalars(const orc_sim* s, double* time, double* dt, int* step,
                     double uinf[2], double* h);
void orc_set_scalars(orc_sim* s, double time, int step, double uinfx, double uinfy);
int  orc_get_shape(const orc_sim* s, int i, orc_shape_state* out);

/* Operators: same names and order as the reference pipeline. */
void   orc_reset(orc_sim* s);                 /* Simulation::reset           */
double orc_find_max_u(orc_sim* s);            /* findMaxU::run               */
double orc_calc_max_timestep(orc_sim* s);     /* Simulation::calcMaxTimestep */
int    orc_put_objects(orc_sim* s, double dt);/* PutObjectsOnGrid            */
void   orc_advdiff(orc_sim* s, double dt);    /* advDiff                     */
void   orc_pressure_rhs(orc_sim* s, double dt);
void   orc_cub2rhs(orc_sim* s);               /* fade + copy to FFT buffer   */
void   orc_poisson_solve(orc_sim* s);         /* PoissonSolver::solve        */
void   orc_pressure_correction(orc_sim* s, double dt);
void   orc_body_update(orc_sim* s, double dt);/* integrateMomenta+update+coll*/
void   orc_penalize(orc_sim* s, double dt);
void   orc_pressure_single(orc_sim* s, double dt);
/* PressureIterator_unif::operator() (PressureIterator_unif.cpp:216-283);
 * returns the iterations done.  orc_iter_stats: iterations and the relative
 * momentum change of the last call, and the adaptive tolerance after it. */
int    orc_pressure_iterator(orc_sim* s, double dt);
void   orc_iter_stats(const orc_sim* s, int* iters, double* relDF, double* target);
/* Test hook: enabled = 0 skips preventCollidingObstacles in the body update,
 * so a test can see the velocities the impulse starts from. */
void   orc_set_collisions(orc_sim* s, int enabled);
int    orc_set_surface_velocity(orc_sim* s, int shape, int M, const double* coef);
void   orc_compute_forces(orc_sim* s);
int    orc_advance(orc_sim* s, double dt);    /* Simulation::advance         */
int    orc_simulate(orc_sim* s, int nsteps);  /* nsteps x (dt + advance)     */
void   orc_apply_obj_vel(orc_sim* s);

/* Poisson solve on a dense NYxNX rhs (row-major), writes dense solution.
 * Same arithmetic as orc_poisson_solve without the boundary fade. */
int orc_poisson_dense(orc_sim* s, const double* rhs, double* sol);

/* 1D FFT used by the solver (for tests): in-place complex, sign -1 forward,
 * +1 backward, unnormalized; interleaved re/im. */
int orc_fft1d(double* data, int n, int sign);

/* Swimmer state of shape k (carlingfish): Nm midline points; out (if not
 * NULL) receives 10 arrays of Nm doubles in the order rS rX rY vX vY norX
 * norY vNorX vNorY width (after the last PutObjectsOnGrid: CoM frame,
 * rotated by theta_internal), then area_internal J_internal CoM_internal[2]
 * vCoM_internal[2] theta_internal angvel_internal.  Returns Nm, -1 if shape k
 * is not a swimmer. */
int  orc_fish_info(const orc_sim* s, int k, double* out);
/* sin and cos as used on the swimmer path (for tests) */
void orc_det_sincos(double x, double* sn, double* cs);
/* 1: the reference's own arithmetic on the swimmer path (libm sine/cosine,
 * index-order sums, sequential hat P2M); 0 (default): the device library's. */
void orc_set_reference_arithmetic(int on);
/* sum of n values in the fixed order used on the swimmer path (for tests) */
double orc_lane_sum(const double* v, int n);

/* timing of the last simulate() call, seconds: total and FFT-only */
void orc_get_timing(const orc_sim* s, double* total, double* fft);
int  orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
