// internal.hpp -- device data structures of libcup2d_hip (gfx950).
//
// Device layout: every field is a dense row-major plane per component,
// index y*NX + x, with NX = 32*bpdx, NY = 32*bpdy.  The reference's block
// grid (32x32 blocks in blockID order, AoS vector elements) is only a host
// interchange format; ghosts are the neighbouring row/column of the dense
// plane, domain faces are filled at load time with BlockLabOpen's
// zero-gradient rule (Definitions.h:125-157).
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/cup2d.h"

namespace cup2d {

constexpr int BS = CUP2D_BS;
constexpr int MAX_SHAPES = 64;
constexpr int NPART = 24;  // doubles of partial sums per obstacle slot
constexpr int MAX_FISH = 8;
constexpr int FISH_MAX_NM = 1536;  // midline points per swimmer (LDS of the midline kernel)
constexpr int FISH_SEGD = 16;      // doubles per AreaSegment record

// Constant per-shape parameters (Shape.h:22-41, ShapesSimple.h:16-17).
struct ShapeParams {
  int kind;        // 0 = disk, 1 = carlingfish
  int fish;        // index into the swimmer table, -1 for other kinds
  int bFixedx, bFixedy, bForcedx, bForcedy, bBlockAng;
  double radius, rhoS;
  double forcedu, forcedv, forcedomega, timeForced, tAccel;
  double origC[2], origAng;
  double appliedForceX, appliedForceY, appliedTorque;
};

// Step scalars held on the device (SimulationData.h:41-70).
struct StepState {
  double time, dt;
  double uinf[2];
  int step;
  int error;     // sticky: CUP2D_ERR_BODY_OUT / CUP2D_ERR_CAPACITY
  double maxU;   // last findMaxU result
  double fade_corr;  // cub2rhs correction factor of the last solve
  double flux_corr;  // advDiff boundary-flux correction of the last step
  double dt_done;    // dt of the step whose forces are pending (set by the body update)
  double tnext[2];   // time + dt of the step with that parity, set by its list builder
  double pmean;      // FP32 refinement: mean of the pressure before the sweep's update
  double res[4];     // FP32 refinement: |b - A p| / |b| at the start of each sweep
  int rows_redo;     // an obstacle block lies in the fade band: the early row pass is redone
  int rows_early_on; // the early row pass runs this step (few block rows hold obstacles)
};

// Body state captured by the body update for ComputeForces, which may run on
// the side stream while the next step already moves the bodies.
struct ForceSnap {
  double C[2], u, v, omega;
  double time;  // sim.time of the step (the step end may run before ComputeForces)
  int step, pad_;
};

// Per-step force records (the columns of forceValues_<id>.dat and
// powerValues_<id>.dat, Shape.cpp:305-328): time step Fx Fy FxPres FyPres
// FxVisc FyVisc tau tauPres tauVisc drag thrust perimeter circulation
// Pthrust Pdrag PoutBnd Pout defPowerBnd defPower EffPDefBnd EffPDef
// n_surface_points.
constexpr int FLOG_REC = CUP2D_FORCE_LOG_REC;

// Imposed surface velocity of one shape (cup2d_set_surface_velocity): a
// Fourier series in the angle phi about the shape's centre of mass,
//   u_n(phi) = sum_m a_m cos(m phi) + b_m sin(m phi)   (along the outward normal)
//   u_t(phi) = sum_m c_m cos(m phi) + d_m sin(m phi)   (along the tangent),
// m = 1..M, added to the shape's deformation velocity where that enters the
// step (integrateMomenta, penalize, ComputeForces' deformation power).
constexpr int SURF_MAX_MODES = 32;
struct SurfVel {
  int shape;           // -1: none
  int M;
  const double* coef;  // device [4M]: a_1..a_M b_1..b_M c_1..c_M d_1..d_M
};

// Constant run parameters passed by value to kernels.
struct Grid {
  int NX, NY, bpdx, bpdy;
  double h, hb, extent0, extent1;
  double nu, CFL, lambda, fadeLenX, fadeLenY;
  int implicit_penal, nshapes, nfish;
};

// One swimmer (FishData, FishLibrary.h:41-147, with the Fish internal-frame
// state, Fish.h:19-25).  Arrays are device memory.
struct FishDev {
  int shape, Nm, Nseg, envelope;
  double length, Tperiod, phaseShift, env_a0, env_a2, waveLength;
  double* rS;     // [Nm] arclength of the midline points (constant)
  double* width;  // [Nm] half-width (constant)
  double* mid;    // [8][Nm] rX rY vX vY norX norY vNorX vNorY of this step
  double* seg;    // [Nseg][FISH_SEGD] AreaSegment: safe, lab box, object box, |normalI|, |normalJ|
  double* box;    // [4] union of the segments' lab boxes (x0 x1 y0 y1)
  double* st;     // [16] area J CoM_internal[2] vCoM_internal[2] theta_internal angvel_internal,
                  // velocity scale of the fixed-point P2M sums
  double* midA;   // [2][4][Nm] rX rY vX vY of the time-only part of the midline (k_fish_mid_a),
                  // one buffer per step parity
  double* stA;    // [2][8] area J CoM_internal[2] vCoM_internal[2] angvel_internal of midA
  double* wk;     // [8][Nm] working arrays of k_fish_mid_a
};

// Obstacle pool (ObstacleBlock.h:25-137, one slot per (shape, block)).
struct Obstacles {
  int cap;            // slot capacity
  int* slot_shape;    // [cap]
  int* slot_block;    // [cap] blockID
  int* ub_block;      // [cap] distinct block ids (ascending)
  int* ub_start;      // [cap]
  int* ub_count;      // [cap]
  int* ub_slots;      // [cap] slots grouped per distinct block, shape order
  int* cnt;           // [2]: slots and distinct blocks in these lists
  int* slot_of_shape_start;  // [MAX_SHAPES+1] slot range per shape
  double* chi;        // [cap][BS*BS]
  double* dist;
  double* rho;
  double* udef;       // [cap][2][BS*BS]
  double* sdelta;     // [cap][BS*BS] surface delta D (>0 marks a surface point)
  double* sdx;        // [cap][BS*BS] dchidx
  double* sdy;        // [cap][BS*BS] dchidy
  double* part;       // [cap][NPART]
};

struct FFTPlan;  // fft.hip

// Per-kernel HIP-event timers (the role of SimulationData::startProfiler /
// stopProfiler, SimulationData.cpp:187-205): when enabled, every launch of
// the kernel is bracketed by two events on the context's stream.
enum TimerId {
  T_MAXU = 0, T_DT, T_ADVDIFF, T_FLUXFIX, T_RHS, T_RHS_OBST, T_FADE, T_FFT_ROWS_FWD,
  T_FFT_COLS, T_FFT_ROWS_INV, T_PROJECT, T_PUT_OBJECTS, T_BODY, T_PENALIZE, T_FORCES, T_REFINE,
  T_NTIMERS
};
struct Timer {
  bool on = false;
  std::vector<hipEvent_t> ev;  // start/stop pairs
  int used = 0;                // pairs recorded since the last read
  int every = 1;               // time every n-th launch
  long long seen = 0;          // launches since enabled
  bool open = false;           // a start event of this launch was recorded
};

}  // namespace cup2d

struct cup2d_ctx {
  cup2d_cfg cfg;
  cup2d::Grid g;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  // side stream: the PutObjectsOnGrid chain after the list builder runs there,
  // overlapping advDiff and the RHS (neither reads obstacle data)
  hipStream_t aux = nullptr;
  hipEvent_t ev_begin = nullptr, ev_obst = nullptr, ev_pen = nullptr, ev_frc = nullptr;
  hipEvent_t ev_lists = nullptr;  // this step's midline and lists done (main stream)
  // third stream: the swimmers' time-only midline part of the next step runs
  // there beside advDiff (a kernel with many workgroups, so one more resident
  // workgroup costs it nothing); the side stream joins it (ev_mid) while idle
  hipStream_t mids = nullptr;
  hipEvent_t ev_mid = nullptr;
  bool obst_pending = false;
  bool commit_pending = false;  // the placement commit rides in the next obstacle RHS launch
  // ComputeForces and the step end (time += dt) of the last whole step, not
  // enqueued yet.  The next whole step applies the step end in its list
  // builder and runs ComputeForces on the side stream after its obstacle
  // chain (joined before the pressure is overwritten); any other call first
  // enqueues both on the main stream (flush_forces).
  bool forces_pending = false;
  bool end_pending = false;
  bool frc_join = false;  // the main stream must wait for ev_frc
  // the swimmers' time-only midline part for the next whole step is computed
  // (valid for the next advance only; any other call clears it)
  bool midA_ready = false;
  int host_step = 0;                       // step index for the CFL ramp (Simulation.cpp:264)
  cup2d::ForceSnap* fsnap = nullptr;       // device [MAX_SHAPES]
  // device ring of force records: [flog_cap][nshapes][FLOG_REC], filled by
  // every ComputeForces in order; flog_cnt counts the records written
  double* flog = nullptr;
  int flog_cap = 0;
  unsigned long long* flog_cnt = nullptr;
  int device = 0;
  std::string err;

  // Precision (cup2d_cfg.precision): with CUP2D_PRECISION_FP32 the velocity
  // planes (u, v, tu, tv) and the Poisson RHS plane (rhs) hold float, the
  // stencils and the Poisson FFT run in float, and the pressure stays double:
  // refine_sweeps FP64 defect-correction sweeps (residual of the FP64
  // spectral operator, FP32 solve of it, FP64 update) follow the FP32 solve.
  // The planes keep their double-sized allocations; at FP32 they are read
  // through tp<float>.
  bool f32 = false;
  int refine_sweeps = 0;
  // Captured whole steps (cup2d_simulate_graph): two steps per graph, from
  // and back to the quiescent state (nothing pending on the side streams);
  // one cached graph per step parity, valid while the velocity and list
  // buffers it was captured with are the current ones.
  struct StepGraph {
    hipGraphExec_t exec = nullptr;
    const double* u0 = nullptr;
    const int* ob0 = nullptr;
  };
  StepGraph graphs[2];
  struct AdjointState* adj = nullptr;  // cup2d_adjoint_* buffers (api.hip)
  double* surf_coef = nullptr;  // device [4 SURF_MAX_MODES]
  int surf_shape = -1, surf_modes = 0;
  bool graph_capture = false;  // enqueueing into a capture: device-side step index
  hipEvent_t ev_fork = nullptr, ev_qaux = nullptr, ev_qmid = nullptr;
  double* rres = nullptr;  // FP32: the refinement residual plane (float)
  double* respart = nullptr;  // FP32: residual-norm partials of the inverse row pass
  int rows_inv_wgs = 0;       // workgroups of the last inverse row pass

  // fields (dense planes)
  double* u = nullptr;   // velocity x
  double* v = nullptr;   // velocity y
  double* tu = nullptr;  // tmpV x (advDiff output buffer, swapped with u)
  double* tv = nullptr;
  double* pres = nullptr;
  double* rhs = nullptr;   // sim.tmp during PressureSingle (the Poisson RHS)
  double* chi = nullptr;
  double* udx = nullptr;   // sim.uDef
  double* udy = nullptr;
  double* invrho = nullptr;
  double* vfu = nullptr;   // sim.vFluid (iterative penalization)
  double* vfv = nullptr;
  // PressureIterator_unif state (PressureIterator_unif.h:17-18)
  int iter_old_nsteps = 10000;
  double iter_target = 1e-4;
  double* staging = nullptr;  // host-layout staging (2 comps)
  double* red = nullptr;      // reduction partials
  size_t red_n = 0;
  int nwg = 0;                // workgroups of the stencil kernels
  double* fluxpart = nullptr; // [nwg] advDiff boundary-flux partials
  double* maxpart = nullptr;  // [nwg + obstacle blocks] findMaxU partials
  int maxpart_n = 0;
  unsigned char* obmap = nullptr;  // [blocks] 1 = obstacle block this step
  bool flux_deferred = false; // advDiff's flux correction still to be applied

  cup2d::StepState* st = nullptr;          // device
  cup2d::ShapeParams* sp = nullptr;        // device [MAX_SHAPES]
  cup2d_shape_state* ss = nullptr;         // device [MAX_SHAPES]
  double* ramp = nullptr;                  // device [100] CFL ramp factors
  std::vector<cup2d::ShapeParams> hsp;     // host copy of the parameters
  // Obstacle lists and slot data of this step (ob) and of the previous one
  // (ob_prev), swapped at every PutObjectsOnGrid: ComputeForces of a step
  // reads its lists on the side stream while the next step builds new ones.
  cup2d::Obstacles ob{}, ob_prev{};
  int* ubidx = nullptr;                    // [blocks] distinct-block index this step, -1 if none
  cup2d::FishDev* fishd = nullptr;         // device [MAX_FISH]
  std::vector<cup2d::FishDev> hfish;       // host copy (device pointers)
  int band_nblocks = 0;
  int* band_blocks = nullptr;              // blocks touched by the cub2rhs fade
  int* band_index = nullptr;               // [blocks] position in band_blocks, -1 if none
  bool fade_early = false;                 // this step's fade sums already launched
  bool rows_early = false;  // this step's forward row FFT already ran on the side stream
  cup2d::FFTPlan* fft = nullptr;
  std::string fft_err;                     // why no FFT plan exists (size limits)
  bool shapes_frozen = false;
  cup2d::Timer timers[cup2d::T_NTIMERS];
  // events of the next timed launch (KernelTimer), recorded by its dispatch
  hipEvent_t lt_start = nullptr, lt_stop = nullptr;
};

// A velocity / RHS plane as its element type (see cup2d_ctx::f32).
namespace cup2d {
template <typename T>
inline T* tp(double* p) { return reinterpret_cast<T*>(p); }
template <typename T>
inline const T* tp(const double* p) { return reinterpret_cast<const T*>(p); }
}  // namespace cup2d
namespace cup2d {
inline SurfVel surf_of(const cup2d_ctx* c) {
  return SurfVel{c->surf_modes > 0 ? c->surf_shape : -1, c->surf_modes, c->surf_coef};
}
}  // namespace cup2d
// Runs the statement with T = float at FP32 precision, T = double otherwise.
#define CUP2D_REAL_DISPATCH(ctx, ...) \
  do {                                \
    if ((ctx)->f32) {                 \
      using T = float;                \
      __VA_ARGS__;                    \
    } else {                          \
      using T = double;               \
      __VA_ARGS__;                    \
    }                                 \
  } while (0)

#define CUP2D_HIP_CHECK(ctx, call)                                             \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess) {                                                    \
      (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);          \
      return CUP2D_ERR_HIP;                                                    \
    }                                                                          \
  } while (0)

// launch helpers / entry points implemented in the .hip files
namespace cup2d {
int launch_check(cup2d_ctx* c, const char* what);

extern const char* const TIMER_NAMES[T_NTIMERS];
void timer_mark(cup2d_ctx* c, int id, bool start);
void timer_mark_on(cup2d_ctx* c, int id, bool start, hipStream_t s);
struct TimerScope {
  cup2d_ctx* c;
  int id;
  hipStream_t s;
  TimerScope(cup2d_ctx* c_, int id_, hipStream_t s_ = nullptr) : c(c_), id(id_), s(s_) {
    timer_mark_on(c, id, true, s);
  }
  ~TimerScope() { timer_mark_on(c, id, false, s); }
};
// Timer of a single kernel launched with CUP2D_TIMED_LAUNCH inside the scope:
// the dispatch itself records the start and stop events (no marker packets
// around it on the stream).
void kernel_timer_begin(cup2d_ctx* c, int id);
void kernel_timer_end(cup2d_ctx* c, int id);
struct KernelTimer {
  cup2d_ctx* c;
  int id;
  KernelTimer(cup2d_ctx* c_, int id_) : c(c_), id(id_) { kernel_timer_begin(c, id); }
  ~KernelTimer() { kernel_timer_end(c, id); }
};
#define CUP2D_TIMED_LAUNCH(ctx, kernel, grid, block, shm, stream, ...)                     \
  do {                                                                                    \
    hipExtLaunchKernelGGL(kernel, grid, block, shm, stream, (ctx)->lt_start, (ctx)->lt_stop, \
                          0, __VA_ARGS__);                                                \
    (ctx)->lt_start = (ctx)->lt_stop = nullptr;                                           \
  } while (0)

// stencils.hip
int grid_workgroups(int NX, int NY);
int enqueue_find_max_u(cup2d_ctx* c);
int enqueue_find_max_u_partials(cup2d_ctx* c);
int enqueue_dt(cup2d_ctx* c);
int enqueue_advdiff(cup2d_ctx* c, int defer_fix);
int enqueue_advdiff_on(cup2d_ctx* c, int defer_fix, hipStream_t s);
int enqueue_pressure_rhs(cup2d_ctx* c);
int stencil_edge_count(const cup2d_ctx* c);
int enqueue_fade_sums(cup2d_ctx* c, int compute = 0);
int enqueue_fade_sums_early(cup2d_ctx* c, hipStream_t s);
int enqueue_fade_reduce_on(cup2d_ctx* c, hipStream_t s);  // the fade factor from the sums
int fade_partials(const cup2d_ctx* c);
int enqueue_pressure_correction(cup2d_ctx* c, int want_max);
int enqueue_set_dt(cup2d_ctx* c, double dt);
int enqueue_ic(cup2d_ctx* c);
int enqueue_apply_obj_vel(cup2d_ctx* c);
int enqueue_step_end(cup2d_ctx* c);
// f32: the planes hold float (the velocity and RHS planes at FP32 precision)
int enqueue_blocks_to_dense(cup2d_ctx* c, const double* stage, double* a, double* b, int nc,
                            bool f32 = false);
int enqueue_dense_to_blocks(cup2d_ctx* c, const double* a, const double* b, double* stage, int nc,
                            bool f32 = false);
int enqueue_stage_to_f32(cup2d_ctx* c, const double* stage, double* plane);
int enqueue_f32_to_stage(cup2d_ctx* c, const double* plane, double* stage);

// fft.hip
int fft_create(cup2d_ctx* c);
void fft_destroy(cup2d_ctx* c);
// src: 0 = rhs as given, 1 = grid RHS with the cub2rhs fade, 2 = RHS formed
// from the advDiff output (deferred flux correction, obstacle blocks from the
// RHS plane) then faded.  With src = 2 after enqueue_rows_early, the row
// pass only adds the transform of the obstacle blocks' correction.
int enqueue_poisson_fft(cup2d_ctx* c, double* rhs, double* sol, int src);
// The forward row FFT of the RHS formed from the advDiff output with every
// block taken as free of obstacles, on stream s (the side stream, beside the
// placement chain), after the early fade sums; the pressure solve then adds
// the row transform of the obstacle RHS correction (the row transform is
// linear), or redoes the pass when an obstacle block lies in the fade band.
int enqueue_rows_early(cup2d_ctx* c, hipStream_t s);

// obstacles.hip
int obstacles_alloc(cup2d_ctx* c);
void obstacles_free(cup2d_ctx* c);
constexpr int OBST_LIST_CAP = 2048;  // slots handled by the list builder
int enqueue_put_objects(cup2d_ctx* c, int compute_dt, int overlap = 0);
int join_obstacles(cup2d_ctx* c);
int flush_commit(cup2d_ctx* c);
// compute: form the RHS of the obstacle blocks from the advDiff output first
// (the whole-step path, where the row FFT forms it for every other cell).
int enqueue_rhs_obstacles(cup2d_ctx* c, int compute = 0);
int enqueue_body_update(cup2d_ctx* c);
int enqueue_penalize(cup2d_ctx* c, double* maxpart);
int enqueue_body_penalize(cup2d_ctx* c, double* maxpart);
int enqueue_forces(cup2d_ctx* c, int end_step, int on_aux = 0);
// enqueue pending ComputeForces and step end on the main stream
int flush_forces(cup2d_ctx* c);
// main stream waits for a ComputeForces running on the side stream
int join_forces(cup2d_ctx* c);
int enqueue_shapes_reset(cup2d_ctx* c);
int enqueue_obst_reset(cup2d_ctx* c);

// adjoint.hip: one step of the adjoint of the drag objective (see there)
struct AdjStep {
  int shape;
  double dt, uinf0, uinf1, vu0, vu1, wJ;
  const double *u0, *v0;   // the step's input velocity
  const double *wu, *wv;   // its flux-corrected advDiff output
  const double *u1, *v1;   // its output velocity
  const double* p1;        // its pressure
  double *ub, *vb;         // in: adjoint of the output velocity; out: of the input
  double *pb, *rb, *wbu, *wbv;  // work planes
  double* part;            // work partials (5 per block)
  double* theta_part;      // [obstacle slots][4M]
  double* theta;           // [4M] gradient, accumulated
  double* fluxsum;         // one double
};
int enqueue_adjoint_step(cup2d_ctx* c, const AdjStep& a);
int enqueue_adjoint_fill(cup2d_ctx* c, double* a, size_t n, double val);
int enqueue_adjoint_acc_drag(cup2d_ctx* c, int shape, double w, double* J);

// fish.hip
// FishData::FishData + width table (FishLibrary.cpp:14-38); returns Nm or -1
int fish_discretize(const cup2d_shape_desc* d, double h, std::vector<double>& rS,
                    std::vector<double>& width);
int fish_add(cup2d_ctx* c, const cup2d_shape_desc* d, int shape);
void fish_free_all(cup2d_ctx* c);
int enqueue_fish_midline(cup2d_ctx* c, hipStream_t s);
// the time-only part of the next whole step's midline (time + dt)
int enqueue_fish_mid_next(cup2d_ctx* c, int parity, hipEvent_t after, bool join_aux);
int enqueue_fish_place(cup2d_ctx* c, hipStream_t s);
// PressureIterator_unif pieces (PressureIterator_unif.cpp)
int enqueue_copy_vel_to_tmpv(cup2d_ctx* c);
int enqueue_iter_project(cup2d_ctx* c);
int enqueue_iter_body(cup2d_ctx* c);
int enqueue_iter_penalize(cup2d_ctx* c, double* reldf);
int enqueue_iter_rhs(cup2d_ctx* c);
}  // namespace cup2d
