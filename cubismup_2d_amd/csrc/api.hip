// api.hip -- the extern "C" boundary of libcup2d_hip.so (declared in
// include/cup2d.h).  Host-side bookkeeping only: allocation, validation,
// layout conversion through a device staging buffer, and the operator
// sequence of Simulation::advance (Simulation.cpp:279-305).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "device_util.hpp"

using namespace cup2d;

namespace cup2d {
const char* const TIMER_NAMES[T_NTIMERS] = {
    "find_max_u", "dt", "advdiff", "flux_fix", "pressure_rhs", "rhs_obstacles", "fade",
    "fft_rows_fwd", "fft_cols", "fft_rows_inv", "pressure_correction", "put_objects",
    "body_update", "penalize", "forces", "refine"};

void timer_mark(cup2d_ctx* c, int id, bool start) { timer_mark_on(c, id, start, nullptr); }

void kernel_timer_begin(cup2d_ctx* c, int id) {
  c->lt_start = c->lt_stop = nullptr;
  Timer& t = c->timers[id];
  if (!t.on || t.seen++ % t.every != 0) return;
  const size_t need = 2 * (size_t)(t.used + 1);
  while (t.ev.size() < need) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) { t.on = false; return; }
    t.ev.push_back(e);
  }
  c->lt_start = t.ev[2 * (size_t)t.used];
  c->lt_stop = t.ev[2 * (size_t)t.used + 1];
  ++t.used;
}

void kernel_timer_end(cup2d_ctx* c, int id) {
  // a scope left before its launch (an error return) recorded nothing
  if (c->lt_start) --c->timers[id].used;
  c->lt_start = c->lt_stop = nullptr;
}

void timer_mark_on(cup2d_ctx* c, int id, bool start, hipStream_t s) {
  Timer& t = c->timers[id];
  if (!t.on) return;
  if (start) {
    t.open = t.seen++ % t.every == 0;
    if (!t.open) return;
  } else {
    if (!t.open) return;
    t.open = false;
  }
  const size_t need = 2 * (size_t)(t.used + 1);
  while (t.ev.size() < need) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) { t.on = false; return; }
    t.ev.push_back(e);
  }
  (void)hipEventRecord(t.ev[2 * (size_t)t.used + (start ? 0 : 1)], s ? s : c->stream);
  if (!start) ++t.used;
}
}  // namespace cup2d

namespace {
int timer_index(const char* name) {
  if (!name) return -1;
  for (int i = 0; i < T_NTIMERS; ++i)
    if (std::strcmp(name, TIMER_NAMES[i]) == 0) return i;
  return -1;
}
thread_local std::string g_create_error;
}  // namespace

namespace cup2d {
// message of a failed call that has no context (cup2d_last_error(NULL))
void set_ctxfree_error(const char* msg) { g_create_error = msg; }
}  // namespace cup2d

namespace {
int fail(cup2d_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

// every entry point first enqueues a ComputeForces still pending from the
// last whole step, so nothing else it enqueues can overtake it
#define CHECK_CTX(c)                                        \
  do {                                                      \
    if (!(c)) return CUP2D_ERR_ARG;                         \
    (void)hipSetDevice((c)->device);                        \
    const int frc_ = flush_forces(c);                       \
    if (frc_) return frc_;                                  \
  } while (0)

#define RC(x)                 \
  do {                        \
    const int rc_ = (x);      \
    if (rc_) return rc_;      \
  } while (0)

int nfield_comps(int f) {
  switch (f) {
    case CUP2D_VEL: case CUP2D_TMPV: case CUP2D_UDEF: case CUP2D_VFLUID: return 2;
    case CUP2D_CHI: case CUP2D_PRES: case CUP2D_TMP: case CUP2D_INVRHO: return 1;
    default: return 0;
  }
}

// the field's planes hold float (FP32 precision: velocity, tmpV, RHS)
bool f32_field(const cup2d_ctx* c, int f) {
  return c->f32 && (f == CUP2D_VEL || f == CUP2D_TMPV || f == CUP2D_TMP);
}

void field_planes(cup2d_ctx* c, int f, double** a, double** b) {
  *b = nullptr;
  switch (f) {
    case CUP2D_VEL: *a = c->u; *b = c->v; break;
    case CUP2D_TMPV: *a = c->tu; *b = c->tv; break;
    case CUP2D_UDEF: *a = c->udx; *b = c->udy; break;
    case CUP2D_CHI: *a = c->chi; break;
    case CUP2D_PRES: *a = c->pres; break;
    case CUP2D_TMP: *a = c->rhs; break;
    case CUP2D_INVRHO: *a = c->invrho; break;
    case CUP2D_VFLUID: *a = c->vfu; *b = c->vfv; break;
    default: *a = nullptr;
  }
}

// Blocks touched by the cub2rhs fade (PoissonSolver.cpp:71-79).
std::vector<int> band_blocks(const Grid& g) {
  std::vector<int> out;
  for (int by = 0; by < g.bpdy; ++by)
    for (int bx = 0; bx < g.bpdx; ++bx)
      if (band_block(g, bx, by)) out.push_back(bx + g.bpdx * by);
  return out;
}

int ensure_frozen(cup2d_ctx* c) {
  if (c->shapes_frozen) return CUP2D_OK;
  RC(obstacles_alloc(c));
  c->shapes_frozen = true;
  return CUP2D_OK;
}

int sync_all(cup2d_ctx* c) {
  RC(join_obstacles(c));
  RC(flush_commit(c));
  RC(flush_forces(c));
  RC(join_obstacles(c));
  RC(join_forces(c));
  CUP2D_HIP_CHECK(c, hipStreamSynchronize(c->mids));
  CUP2D_HIP_CHECK(c, hipStreamSynchronize(c->aux));
  CUP2D_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  return CUP2D_OK;
}

int sync_and_check(cup2d_ctx* c) {
  RC(sync_all(c));
  int err = 0;
  CUP2D_HIP_CHECK(c, hipMemcpy(&err, &c->st->error, sizeof(int), hipMemcpyDeviceToHost));
  if (err == CUP2D_ERR_BODY_OUT) return fail(c, err, "a body left the domain");
  if (err == CUP2D_ERR_CAPACITY) return fail(c, err, "obstacle-block pool exhausted");
  if (err == CUP2D_ERR_NONFINITE)
    return fail(c, err, "non-finite velocity (NaN or Inf reached findMaxU)");
  if (err) return fail(c, err, "device error");
  return CUP2D_OK;
}

int need_fft(cup2d_ctx* c) {
  if (c->fft) return CUP2D_OK;
  return fail(c, CUP2D_ERR_ARG, c->fft_err);
}

// PressureSingle::operator() (PressureSingle.cpp:346-373).  fused (whole-step
// path, advDiff's flux correction deferred): no separate RHS pass; the RHS of
// the obstacle blocks is formed by the obstacle RHS kernel, every other
// cell's by the forward row FFT, so the RHS plane holds only the obstacle
// and boundary-band blocks afterwards.
int pressure_single(cup2d_ctx* c, int fused) {
  RC(need_fft(c));
  if (fused && c->flux_deferred && c->cfg.poisson != CUP2D_POISSON_COSINE) {
    RC(join_obstacles(c));
    RC(enqueue_rhs_obstacles(c, 1));
    // after an early row pass the fade factor is already reduced (the patch
    // pass reduces the final one itself when it has to redo the pass)
    if (c->rows_early)
      c->fade_early = false;
    else
      RC(enqueue_fade_sums(c, 1));
    RC(enqueue_poisson_fft(c, c->rhs, c->pres, 2));
  } else {
    // the side stream may still be writing the advDiff output that the RHS
    // reads (swimmer layout: advDiff runs beside the placement chain)
    RC(join_obstacles(c));
    RC(enqueue_pressure_rhs(c));
    RC(enqueue_rhs_obstacles(c));
    RC(enqueue_fade_sums(c));
    RC(enqueue_poisson_fft(c, c->rhs, c->pres, 1));
  }
  RC(enqueue_pressure_correction(c, 1));
  RC(enqueue_body_penalize(c, c->maxpart + c->nwg));
  return CUP2D_OK;
}

// PressureIterator_unif::operator() (PressureIterator_unif.cpp:216-283).  The
// convergence test needs relDF on the host: one synchronization per iteration.
int pressure_iterator(cup2d_ctx* c, int* iters_out, double* reldf_out) {
  RC(need_fft(c));
  RC(join_obstacles(c));
  double* d_reldf = c->red + 4096 - 8;
  RC(enqueue_copy_vel_to_tmpv(c));
  int iter;
  double relDF = 1e3;
  for (iter = 0; iter < 1000; ++iter) {
    RC(enqueue_iter_project(c));
    RC(enqueue_iter_body(c));
    RC(enqueue_iter_penalize(c, d_reldf));
    RC(enqueue_iter_rhs(c));
    RC(enqueue_fade_sums(c));
    RC(enqueue_poisson_fft(c, c->rhs, c->pres, 1));
    CUP2D_HIP_CHECK(c, hipMemcpyAsync(&relDF, d_reldf, sizeof(double), hipMemcpyDeviceToHost,
                                      c->stream));
    CUP2D_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    const bool conv = relDF < c->iter_target || iter > 100 || iter > 2 * c->iter_old_nsteps;
    if (conv) {
      RC(enqueue_pressure_correction(c, 0));  // finalizePressure (:186-212)
      break;
    }
  }
  c->iter_old_nsteps = iter + 1;
  if (c->iter_old_nsteps > 30) c->iter_target = std::max(relDF, c->iter_target);
  if (c->iter_old_nsteps > 10) c->iter_target *= 1.01;
  if (c->iter_old_nsteps <= 2) c->iter_target *= 0.99;
  c->iter_target = std::min(1e-3, std::max(1e-5, c->iter_target));
  if (iters_out) *iters_out = iter + 1;
  if (reldf_out) *reldf_out = relDF;
  return CUP2D_OK;
}

// Flags of the events that order the main and side streams.  They only order
// device work between two streams of one device and the host never waits on
// them, so recording them needs no system-scope fence (which would write the
// L2 back on every record).
#define CUP2D_SYNC_EVENT_FLAGS (hipEventDisableTiming | hipEventDisableSystemFence)


// Simulation::advance.  compute_dt: dt from the findMaxU partials of the
// previous step (device loop) instead of the value already on the device.
int advance_enqueued(cup2d_ctx* c, int compute_dt) {
  // (with overlap, also the swimmers' time-only midline part of the next
  // step, off the critical path).  With swimmers the placement chain is
  // longer than advDiff: it keeps the main stream and advDiff goes beside it.
  const bool chain_main = !c->hfish.empty() && !c->cfg.iterative_penal;
  RC(enqueue_put_objects(c, compute_dt, chain_main ? 2 : 1));
  if (c->cfg.iterative_penal) {
    RC(enqueue_advdiff(c, 0));
    RC(pressure_iterator(c, nullptr, nullptr));
    RC(enqueue_find_max_u_partials(c));  // for the next step's dt
  } else {
    if (!chain_main) RC(enqueue_advdiff(c, 1));
    RC(pressure_single(c, 1));
  }
  RC(enqueue_forces(c, 1, 1));
  c->host_step += 1;
  return CUP2D_OK;
}

// The quiescent state: everything the last step left pending (the
// placement commit, its ComputeForces and step end, the side-stream chain)
// enqueued, and both side streams joined into the main stream, with no host
// wait, so it can end a captured graph.  The next step's precomputed
// midline part is dropped (the step recomputes it).
int enqueue_quiesce(cup2d_ctx* c) {
  RC(join_obstacles(c));
  RC(flush_commit(c));
  RC(flush_forces(c));
  RC(join_obstacles(c));
  RC(join_forces(c));
  CUP2D_HIP_CHECK(c, hipEventRecord(c->ev_qaux, c->aux));
  CUP2D_HIP_CHECK(c, hipStreamWaitEvent(c->stream, c->ev_qaux, 0));
  CUP2D_HIP_CHECK(c, hipEventRecord(c->ev_qmid, c->mids));
  CUP2D_HIP_CHECK(c, hipStreamWaitEvent(c->stream, c->ev_qmid, 0));
  c->midA_ready = false;
  return CUP2D_OK;
}

// Two whole steps (Simulation::advance with dt from the device) from the
// quiescent state back to it, enqueued into the capture that c->stream is
// part of.  The side streams are forked from the main stream first, so every
// stream the steps use belongs to the capture.  Host-side state is as before
// (two buffer swaps each); host_step is restored, the replays advance it.
int capture_two_steps(cup2d_ctx* c) {
  CUP2D_HIP_CHECK(c, hipEventRecord(c->ev_fork, c->stream));
  CUP2D_HIP_CHECK(c, hipStreamWaitEvent(c->aux, c->ev_fork, 0));
  CUP2D_HIP_CHECK(c, hipStreamWaitEvent(c->mids, c->ev_fork, 0));
  c->graph_capture = true;
  int rc = advance_enqueued(c, 1);
  if (!rc) rc = advance_enqueued(c, 1);
  if (!rc) rc = enqueue_quiesce(c);
  c->graph_capture = false;
  c->host_step -= 2;
  return rc;
}

bool timers_on(const cup2d_ctx* c) {
  for (const Timer& t : c->timers)
    if (t.on) return true;
  return false;
}

int graph_eligible(cup2d_ctx* c) {
  if (!c->shapes_frozen) return fail(c, CUP2D_ERR_STATE, "call reset first");
  if (c->cfg.iterative_penal)
    return fail(c, CUP2D_ERR_STATE, "captured steps need PressureSingle (the iterator waits "
                                    "on the host)");
  if (timers_on(c)) return fail(c, CUP2D_ERR_STATE, "captured steps need the kernel timers off");
  RC(need_fft(c));
  return CUP2D_OK;
}

void drop_graphs(cup2d_ctx* c) {
  for (auto& gr : c->graphs) {
    if (gr.exec) (void)hipGraphExecDestroy(gr.exec);
    gr = cup2d_ctx::StepGraph{};
  }
}

// The cached two-step graph of the current parity and buffers (captured now
// if there is none).  Needs the quiescent state.
int step_graph(cup2d_ctx* c, hipGraphExec_t* out) {
  cup2d_ctx::StepGraph& gr = c->graphs[c->host_step & 1];
  if (gr.exec && gr.u0 == c->u && gr.ob0 == c->ob.cnt) {
    *out = gr.exec;
    return CUP2D_OK;
  }
  if (gr.exec) (void)hipGraphExecDestroy(gr.exec);
  gr = cup2d_ctx::StepGraph{};
  CUP2D_HIP_CHECK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  const int rc = capture_two_steps(c);
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(c->stream, &graph);
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  CUP2D_HIP_CHECK(c, e);
  const hipError_t ei = hipGraphInstantiate(&gr.exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  CUP2D_HIP_CHECK(c, ei);
  gr.u0 = c->u;
  gr.ob0 = c->ob.cnt;
  *out = gr.exec;
  return CUP2D_OK;
}
}  // namespace

namespace cup2d {
// graphs bake buffer pointers and launch arguments: dropped whenever those
// may change (reset, force log, shapes)
void invalidate_graphs(cup2d_ctx* c) { drop_graphs(c); }
}  // namespace cup2d

// ------------------------------------------------------------ the adjoint
// Buffers of cup2d_adjoint_forward / _backward (see adjoint.hip for the
// operator sequence).  Checkpoints hold the velocity and the step and shape
// state every `every` steps; a segment's recompute stores, per step, the
// input velocity, the flux-corrected advDiff output and the pressure.
struct AdjointState {
  int nsteps = 0, every = 0, nseg = 0, cap_steps = 0;
  std::vector<double> dts;
  double uinf0 = 0, uinf1 = 0, vu0 = 1, vu1 = 0;
  size_t n = 0;                        // cells
  std::vector<double*> ck;             // [nseg + 1] (u, v) pairs; the last: the final state
  std::vector<StepState> ck_st;
  std::vector<std::vector<cup2d_shape_state>> ck_ss;
  double* seg = nullptr;               // (every + 1) (u, v), every (wu, wv), every p
  double* work = nullptr;              // ub vb pb rb wbu wbv ub2 vb2
  double* part = nullptr;              // 5 per block
  double* theta_part = nullptr;
  double* theta = nullptr;             // [4M]
  double* scal = nullptr;              // J, flux sum
  bool ready = false;                  // a forward run whose backward may follow
};

namespace {
void adjoint_free(cup2d_ctx* c) {
  AdjointState* a = c->adj;
  if (!a) return;
  for (double* p : a->ck) (void)hipFree(p);
  for (double* p : {a->seg, a->work, a->part, a->theta_part, a->theta, a->scal})
    if (p) (void)hipFree(p);
  delete a;
  c->adj = nullptr;
}

// the forced, fixed single disk with an imposed surface velocity, explicit
// penalization, FP64: the configuration the adjoint is written for
int adjoint_eligible(cup2d_ctx* c) {
  if (!c->shapes_frozen) return fail(c, CUP2D_ERR_STATE, "call reset first");
  if (c->f32) return fail(c, CUP2D_ERR_STATE, "the adjoint runs at FP64 precision");
  if (c->cfg.iterative_penal || c->cfg.implicit_penal)
    return fail(c, CUP2D_ERR_STATE, "the adjoint needs PressureSingle with explicit penalization");
  if (c->g.nshapes != 1 || c->hsp[0].kind != 0)
    return fail(c, CUP2D_ERR_STATE, "the adjoint is written for one disk");
  const ShapeParams& p = c->hsp[0];
  if (!(p.bForcedx && p.bForcedy && p.bBlockAng && p.bFixedx && p.bFixedy))
    return fail(c, CUP2D_ERR_STATE, "the adjoint's disk must be forced and fixed (its "
                                    "velocity and position are then independent of the flow)");
  if (c->surf_shape != 0 || c->surf_modes <= 0)
    return fail(c, CUP2D_ERR_STATE, "set the disk's surface velocity first");
  return need_fft(c);
}

// one forward step at fixed dt, operator by operator on the main stream
// (put objects, advDiff with its flux correction, PressureSingle, forces and
// the step end); rec != nullptr stores the flux-corrected advDiff output and
// the pressure there.
int adjoint_fwd_step(cup2d_ctx* c, double dt, double* rec_w, double* rec_p) {
  const size_t n = (size_t)c->g.NX * c->g.NY;
  RC(enqueue_set_dt(c, dt));
  RC(enqueue_put_objects(c, 0, 0));
  RC(enqueue_advdiff(c, 0));
  if (rec_w) {
    CUP2D_HIP_CHECK(c, hipMemcpyAsync(rec_w, c->u, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
    CUP2D_HIP_CHECK(c, hipMemcpyAsync(rec_w + n, c->v, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
  }
  RC(pressure_single(c, 0));
  RC(enqueue_forces(c, 1, 0));
  if (rec_p)
    CUP2D_HIP_CHECK(c, hipMemcpyAsync(rec_p, c->pres, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
  return CUP2D_OK;
}

int adjoint_save(cup2d_ctx* c, int k) {
  AdjointState* a = c->adj;
  const size_t n = a->n;
  CUP2D_HIP_CHECK(c, hipMemcpyAsync(a->ck[k], c->u, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
  CUP2D_HIP_CHECK(c, hipMemcpyAsync(a->ck[k] + n, c->v, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
  CUP2D_HIP_CHECK(c, hipMemcpyAsync(&a->ck_st[k], c->st, sizeof(StepState), hipMemcpyDeviceToHost, c->stream));
  a->ck_ss[k].resize(c->g.nshapes);
  CUP2D_HIP_CHECK(c, hipMemcpyAsync(a->ck_ss[k].data(), c->ss, sizeof(cup2d_shape_state) * c->g.nshapes,
                                    hipMemcpyDeviceToHost, c->stream));
  CUP2D_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  return CUP2D_OK;
}

int adjoint_restore(cup2d_ctx* c, int k) {
  AdjointState* a = c->adj;
  const size_t n = a->n;
  RC(sync_all(c));
  CUP2D_HIP_CHECK(c, hipMemcpy(c->u, a->ck[k], sizeof(double) * n, hipMemcpyDeviceToDevice));
  CUP2D_HIP_CHECK(c, hipMemcpy(c->v, a->ck[k] + n, sizeof(double) * n, hipMemcpyDeviceToDevice));
  CUP2D_HIP_CHECK(c, hipMemcpy(c->st, &a->ck_st[k], sizeof(StepState), hipMemcpyHostToDevice));
  CUP2D_HIP_CHECK(c, hipMemcpy(c->ss, a->ck_ss[k].data(), sizeof(cup2d_shape_state) * c->g.nshapes,
                               hipMemcpyHostToDevice));
  c->host_step = a->ck_st[k].step;
  c->midA_ready = false;
  return CUP2D_OK;
}
}  // namespace

extern "C" {

const char* cup2d_last_error(const cup2d_ctx* c) {
  return c ? c->err.c_str() : g_create_error.c_str();
}

int cup2d_create(const cup2d_cfg* cfg, cup2d_ctx** out) {
  if (!cfg || !out) { g_create_error = "null argument"; return CUP2D_ERR_ARG; }
  *out = nullptr;
  if (cfg->bpdx <= 0 || cfg->bpdy <= 0 || !(cfg->extent > 0) || !(cfg->nu > 0) ||
      !(cfg->CFL > 0) || cfg->poisson < 0 || cfg->poisson > 2) {
    g_create_error = "invalid configuration (bpdx, bpdy, extent, nu, CFL > 0; poisson 0|1|2)";
    return CUP2D_ERR_ARG;
  }
  if (cfg->precision != CUP2D_PRECISION_FP64 && cfg->precision != CUP2D_PRECISION_FP32) {
    g_create_error = "precision must be CUP2D_PRECISION_FP64 or CUP2D_PRECISION_FP32";
    return CUP2D_ERR_ARG;
  }
  if (cfg->refine_sweeps < 0 || cfg->refine_sweeps > 4) {
    g_create_error = "refine_sweeps must be 0..4";
    return CUP2D_ERR_ARG;
  }
  if (cfg->precision == CUP2D_PRECISION_FP32 && cfg->iterative_penal) {
    g_create_error = "iterative penalization needs FP64 precision";
    return CUP2D_ERR_ARG;
  }
  if (cfg->precision == CUP2D_PRECISION_FP32 && cfg->refine_sweeps > 0 &&
      cfg->poisson == CUP2D_POISSON_FREESPACE) {
    g_create_error = "FP64 refinement sweeps need the periodic or cosine solver";
    return CUP2D_ERR_ARG;
  }
  cup2d_ctx* c = new cup2d_ctx();
  c->cfg = *cfg;
  c->f32 = cfg->precision == CUP2D_PRECISION_FP32;
  c->refine_sweeps = c->f32 ? cfg->refine_sweeps : 0;
  if (hipGetDevice(&c->device) != hipSuccess) {
    g_create_error = "no HIP device";
    delete c;
    return CUP2D_ERR_HIP;
  }
  Grid& g = c->g;
  g.bpdx = cfg->bpdx; g.bpdy = cfg->bpdy;
  g.NX = BS * g.bpdx; g.NY = BS * g.bpdy;
  g.hb = cfg->extent / (double)(g.bpdx > g.bpdy ? g.bpdx : g.bpdy);
  g.h = g.hb / BS;
  g.extent0 = (double)g.bpdx * g.h * BS;  // SimulationData.cpp:34-35
  g.extent1 = (double)g.bpdy * g.h * BS;
  g.nu = cfg->nu; g.CFL = cfg->CFL; g.lambda = cfg->lambda;
  g.fadeLenX = cfg->fadeLen * cfg->extent;  // Simulation.cpp:97-98
  g.fadeLenY = cfg->fadeLen * cfg->extent;
  g.implicit_penal = cfg->implicit_penal;
  g.nshapes = 0;
  g.nfish = 0;
  int rc = CUP2D_OK;
  auto bail = [&](int code) {
    g_create_error = c->err;
    cup2d_destroy(c);
    return code;
  };
  if (cfg->stream) {
    c->stream = (hipStream_t)cfg->stream;
  } else {
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
      c->err = "hipStreamCreate failed";
      return bail(CUP2D_ERR_HIP);
    }
    c->own_stream = true;
  }
  if (hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&c->mids, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_mid, CUP2D_SYNC_EVENT_FLAGS) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_begin, CUP2D_SYNC_EVENT_FLAGS) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_obst, CUP2D_SYNC_EVENT_FLAGS) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_pen, CUP2D_SYNC_EVENT_FLAGS) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_frc, CUP2D_SYNC_EVENT_FLAGS) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_lists, CUP2D_SYNC_EVENT_FLAGS) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_fork, CUP2D_SYNC_EVENT_FLAGS) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_qaux, CUP2D_SYNC_EVENT_FLAGS) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_qmid, CUP2D_SYNC_EVENT_FLAGS) != hipSuccess ||
      hipMalloc(&c->fsnap, sizeof(ForceSnap) * MAX_SHAPES) != hipSuccess ||
      hipMalloc(&c->surf_coef, sizeof(double) * 4 * SURF_MAX_MODES) != hipSuccess) {
    c->err = "side stream creation failed";
    return bail(CUP2D_ERR_HIP);
  }
  const size_t n = (size_t)g.NX * g.NY;
  double** planes[] = {&c->u, &c->v, &c->tu, &c->tv, &c->pres, &c->rhs,
                       &c->chi, &c->udx, &c->udy, &c->invrho, &c->vfu, &c->vfv};
  for (double** p : planes)
    if (hipMalloc(p, sizeof(double) * n) != hipSuccess) { c->err = "out of device memory"; return bail(CUP2D_ERR_HIP); }
  if (hipMalloc(&c->staging, sizeof(double) * 2 * n) != hipSuccess) { c->err = "out of device memory"; return bail(CUP2D_ERR_HIP); }
  if (c->f32 && (hipMalloc(&c->rres, sizeof(float) * n) != hipSuccess ||
                 hipMalloc(&c->respart, sizeof(double) * 5 * g.NY) != hipSuccess)) {
    c->err = "out of device memory";
    return bail(CUP2D_ERR_HIP);
  }
  c->red_n = 1 << 16;
  if (hipMalloc(&c->red, sizeof(double) * c->red_n) != hipSuccess) { c->err = "out of device memory"; return bail(CUP2D_ERR_HIP); }
  if (hipMalloc(&c->st, sizeof(StepState)) != hipSuccess ||
      hipMalloc(&c->sp, sizeof(ShapeParams) * MAX_SHAPES) != hipSuccess ||
      hipMalloc(&c->ss, sizeof(cup2d_shape_state) * MAX_SHAPES) != hipSuccess ||
      hipMalloc(&c->ramp, sizeof(double) * 100) != hipSuccess) {
    c->err = "out of device memory";
    return bail(CUP2D_ERR_HIP);
  }
  (void)hipMemset(c->ss, 0, sizeof(cup2d_shape_state) * MAX_SHAPES);
  (void)hipMemset(c->sp, 0, sizeof(ShapeParams) * MAX_SHAPES);
  // Simulation.cpp:264-269: rampCFL for the first 100 steps
  double ramp[100];
  for (int s = 0; s < 100; ++s) {
    const double x = (s + 1.0) / 100;
    ramp[s] = std::exp(std::log(1e-3) * (1 - x) + std::log(cfg->CFL) * x);
  }
  if (hipMemcpy(c->ramp, ramp, sizeof(ramp), hipMemcpyHostToDevice) != hipSuccess) { c->err = "copy failed"; return bail(CUP2D_ERR_HIP); }
  const std::vector<int> band = band_blocks(g);
  c->band_nblocks = (int)band.size();
  if (c->band_nblocks > 0) {
    std::vector<int> index((size_t)g.bpdx * g.bpdy, -1);
    for (size_t j = 0; j < band.size(); ++j) index[band[j]] = (int)j;
    if (hipMalloc(&c->band_blocks, sizeof(int) * band.size()) != hipSuccess ||
        hipMemcpy(c->band_blocks, band.data(), sizeof(int) * band.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMalloc(&c->band_index, sizeof(int) * index.size()) != hipSuccess ||
        hipMemcpy(c->band_index, index.data(), sizeof(int) * index.size(), hipMemcpyHostToDevice) != hipSuccess) {
      c->err = "out of device memory";
      return bail(CUP2D_ERR_HIP);
    }
  }
  if ((size_t)3 * c->band_nblocks > c->red_n) { c->err = "grid too large for fade partials"; return bail(CUP2D_ERR_ARG); }
  c->nwg = grid_workgroups(g.NX, g.NY);
  c->maxpart_n = c->nwg + OBST_LIST_CAP;
  if (hipMalloc(&c->fluxpart, sizeof(double) * c->nwg) != hipSuccess ||
      hipMalloc(&c->maxpart, sizeof(double) * c->maxpart_n) != hipSuccess ||
      hipMalloc(&c->obmap, (size_t)g.bpdx * g.bpdy) != hipSuccess ||
      hipMemset(c->obmap, 0, (size_t)g.bpdx * g.bpdy) != hipSuccess ||
      hipMalloc(&c->ubidx, sizeof(int) * g.bpdx * g.bpdy) != hipSuccess ||
      hipMemset(c->ubidx, 0xff, sizeof(int) * g.bpdx * g.bpdy) != hipSuccess ||
      hipMemset(c->fluxpart, 0, sizeof(double) * c->nwg) != hipSuccess ||
      hipMemset(c->maxpart, 0, sizeof(double) * c->maxpart_n) != hipSuccess) {
    c->err = "out of device memory";
    return bail(CUP2D_ERR_HIP);
  }
  rc = fft_create(c);
  if (rc == CUP2D_ERR_ARG) {
    // unsupported FFT size: stencil operators still work, Poisson calls fail
    c->fft_err = c->err;
    fft_destroy(c);
  } else if (rc) {
    return bail(rc);
  }
  rc = enqueue_ic(c);
  if (rc) return bail(rc);
  if (hipStreamSynchronize(c->stream) != hipSuccess) { c->err = "initialisation failed"; return bail(CUP2D_ERR_HIP); }
  *out = c;
  return CUP2D_OK;
}

int cup2d_destroy(cup2d_ctx* c) {
  if (!c) return CUP2D_ERR_ARG;
  (void)hipSetDevice(c->device);
  if (c->mids) (void)hipStreamSynchronize(c->mids);
  if (c->aux) (void)hipStreamSynchronize(c->aux);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  fft_destroy(c);
  obstacles_free(c);
  double* ptrs[] = {c->u, c->v, c->tu, c->tv, c->pres, c->rhs, c->chi, c->vfu, c->vfv,
                    c->udx, c->udy, c->invrho, c->staging, c->red, c->ramp,
                    c->fluxpart, c->maxpart, c->rres, c->respart};
  if (c->obmap) (void)hipFree(c->obmap);
  if (c->ubidx) (void)hipFree(c->ubidx);
  fish_free_all(c);
  for (double* p : ptrs)
    if (p) (void)hipFree(p);
  if (c->st) (void)hipFree(c->st);
  if (c->sp) (void)hipFree(c->sp);
  if (c->ss) (void)hipFree(c->ss);
  if (c->band_blocks) (void)hipFree(c->band_blocks);
  if (c->band_index) (void)hipFree(c->band_index);
  for (Timer& t : c->timers)
    for (hipEvent_t e : t.ev) (void)hipEventDestroy(e);
  if (c->aux) { (void)hipStreamSynchronize(c->aux); (void)hipStreamDestroy(c->aux); }
  if (c->mids) (void)hipStreamDestroy(c->mids);
  if (c->ev_mid) (void)hipEventDestroy(c->ev_mid);
  if (c->ev_begin) (void)hipEventDestroy(c->ev_begin);
  if (c->ev_obst) (void)hipEventDestroy(c->ev_obst);
  if (c->ev_pen) (void)hipEventDestroy(c->ev_pen);
  if (c->ev_frc) (void)hipEventDestroy(c->ev_frc);
  if (c->ev_lists) (void)hipEventDestroy(c->ev_lists);
  drop_graphs(c);
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_qaux) (void)hipEventDestroy(c->ev_qaux);
  if (c->ev_qmid) (void)hipEventDestroy(c->ev_qmid);
  if (c->fsnap) (void)hipFree(c->fsnap);
  if (c->surf_coef) (void)hipFree(c->surf_coef);
  adjoint_free(c);
  if (c->flog) (void)hipFree(c->flog);
  if (c->flog_cnt) (void)hipFree(c->flog_cnt);
  if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return CUP2D_OK;
}

int cup2d_add_shape(cup2d_ctx* c, const cup2d_shape_desc* d, int* shape_id) {
  CHECK_CTX(c);
  if (!d) return fail(c, CUP2D_ERR_ARG, "null shape");
  if (c->shapes_frozen) return fail(c, CUP2D_ERR_STATE, "shapes must be added before reset");
  if (c->flog) return fail(c, CUP2D_ERR_STATE, "shapes must be added before the force log is enabled");
  if (d->kind != 0 && d->kind != 1)
    return fail(c, CUP2D_ERR_ARG, "shape kind must be 0 (disk) or 1 (carlingfish)");
  if (d->kind == 0 && !(d->radius > 0)) return fail(c, CUP2D_ERR_ARG, "radius must be > 0");
  if (d->kind == 1) {
    if (!(d->length > 0) || !(d->Tperiod > 0))
      return fail(c, CUP2D_ERR_ARG, "carlingfish needs L > 0 and T > 0");
    if (d->envelope != 0 && d->envelope != 1) return fail(c, CUP2D_ERR_ARG, "envelope must be 0 or 1");
    if (d->envelope == 1 && !(d->waveLength > 0))
      return fail(c, CUP2D_ERR_ARG, "waveLength must be > 0");
    if (d->profile != 0 && d->profile != 1) return fail(c, CUP2D_ERR_ARG, "profile must be 0 or 1");
    if (d->profile == 1 && !(d->naca_t > 0)) return fail(c, CUP2D_ERR_ARG, "naca_t must be > 0");
  }
  if (c->g.nshapes >= MAX_SHAPES) return fail(c, CUP2D_ERR_ARG, "too many shapes");
  ShapeParams p{};
  p.kind = d->kind;
  p.fish = -1;
  p.bFixedx = d->bFixedx; p.bFixedy = d->bFixedy;
  p.bForcedx = d->bForcedx; p.bForcedy = d->bForcedy; p.bBlockAng = d->bBlockAng;
  p.radius = d->radius; p.rhoS = d->rhoS;
  // Shape::Shape (Shape.cpp:331-345)
  p.forcedu = -d->xvel; p.forcedv = -d->yvel; p.forcedomega = -d->angvel;
  p.timeForced = d->timeForced; p.tAccel = d->tAccel;
  p.origC[0] = d->xpos; p.origC[1] = d->ypos;
  p.origAng = d->angle_deg * M_PI / 180;
  const int id = c->g.nshapes;
  if (d->kind == 1) {
    // radius: distance from the swimmer's centre within which it always
    // lies (sizes the obstacle-block pool only)
    const double amp = d->envelope == 0
                           ? 4. / 33. * (d->length + 0.03125 * d->length)
                           : d->length * std::max(std::fabs(d->env_a0), std::fabs(d->env_a0 + d->env_a2));
    const double wmax = d->profile == 0 ? 0.04 * d->length : 0.6 * d->naca_t * d->length;
    p.radius = d->length + amp + wmax;
    p.fish = (int)c->hfish.size();
    const int rc = fish_add(c, d, id);
    if (rc) return rc;
  }
  CUP2D_HIP_CHECK(c, hipMemcpy(c->sp + id, &p, sizeof(p), hipMemcpyHostToDevice));
  c->hsp.push_back(p);
  c->g.nshapes++;
  if (shape_id) *shape_id = id;
  return CUP2D_OK;
}

int cup2d_grid_size(const cup2d_ctx* c, int* NX, int* NY) {
  if (!c || !NX || !NY) return CUP2D_ERR_ARG;
  *NX = c->g.NX;
  *NY = c->g.NY;
  return CUP2D_OK;
}

int cup2d_upload_field(cup2d_ctx* c, int f, int layout, const double* host) {
  CHECK_CTX(c);
  const int nc = nfield_comps(f);
  if (!nc || !host || (layout != CUP2D_LAYOUT_BLOCKS && layout != CUP2D_LAYOUT_DENSE))
    return fail(c, CUP2D_ERR_ARG, "upload_field: bad field, layout or pointer");
  double *a, *b;
  field_planes(c, f, &a, &b);
  const size_t n = (size_t)c->g.NX * c->g.NY;
  const bool f32 = f32_field(c, f);
  if (layout == CUP2D_LAYOUT_DENSE && f32) {
    CUP2D_HIP_CHECK(c, hipMemcpyAsync(c->staging, host, sizeof(double) * n * nc, hipMemcpyHostToDevice, c->stream));
    RC(enqueue_stage_to_f32(c, c->staging, a));
    if (nc > 1) RC(enqueue_stage_to_f32(c, c->staging + n, b));
  } else if (layout == CUP2D_LAYOUT_DENSE) {
    CUP2D_HIP_CHECK(c, hipMemcpyAsync(a, host, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    if (nc > 1)
      CUP2D_HIP_CHECK(c, hipMemcpyAsync(b, host + n, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  } else {
    CUP2D_HIP_CHECK(c, hipMemcpyAsync(c->staging, host, sizeof(double) * n * nc, hipMemcpyHostToDevice, c->stream));
    RC(enqueue_blocks_to_dense(c, c->staging, a, b, nc, f32));
  }
  CUP2D_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  return CUP2D_OK;
}

int cup2d_download_field(cup2d_ctx* c, int f, int layout, double* host) {
  CHECK_CTX(c);
  const int nc = nfield_comps(f);
  if (!nc || !host || (layout != CUP2D_LAYOUT_BLOCKS && layout != CUP2D_LAYOUT_DENSE))
    return fail(c, CUP2D_ERR_ARG, "download_field: bad field, layout or pointer");
  double *a, *b;
  field_planes(c, f, &a, &b);
  const size_t n = (size_t)c->g.NX * c->g.NY;
  const bool f32 = f32_field(c, f);
  if (layout == CUP2D_LAYOUT_DENSE && f32) {
    RC(enqueue_f32_to_stage(c, a, c->staging));
    if (nc > 1) RC(enqueue_f32_to_stage(c, b, c->staging + n));
    CUP2D_HIP_CHECK(c, hipMemcpyAsync(host, c->staging, sizeof(double) * n * nc, hipMemcpyDeviceToHost, c->stream));
  } else if (layout == CUP2D_LAYOUT_DENSE) {
    CUP2D_HIP_CHECK(c, hipMemcpyAsync(host, a, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    if (nc > 1)
      CUP2D_HIP_CHECK(c, hipMemcpyAsync(host + n, b, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  } else {
    RC(enqueue_dense_to_blocks(c, a, b, c->staging, nc, f32));
    CUP2D_HIP_CHECK(c, hipMemcpyAsync(host, c->staging, sizeof(double) * n * nc, hipMemcpyDeviceToHost, c->stream));
  }
  CUP2D_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  return CUP2D_OK;
}

int cup2d_field_device_ptr(cup2d_ctx* c, int f, int comp, void** dptr) {
  CHECK_CTX(c);
  const int nc = nfield_comps(f);
  if (!nc || comp < 0 || comp >= nc || !dptr) return fail(c, CUP2D_ERR_ARG, "bad field/comp");
  double *a, *b;
  field_planes(c, f, &a, &b);
  *dptr = comp == 0 ? a : b;
  return CUP2D_OK;
}

int cup2d_get_scalars(cup2d_ctx* c, double* time, double* dt, int* step, double uinf[2],
                      double* h) {
  CHECK_CTX(c);
  RC(sync_all(c));
  StepState s;
  CUP2D_HIP_CHECK(c, hipMemcpy(&s, c->st, sizeof(s), hipMemcpyDeviceToHost));
  if (time) *time = s.time;
  if (dt) *dt = s.dt;
  if (step) *step = s.step;
  if (uinf) { uinf[0] = s.uinf[0]; uinf[1] = s.uinf[1]; }
  if (h) *h = c->g.h;
  return CUP2D_OK;
}

int cup2d_set_scalars(cup2d_ctx* c, double time, int step, double uinfx, double uinfy) {
  CHECK_CTX(c);
  RC(sync_all(c));
  c->host_step = step;
  c->midA_ready = false;  // computed for the old time
  StepState s;
  CUP2D_HIP_CHECK(c, hipMemcpy(&s, c->st, sizeof(s), hipMemcpyDeviceToHost));
  s.time = time; s.step = step; s.uinf[0] = uinfx; s.uinf[1] = uinfy;
  CUP2D_HIP_CHECK(c, hipMemcpy(c->st, &s, sizeof(s), hipMemcpyHostToDevice));
  return CUP2D_OK;
}

int cup2d_get_shape(cup2d_ctx* c, int i, cup2d_shape_state* out) {
  CHECK_CTX(c);
  if (i < 0 || i >= c->g.nshapes || !out) return fail(c, CUP2D_ERR_ARG, "bad shape id");
  RC(sync_all(c));
  CUP2D_HIP_CHECK(c, hipMemcpy(out, c->ss + i, sizeof(*out), hipMemcpyDeviceToHost));
  return CUP2D_OK;
}

int cup2d_fish_midline(cup2d_ctx* c, int i, int* Nm, double* out) {
  CHECK_CTX(c);
  if (i < 0 || i >= c->g.nshapes || !Nm) return fail(c, CUP2D_ERR_ARG, "bad shape id");
  const int f = c->hsp[i].fish;
  if (f < 0) return fail(c, CUP2D_ERR_ARG, "shape is not a swimmer");
  const FishDev& F = c->hfish[f];
  *Nm = F.Nm;
  if (!out) return CUP2D_OK;
  RC(sync_all(c));
  const size_t n = (size_t)F.Nm;
  CUP2D_HIP_CHECK(c, hipMemcpy(out, F.rS, sizeof(double) * n, hipMemcpyDeviceToHost));
  CUP2D_HIP_CHECK(c, hipMemcpy(out + n, F.mid, sizeof(double) * 8 * n, hipMemcpyDeviceToHost));
  CUP2D_HIP_CHECK(c, hipMemcpy(out + 9 * n, F.width, sizeof(double) * n, hipMemcpyDeviceToHost));
  CUP2D_HIP_CHECK(c, hipMemcpy(out + 10 * n, F.st, sizeof(double) * 8, hipMemcpyDeviceToHost));
  return CUP2D_OK;
}

int cup2d_reset(cup2d_ctx* c) {
  CHECK_CTX(c);
  RC(ensure_frozen(c));
  RC(sync_all(c));
  drop_graphs(c);
  c->host_step = 0;
  c->midA_ready = false;
  RC(enqueue_shapes_reset(c));
  RC(enqueue_obst_reset(c));
  RC(enqueue_ic(c));
  RC(enqueue_put_objects(c, 0));
  RC(enqueue_apply_obj_vel(c));
  return sync_and_check(c);
}

int cup2d_find_max_u(cup2d_ctx* c, double* maxU) {
  CHECK_CTX(c);
  RC(enqueue_find_max_u(c));
  CUP2D_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  StepState s;
  CUP2D_HIP_CHECK(c, hipMemcpy(&s, c->st, sizeof(s), hipMemcpyDeviceToHost));
  if (maxU) *maxU = s.maxU;
  return sync_and_check(c);
}

int cup2d_calc_max_timestep(cup2d_ctx* c, double* dt) {
  CHECK_CTX(c);
  RC(enqueue_dt(c));
  CUP2D_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  StepState s;
  CUP2D_HIP_CHECK(c, hipMemcpy(&s, c->st, sizeof(s), hipMemcpyDeviceToHost));
  if (dt) *dt = s.dt;
  return sync_and_check(c);
}

int cup2d_put_objects(cup2d_ctx* c, double dt) {
  CHECK_CTX(c);
  RC(ensure_frozen(c));
  RC(enqueue_set_dt(c, dt));
  RC(enqueue_put_objects(c, 0));
  return sync_and_check(c);
}

int cup2d_advdiff(cup2d_ctx* c, double dt) {
  CHECK_CTX(c);
  RC(enqueue_set_dt(c, dt));
  RC(enqueue_advdiff(c, 0));
  return sync_and_check(c);
}

int cup2d_pressure_rhs(cup2d_ctx* c, double dt) {
  CHECK_CTX(c);
  RC(enqueue_set_dt(c, dt));
  RC(enqueue_pressure_rhs(c));
  if (c->shapes_frozen) RC(enqueue_rhs_obstacles(c));
  return sync_and_check(c);
}

int cup2d_poisson_solve(cup2d_ctx* c) {
  CHECK_CTX(c);
  RC(need_fft(c));
  RC(enqueue_fade_sums(c));
  RC(enqueue_poisson_fft(c, c->rhs, c->pres, 1));
  return sync_and_check(c);
}

int cup2d_pressure_correction(cup2d_ctx* c, double dt) {
  CHECK_CTX(c);
  RC(enqueue_set_dt(c, dt));
  RC(enqueue_pressure_correction(c, 0));
  return sync_and_check(c);
}

int cup2d_body_update(cup2d_ctx* c, double dt) {
  CHECK_CTX(c);
  if (!c->shapes_frozen) return fail(c, CUP2D_ERR_STATE, "call reset first");
  RC(enqueue_set_dt(c, dt));
  RC(enqueue_body_update(c));
  return sync_and_check(c);
}

int cup2d_penalize(cup2d_ctx* c, double dt) {
  CHECK_CTX(c);
  if (!c->shapes_frozen) return fail(c, CUP2D_ERR_STATE, "call reset first");
  RC(enqueue_set_dt(c, dt));
  RC(enqueue_penalize(c, c->maxpart + c->nwg));
  return sync_and_check(c);
}

int cup2d_pressure_single(cup2d_ctx* c, double dt) {
  CHECK_CTX(c);
  if (!c->shapes_frozen) return fail(c, CUP2D_ERR_STATE, "call reset first");
  RC(enqueue_set_dt(c, dt));
  RC(pressure_single(c, 0));
  return sync_and_check(c);
}

int cup2d_pressure_iterator(cup2d_ctx* c, double dt, int* iters, double* relDF) {
  CHECK_CTX(c);
  if (!c->shapes_frozen) return fail(c, CUP2D_ERR_STATE, "call reset first");
  RC(enqueue_set_dt(c, dt));
  RC(pressure_iterator(c, iters, relDF));
  return sync_and_check(c);
}

int cup2d_compute_forces(cup2d_ctx* c) {
  CHECK_CTX(c);
  if (!c->shapes_frozen) return fail(c, CUP2D_ERR_STATE, "call reset first");
  RC(enqueue_forces(c, 0));
  return sync_and_check(c);
}

int cup2d_advance(cup2d_ctx* c, double dt) {
  CHECK_CTX(c);
  if (!c->shapes_frozen) return fail(c, CUP2D_ERR_STATE, "call reset first");
  RC(enqueue_set_dt(c, dt));
  RC(advance_enqueued(c, 0));
  return sync_and_check(c);
}

int cup2d_simulate(cup2d_ctx* c, int nsteps) {
  CHECK_CTX(c);
  if (!c->shapes_frozen) return fail(c, CUP2D_ERR_STATE, "call reset first");
  if (nsteps < 0) return fail(c, CUP2D_ERR_ARG, "nsteps < 0");
  if (nsteps == 0) return CUP2D_OK;
  RC(enqueue_find_max_u_partials(c));
  for (int s = 0; s < nsteps; ++s) RC(advance_enqueued(c, 1));
  return CUP2D_OK;
}

int cup2d_set_surface_velocity(cup2d_ctx* c, int shape_id, int nmodes, const double* coef) {
  CHECK_CTX(c);
  if (nmodes < 0 || nmodes > SURF_MAX_MODES || (nmodes > 0 && !coef))
    return fail(c, CUP2D_ERR_ARG, "surface velocity: nmodes must be 0..32 with 4*nmodes coefficients");
  if (nmodes > 0 && (shape_id < 0 || shape_id >= c->g.nshapes))
    return fail(c, CUP2D_ERR_ARG, "bad shape id");
  if (nmodes > 0 && c->hsp[shape_id].kind != 0)
    return fail(c, CUP2D_ERR_ARG, "the surface velocity is defined on disks");
  if (c->cfg.iterative_penal)
    return fail(c, CUP2D_ERR_STATE, "the surface velocity needs PressureSingle");
  RC(sync_all(c));
  drop_graphs(c);
  if (nmodes > 0)
    CUP2D_HIP_CHECK(c, hipMemcpy(c->surf_coef, coef, sizeof(double) * 4 * nmodes,
                                 hipMemcpyHostToDevice));
  c->surf_shape = nmodes > 0 ? shape_id : -1;
  c->surf_modes = nmodes;
  return CUP2D_OK;
}

int cup2d_adjoint_forward(cup2d_ctx* c, int nsteps, int every, const double* dts_in,
                          double* dts_out, double* J) {
  CHECK_CTX(c);
  RC(adjoint_eligible(c));
  if (nsteps < 1 || every < 1) return fail(c, CUP2D_ERR_ARG, "nsteps and every must be >= 1");
  RC(sync_all(c));
  // the obstacle blocks and the fade band must not overlap (the adjoint of
  // the fade holds the RHS formed by the divergence alone)
  {
    std::vector<unsigned char> om((size_t)c->g.bpdx * c->g.bpdy);
    CUP2D_HIP_CHECK(c, hipMemcpy(om.data(), c->obmap, om.size(), hipMemcpyDeviceToHost));
    for (int by = 0; by < c->g.bpdy; ++by)
      for (int bx = 0; bx < c->g.bpdx; ++bx)
        if (om[bx + c->g.bpdx * by] && band_block(c->g, bx, by))
          return fail(c, CUP2D_ERR_STATE, "the disk's blocks reach the boundary fade band");
  }
  const size_t n = (size_t)c->g.NX * c->g.NY;
  const int nseg = (nsteps + every - 1) / every;
  AdjointState* a = c->adj;
  if (!a || a->n != n || a->every != every || a->nseg < nseg) {
    adjoint_free(c);
    a = c->adj = new AdjointState();
    a->n = n;
    a->every = every;
    a->nseg = nseg;
    a->ck.assign(nseg + 1, nullptr);
    for (double*& p : a->ck) CUP2D_HIP_CHECK(c, hipMalloc(&p, sizeof(double) * 2 * n));
    CUP2D_HIP_CHECK(c, hipMalloc(&a->seg, sizeof(double) * n * (2 * (every + 1) + 3 * every)));
    CUP2D_HIP_CHECK(c, hipMalloc(&a->work, sizeof(double) * n * 8));
    const int G = c->ob.cap > 0 ? c->ob.cap : 1;
    const int nblk = std::max(G, c->band_nblocks);
    CUP2D_HIP_CHECK(c, hipMalloc(&a->part, sizeof(double) * 5 * (size_t)(nblk + 1)));
    CUP2D_HIP_CHECK(c, hipMalloc(&a->theta_part, sizeof(double) * 4 * SURF_MAX_MODES * (size_t)G));
    CUP2D_HIP_CHECK(c, hipMalloc(&a->theta, sizeof(double) * 4 * SURF_MAX_MODES));
    CUP2D_HIP_CHECK(c, hipMalloc(&a->scal, sizeof(double) * 4));
  }
  a->ck_st.assign(nseg + 1, StepState{});
  a->ck_ss.assign(nseg + 1, {});
  a->nsteps = nsteps;
  a->dts.assign(nsteps, 0.0);
  a->ready = false;
  drop_graphs(c);
  RC(enqueue_adjoint_fill(c, a->scal, 1, 0.0));
  RC(adjoint_save(c, 0));
  const double wJ = 1.0 / nsteps;
  for (int k = 0; k < nsteps; ++k) {
    double dt;
    if (dts_in) {
      dt = dts_in[k];
    } else {  // Simulation::calcMaxTimestep
      RC(enqueue_dt(c));
      StepState s;
      CUP2D_HIP_CHECK(c, hipMemcpyAsync(&s, c->st, sizeof(s), hipMemcpyDeviceToHost, c->stream));
      CUP2D_HIP_CHECK(c, hipStreamSynchronize(c->stream));
      dt = s.dt;
    }
    a->dts[k] = dt;
    RC(adjoint_fwd_step(c, dt, nullptr, nullptr));
    c->host_step += 1;
    RC(enqueue_adjoint_acc_drag(c, 0, wJ, a->scal));
    if ((k + 1) % every == 0 && k + 1 < nsteps) RC(adjoint_save(c, (k + 1) / every));
  }
  RC(adjoint_save(c, nseg));  // the final state, restored after the backward sweep
  {
    const StepState& s = a->ck_st[nseg];
    a->uinf0 = s.uinf[0];
    a->uinf1 = s.uinf[1];
    const cup2d_shape_state& sh = a->ck_ss[nseg][0];
    const double vn = std::sqrt(sh.u * sh.u + sh.v * sh.v);
    a->vu0 = vn > 0 ? sh.u / vn : 0;
    a->vu1 = vn > 0 ? sh.v / vn : 0;
  }
  double Jh = 0;
  CUP2D_HIP_CHECK(c, hipMemcpy(&Jh, a->scal, sizeof(double), hipMemcpyDeviceToHost));
  if (J) *J = Jh;
  if (dts_out) std::memcpy(dts_out, a->dts.data(), sizeof(double) * nsteps);
  a->ready = true;
  return sync_and_check(c);
}

int cup2d_adjoint_backward(cup2d_ctx* c, double* grad) {
  CHECK_CTX(c);
  AdjointState* a = c->adj;
  if (!a || !a->ready) return fail(c, CUP2D_ERR_STATE, "run cup2d_adjoint_forward first");
  if (!grad) return fail(c, CUP2D_ERR_ARG, "null gradient");
  RC(adjoint_eligible(c));
  const size_t n = a->n;
  const int K = a->every, N = a->nsteps;
  double* ub = a->work;
  double* vb = ub + n;
  double* pb = vb + n;
  double* rb = pb + n;
  double* wbu = rb + n;
  double* wbv = wbu + n;
  double* ub2 = wbv + n;
  double* vb2 = ub2 + n;
  double* segU = a->seg;                       // (K + 1) x (u, v)
  double* segW = segU + 2 * n * (K + 1);       // K x (wu, wv)
  double* segP = segW + 2 * n * K;             // K x p
  const int M = c->surf_modes;
  RC(enqueue_adjoint_fill(c, ub, 2 * n, 0.0));
  RC(enqueue_adjoint_fill(c, a->theta, 4 * M, 0.0));
  const int nseg = (N + K - 1) / K;
  for (int sgi = nseg - 1; sgi >= 0; --sgi) {
    const int k0 = sgi * K, len = std::min(K, N - k0);
    RC(adjoint_restore(c, sgi));
    // recompute the segment, storing each step's input, advDiff output and pressure
    for (int j = 0; j < len; ++j) {
      CUP2D_HIP_CHECK(c, hipMemcpyAsync(segU + 2 * n * j, c->u, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
      CUP2D_HIP_CHECK(c, hipMemcpyAsync(segU + 2 * n * j + n, c->v, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
      RC(adjoint_fwd_step(c, a->dts[k0 + j], segW + 2 * n * j, segP + n * j));
    }
    CUP2D_HIP_CHECK(c, hipMemcpyAsync(segU + 2 * n * len, c->u, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
    CUP2D_HIP_CHECK(c, hipMemcpyAsync(segU + 2 * n * len + n, c->v, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
    for (int j = len - 1; j >= 0; --j) {
      AdjStep s{};
      s.shape = 0;
      s.dt = a->dts[k0 + j];
      s.uinf0 = a->uinf0;
      s.uinf1 = a->uinf1;
      s.vu0 = a->vu0;
      s.vu1 = a->vu1;
      s.wJ = 1.0 / N;
      s.u0 = segU + 2 * n * j;
      s.v0 = s.u0 + n;
      s.wu = segW + 2 * n * j;
      s.wv = s.wu + n;
      s.u1 = segU + 2 * n * (j + 1);
      s.v1 = s.u1 + n;
      s.p1 = segP + n * j;
      s.ub = ub; s.vb = vb; s.pb = pb; s.rb = rb; s.wbu = wbu; s.wbv = wbv;
      s.part = a->part;
      s.theta_part = a->theta_part;
      s.theta = a->theta;
      s.fluxsum = a->scal + 1;
      (void)ub2; (void)vb2;
      RC(enqueue_adjoint_step(c, s));
    }
  }
  CUP2D_HIP_CHECK(c, hipMemcpyAsync(grad, a->theta, sizeof(double) * 4 * M, hipMemcpyDeviceToHost, c->stream));
  RC(adjoint_restore(c, nseg));  // back to the state after the forward run
  return sync_and_check(c);
}

int cup2d_refine_residuals(cup2d_ctx* c, double* res, int* nsweeps) {
  CHECK_CTX(c);
  if (!res || !nsweeps) return fail(c, CUP2D_ERR_ARG, "null pointer");
  RC(sync_all(c));
  StepState s;
  CUP2D_HIP_CHECK(c, hipMemcpy(&s, c->st, sizeof(s), hipMemcpyDeviceToHost));
  const int n = c->fft && !c->fft_err.size() && c->cfg.poisson != CUP2D_POISSON_FREESPACE
                    ? c->refine_sweeps : 0;
  for (int k = 0; k < n; ++k) res[k] = s.res[k];
  *nsweeps = n;
  return CUP2D_OK;
}

int cup2d_simulate_graph(cup2d_ctx* c, int nsteps) {
  CHECK_CTX(c);
  if (nsteps < 0) return fail(c, CUP2D_ERR_ARG, "nsteps < 0");
  RC(graph_eligible(c));
  if (nsteps == 0) return CUP2D_OK;
  RC(enqueue_quiesce(c));
  RC(enqueue_find_max_u_partials(c));
  for (int left = nsteps; left > 0;) {
    if (left == 1) {
      RC(advance_enqueued(c, 1));
      RC(enqueue_quiesce(c));
      break;
    }
    hipGraphExec_t ex = nullptr;
    RC(step_graph(c, &ex));
    CUP2D_HIP_CHECK(c, hipGraphLaunch(ex, c->stream));
    c->host_step += 2;
    left -= 2;
  }
  return launch_check(c, "simulate_graph");
}

int cup2d_synchronize(cup2d_ctx* c) {
  CHECK_CTX(c);
  return sync_and_check(c);
}

int cup2d_poisson_dense(cup2d_ctx* c, const double* rhs, double* sol) {
  CHECK_CTX(c);
  if (!rhs || !sol) return fail(c, CUP2D_ERR_ARG, "null pointer");
  RC(need_fft(c));
  const size_t n = (size_t)c->g.NX * c->g.NY;
  double* din = c->staging;
  double* dout = c->staging + n;
  CUP2D_HIP_CHECK(c, hipMemcpyAsync(din, rhs, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  if (c->f32) {
    // the float RHS in the second half of the staging buffer; the solution
    // (double) replaces the uploaded RHS
    RC(enqueue_stage_to_f32(c, din, dout));
    RC(enqueue_poisson_fft(c, dout, din, 0));
    std::swap(din, dout);
  } else {
    RC(enqueue_poisson_fft(c, din, dout, 0));
  }
  CUP2D_HIP_CHECK(c, hipMemcpyAsync(sol, dout, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  CUP2D_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  return CUP2D_OK;
}

int cup2d_force_log_enable(cup2d_ctx* c, int capacity) {
  CHECK_CTX(c);
  if (capacity < 0) return fail(c, CUP2D_ERR_ARG, "force log capacity < 0");
  RC(sync_all(c));
  drop_graphs(c);  // they hold the log's buffers as launch arguments
  if (c->flog) (void)hipFree(c->flog);
  if (c->flog_cnt) (void)hipFree(c->flog_cnt);
  c->flog = nullptr;
  c->flog_cnt = nullptr;
  c->flog_cap = 0;
  if (capacity == 0 || c->g.nshapes == 0) return CUP2D_OK;
  const size_t n = (size_t)capacity * c->g.nshapes * FLOG_REC;
  CUP2D_HIP_CHECK(c, hipMalloc(&c->flog, sizeof(double) * n));
  CUP2D_HIP_CHECK(c, hipMalloc(&c->flog_cnt, sizeof(unsigned long long)));
  CUP2D_HIP_CHECK(c, hipMemset(c->flog_cnt, 0, sizeof(unsigned long long)));
  c->flog_cap = capacity;
  return CUP2D_OK;
}

int cup2d_force_log_read(cup2d_ctx* c, int shape_id, int max_records, double* out, int* nrecords) {
  CHECK_CTX(c);
  if (!nrecords || max_records < 0 || (max_records > 0 && !out))
    return fail(c, CUP2D_ERR_ARG, "force_log_read: bad arguments");
  if (shape_id < 0 || shape_id >= c->g.nshapes) return fail(c, CUP2D_ERR_ARG, "bad shape id");
  if (!c->flog) return fail(c, CUP2D_ERR_STATE, "force log not enabled");
  RC(sync_all(c));
  unsigned long long cnt = 0;
  CUP2D_HIP_CHECK(c, hipMemcpy(&cnt, c->flog_cnt, sizeof(cnt), hipMemcpyDeviceToHost));
  const unsigned long long held = cnt < (unsigned long long)c->flog_cap ? cnt : c->flog_cap;
  const unsigned long long n = held < (unsigned long long)max_records ? held : max_records;
  const size_t ns = (size_t)c->g.nshapes;
  std::vector<double> all((size_t)c->flog_cap * ns * FLOG_REC);
  CUP2D_HIP_CHECK(c, hipMemcpy(all.data(), c->flog, sizeof(double) * all.size(),
                               hipMemcpyDeviceToHost));
  // the n most recent records, oldest first
  for (unsigned long long j = 0; j < n; ++j) {
    const unsigned long long r = cnt - n + j;
    const double* src = all.data() + ((r % c->flog_cap) * ns + shape_id) * FLOG_REC;
    std::memcpy(out + j * FLOG_REC, src, sizeof(double) * FLOG_REC);
  }
  *nrecords = (int)n;
  return CUP2D_OK;
}

int cup2d_timer_enable(cup2d_ctx* c, const char* kernel, int reserve_launches) {
  CHECK_CTX(c);
  const bool all = kernel && std::strcmp(kernel, "*") == 0;
  const int id = timer_index(kernel);
  if (!all && id < 0) return fail(c, CUP2D_ERR_ARG, "unknown timer name");
  CUP2D_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  for (int i = 0; i < T_NTIMERS; ++i) {
    if (!all && i != id) continue;
    Timer& t = c->timers[i];
    t.on = reserve_launches > 0;
    t.used = 0;
    t.seen = 0;
    t.open = false;
    while (t.ev.size() < 2 * (size_t)(reserve_launches > 0 ? reserve_launches : 0)) {
      hipEvent_t e;
      CUP2D_HIP_CHECK(c, hipEventCreate(&e));
      t.ev.push_back(e);
    }
  }
  return CUP2D_OK;
}

int cup2d_timer_sample(cup2d_ctx* c, const char* kernel, int every) {
  CHECK_CTX(c);
  const bool all = kernel && std::strcmp(kernel, "*") == 0;
  const int id = timer_index(kernel);
  if ((!all && id < 0) || every < 1) return fail(c, CUP2D_ERR_ARG, "timer_sample: bad name or n");
  for (int i = 0; i < T_NTIMERS; ++i)
    if (all || i == id) {
      c->timers[i].every = every;
      c->timers[i].seen = 0;
      c->timers[i].open = false;
    }
  return CUP2D_OK;
}

int cup2d_timer_read(cup2d_ctx* c, const char* kernel, double* total_ms, int* launches) {
  CHECK_CTX(c);
  const int id = timer_index(kernel);
  if (id < 0 || !total_ms || !launches) return fail(c, CUP2D_ERR_ARG, "unknown timer name");
  CUP2D_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  Timer& t = c->timers[id];
  double tot = 0;
  for (int k = 0; k < t.used; ++k) {
    float ms = 0;
    CUP2D_HIP_CHECK(c, hipEventElapsedTime(&ms, t.ev[2 * (size_t)k], t.ev[2 * (size_t)k + 1]));
    tot += ms;
  }
  *total_ms = tot;
  *launches = t.used;
  t.used = 0;
  return CUP2D_OK;
}

}  // extern "C"

// ------------------------------------------------------------- ensembles
// Many independent simulations (the reference runs one Simulation per
// process under an RL driver, main_RL_smartCyl.cpp:102-205) advanced by
// captured graphs: two whole steps of every member per launch.  The members
// are spread over `lanes` lanes; each lane is one captured graph (its
// members one after another) launched on its own stream, the lanes side by
// side.  The autotuner times every lane count on the live run and keeps the
// fastest.
struct cup2d_ensemble {
  std::vector<cup2d_ctx*> m;
  hipStream_t stream = nullptr;     // orders the lanes after and before the members
  std::vector<hipStream_t> lane_stream;
  hipEvent_t ev_start = nullptr, ev_end = nullptr;
  std::vector<hipEvent_t> ev_lane, ev_done;
  hipEvent_t t0 = nullptr, t1 = nullptr;  // timing (autotuner)
  int lanes = 1;
  std::vector<hipGraphExec_t> exec;
  int exec_lanes = -1, exec_parity = -1;
  std::vector<const double*> u0;
  std::vector<const int*> ob0;
  std::string err;
};

namespace {
int efail(cup2d_ensemble* e, int code, const std::string& msg) {
  e->err = msg;
  return code;
}
int member_error(cup2d_ensemble* e, size_t i, int rc) {
  e->err = "member " + std::to_string(i) + ": " + e->m[i]->err;
  return rc;
}
#define EHIP(e, call)                                                     \
  do {                                                                    \
    hipError_t e_ = (call);                                               \
    if (e_ != hipSuccess) return efail((e), CUP2D_ERR_HIP, std::string(#call) + ": " + \
                                       hipGetErrorString(e_));            \
  } while (0)

bool ens_graph_valid(const cup2d_ensemble* e, int lanes) {
  if (e->exec.empty() || e->exec_lanes != lanes ||
      e->exec_parity != (e->m[0]->host_step & 1))
    return false;
  for (size_t i = 0; i < e->m.size(); ++i)
    if (e->u0[i] != e->m[i]->u || e->ob0[i] != e->m[i]->ob.cnt) return false;
  return true;
}

void ens_drop(cup2d_ensemble* e) {
  for (hipGraphExec_t x : e->exec)
    if (x) (void)hipGraphExecDestroy(x);
  e->exec.clear();
  e->exec_lanes = -1;
}

// One graph per lane: lane l holds members l, l + L, l + 2L, ... one after
// another, each forked from the lane's stream (or the previous member) and
// joined back.
int ens_capture(cup2d_ensemble* e, int lanes) {
  ens_drop(e);
  const int n = (int)e->m.size();
  const int L = lanes <= 0 || lanes > n ? n : lanes;
  while ((int)e->lane_stream.size() < L) {
    hipStream_t st = nullptr;
    EHIP(e, hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    e->lane_stream.push_back(st);
  }
  for (int l = 0; l < L; ++l) {
    hipStream_t ls = e->lane_stream[l];
    EHIP(e, hipStreamBeginCapture(ls, hipStreamCaptureModeThreadLocal));
    int rc = CUP2D_OK;
    if (hipEventRecord(e->ev_lane[l], ls) != hipSuccess) rc = CUP2D_ERR_HIP;
    for (int i = l; i < n && !rc; i += L) {
      cup2d_ctx* c = e->m[i];
      if (hipStreamWaitEvent(c->stream, e->ev_lane[l], 0) != hipSuccess) {
        rc = CUP2D_ERR_HIP;
        break;
      }
      rc = capture_two_steps(c);
      if (rc) {
        member_error(e, i, rc);
        break;
      }
      if (hipEventRecord(e->ev_lane[l], c->stream) != hipSuccess) rc = CUP2D_ERR_HIP;
    }
    if (!rc && hipStreamWaitEvent(ls, e->ev_lane[l], 0) != hipSuccess) rc = CUP2D_ERR_HIP;
    hipGraph_t graph = nullptr;
    const hipError_t ee = hipStreamEndCapture(ls, &graph);
    if (rc || ee != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      ens_drop(e);
      if (!rc) return efail(e, CUP2D_ERR_HIP, std::string("hipStreamEndCapture: ") +
                                                   hipGetErrorString(ee));
      if (e->err.empty()) e->err = "capture failed";
      return rc;
    }
    hipGraphExec_t x = nullptr;
    const hipError_t ei = hipGraphInstantiate(&x, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) {
      ens_drop(e);
      return efail(e, CUP2D_ERR_HIP, std::string("hipGraphInstantiate: ") +
                                         hipGetErrorString(ei));
    }
    e->exec.push_back(x);
  }
  e->exec_lanes = lanes;
  e->exec_parity = e->m[0]->host_step & 1;
  for (size_t i = 0; i < e->m.size(); ++i) {
    e->u0[i] = e->m[i]->u;
    e->ob0[i] = e->m[i]->ob.cnt;
  }
  return CUP2D_OK;
}

// members quiescent, with their findMaxU partials, and the ensemble stream
// ordered after their streams
int ens_prepare(cup2d_ensemble* e) {
  for (size_t i = 0; i < e->m.size(); ++i) {
    cup2d_ctx* c = e->m[i];
    int rc = flush_forces(c);
    if (!rc) rc = graph_eligible(c);
    if (!rc) rc = enqueue_quiesce(c);
    if (!rc) rc = enqueue_find_max_u_partials(c);
    if (rc) return member_error(e, i, rc);
    EHIP(e, hipEventRecord(e->ev_lane[0], c->stream));
    EHIP(e, hipStreamWaitEvent(e->stream, e->ev_lane[0], 0));
  }
  return CUP2D_OK;
}

// every member's stream ordered after the ensemble stream
int ens_release(cup2d_ensemble* e) {
  EHIP(e, hipEventRecord(e->ev_end, e->stream));
  for (cup2d_ctx* c : e->m) EHIP(e, hipStreamWaitEvent(c->stream, e->ev_end, 0));
  return CUP2D_OK;
}

int ens_launch(cup2d_ensemble* e, int lanes) {
  if (!ens_graph_valid(e, lanes)) {
    const int rc = ens_capture(e, lanes);
    if (rc) return rc;
  }
  EHIP(e, hipEventRecord(e->ev_start, e->stream));
  for (size_t l = 0; l < e->exec.size(); ++l) {
    hipStream_t ls = e->lane_stream[l];
    EHIP(e, hipStreamWaitEvent(ls, e->ev_start, 0));
    EHIP(e, hipGraphLaunch(e->exec[l], ls));
    EHIP(e, hipEventRecord(e->ev_done[l], ls));
    EHIP(e, hipStreamWaitEvent(e->stream, e->ev_done[l], 0));
  }
  for (cup2d_ctx* c : e->m) c->host_step += 2;
  return CUP2D_OK;
}
}  // namespace

extern "C" {

int cup2d_ensemble_create(cup2d_ctx* const* ctxs, int n, cup2d_ensemble** out) {
  if (!ctxs || n <= 0 || !out) {
    g_create_error = "ensemble: null members or n <= 0";
    return CUP2D_ERR_ARG;
  }
  *out = nullptr;
  for (int i = 0; i < n; ++i) {
    if (!ctxs[i] || ctxs[i]->device != ctxs[0]->device) {
      g_create_error = "ensemble: null member or members on different devices";
      return CUP2D_ERR_ARG;
    }
    for (int j = 0; j < i; ++j)
      if (ctxs[j] == ctxs[i]) {
        g_create_error = "ensemble: a context appears twice";
        return CUP2D_ERR_ARG;
      }
  }
  (void)hipSetDevice(ctxs[0]->device);
  cup2d_ensemble* e = new cup2d_ensemble();
  e->m.assign(ctxs, ctxs + n);
  e->u0.assign(n, nullptr);
  e->ob0.assign(n, nullptr);
  e->ev_lane.assign(n, nullptr);
  e->ev_done.assign(n, nullptr);
  bool ok = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) == hipSuccess &&
            hipEventCreateWithFlags(&e->ev_start, CUP2D_SYNC_EVENT_FLAGS) == hipSuccess &&
            hipEventCreateWithFlags(&e->ev_end, CUP2D_SYNC_EVENT_FLAGS) == hipSuccess &&
            hipEventCreate(&e->t0) == hipSuccess && hipEventCreate(&e->t1) == hipSuccess;
  for (int i = 0; i < n && ok; ++i)
    ok = hipEventCreateWithFlags(&e->ev_lane[i], CUP2D_SYNC_EVENT_FLAGS) == hipSuccess &&
         hipEventCreateWithFlags(&e->ev_done[i], CUP2D_SYNC_EVENT_FLAGS) == hipSuccess;
  if (!ok) {
    g_create_error = "ensemble: stream or event creation failed";
    cup2d_ensemble_destroy(e);
    return CUP2D_ERR_HIP;
  }
  *out = e;
  return CUP2D_OK;
}

int cup2d_ensemble_destroy(cup2d_ensemble* e) {
  if (!e) return CUP2D_ERR_ARG;
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  for (hipStream_t ls : e->lane_stream) (void)hipStreamSynchronize(ls);
  ens_drop(e);
  for (hipStream_t ls : e->lane_stream) (void)hipStreamDestroy(ls);
  for (hipEvent_t v : e->ev_lane)
    if (v) (void)hipEventDestroy(v);
  for (hipEvent_t v : e->ev_done)
    if (v) (void)hipEventDestroy(v);
  for (hipEvent_t v : {e->ev_start, e->ev_end, e->t0, e->t1})
    if (v) (void)hipEventDestroy(v);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
  return CUP2D_OK;
}

const char* cup2d_ensemble_last_error(const cup2d_ensemble* e) {
  return e ? e->err.c_str() : g_create_error.c_str();
}

int cup2d_ensemble_set_lanes(cup2d_ensemble* e, int lanes) {
  if (!e) return CUP2D_ERR_ARG;
  if (lanes < 0 || lanes > (int)e->m.size())
    return efail(e, CUP2D_ERR_ARG, "lanes must be 0..members (0: one lane per member)");
  e->lanes = lanes;
  return CUP2D_OK;
}

int cup2d_ensemble_simulate(cup2d_ensemble* e, int nsteps) {
  if (!e) return CUP2D_ERR_ARG;
  if (nsteps < 0) return efail(e, CUP2D_ERR_ARG, "nsteps < 0");
  if (nsteps == 0) return CUP2D_OK;
  (void)hipSetDevice(e->m[0]->device);
  const int parity = e->m[0]->host_step & 1;
  for (cup2d_ctx* c : e->m)
    if ((c->host_step & 1) != parity)
      return efail(e, CUP2D_ERR_STATE, "members must be at steps of the same parity");
  RC(ens_prepare(e));
  for (int left = nsteps; left >= 2; left -= 2) RC(ens_launch(e, e->lanes));
  RC(ens_release(e));
  if (nsteps % 2) {
    for (size_t i = 0; i < e->m.size(); ++i) {
      cup2d_ctx* c = e->m[i];
      int rc = advance_enqueued(c, 1);
      if (!rc) rc = enqueue_quiesce(c);
      if (rc) return member_error(e, i, rc);
    }
  }
  return CUP2D_OK;
}

int cup2d_ensemble_synchronize(cup2d_ensemble* e) {
  if (!e) return CUP2D_ERR_ARG;
  EHIP(e, hipStreamSynchronize(e->stream));
  for (size_t i = 0; i < e->m.size(); ++i) {
    cup2d_ctx* c = e->m[i];
    (void)hipSetDevice(c->device);
    const int rc = sync_and_check(c);
    if (rc) return member_error(e, i, rc);
  }
  return CUP2D_OK;
}

int cup2d_ensemble_tune(cup2d_ensemble* e, int launches_per_candidate, int* lanes_out,
                        double* ms_per_launch_out) {
  if (!e || launches_per_candidate < 1) return CUP2D_ERR_ARG;
  (void)hipSetDevice(e->m[0]->device);
  const int n = (int)e->m.size();
  std::vector<int> cand;
  for (int l = 1; l < n; l *= 2) cand.push_back(l);
  cand.push_back(n);
  const int parity = e->m[0]->host_step & 1;
  for (cup2d_ctx* c : e->m)
    if ((c->host_step & 1) != parity)
      return efail(e, CUP2D_ERR_STATE, "members must be at steps of the same parity");
  RC(ens_prepare(e));
  double best = 1e300;
  int best_l = n;
  for (int l : cand) {
    RC(ens_launch(e, l));  // captures; the first replay warms up
    EHIP(e, hipEventRecord(e->t0, e->stream));
    for (int k = 0; k < launches_per_candidate; ++k) RC(ens_launch(e, l));
    EHIP(e, hipEventRecord(e->t1, e->stream));
    EHIP(e, hipEventSynchronize(e->t1));
    float ms = 0;
    EHIP(e, hipEventElapsedTime(&ms, e->t0, e->t1));
    const double per = ms / launches_per_candidate;
    if (per < best) {
      best = per;
      best_l = l;
    }
  }
  e->lanes = best_l;
  RC(ens_release(e));
  if (lanes_out) *lanes_out = best_l;
  if (ms_per_launch_out) *ms_per_launch_out = best;
  return CUP2D_OK;
}

}  // extern "C"
