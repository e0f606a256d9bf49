// SimulationData.cpp -- device context lifetime and host <-> device state.
#include "SimulationData.h"

#include <cstdio>

#include "DumpHDF5.h"

#include <algorithm>
#include <cmath>

namespace cup2d {

void check(cup2d_ctx* ctx, int status, const char* what) {
  if (status == CUP2D_OK) return;
  std::string msg = std::string(what) + ": " + cup2d_last_error(ctx);
  throw Error(status, msg);
}

double Shape::getMaxVel() const { return std::sqrt(state.u * state.u + state.v * state.v); }

SimulationData::~SimulationData() {
  if (ctx) cup2d_destroy(ctx);
}

void SimulationData::allocateGrid() {
  if (ctx) throw Error(CUP2D_ERR_STATE, "allocateGrid: grid already allocated");
  if (bpdx <= 0 || bpdy <= 0) throw Error(CUP2D_ERR_ARG, "allocateGrid: bpdx, bpdy must be > 0");
  // extents of the domain: the longer side has length `extent`
  // (SimulationData.cpp:34-35: h is the same in both directions)
  const int mx = std::max(bpdx, bpdy);
  extents[0] = extent * bpdx / mx;
  extents[1] = extent * bpdy / mx;
  int kind;
  if (poissonType == "periodic" || poissonType == "cpu_periodic")
    kind = CUP2D_POISSON_PERIODIC;
  else if (poissonType == "freespace" || poissonType == "cpu_freespace")
    kind = CUP2D_POISSON_FREESPACE;
  else if (poissonType == "" || poissonType == "cosine" || poissonType == "cpu_cosine")
    kind = CUP2D_POISSON_COSINE;  // the default: FFTW_dirichlet (PoissonSolver.cpp:57-58)
  else
    throw Error(CUP2D_ERR_ARG, "poissonType '" + poissonType +
                                   "' has no device solver (periodic, freespace, cosine)");
  cup2d_cfg cfg{};
  cfg.bpdx = bpdx;
  cfg.bpdy = bpdy;
  cfg.extent = extent;
  cfg.nu = nu;
  cfg.CFL = CFL;
  cfg.lambda = lambda;
  cfg.fadeLen = fadeLen;
  fadeLenX = fadeLen * extent;
  fadeLenY = fadeLen * extent;
  cfg.poisson = kind;
  cfg.implicit_penal = 0;
  cfg.iterative_penal = iterativePenalization ? 1 : 0;
  cfg.stream = stream;
  // -precision fp32 is the reference's float build (_FLOAT_PRECISION_,
  // Definitions.h:18-22) with the Poisson solve refined in FP64
  if (precision == "fp64" || precision == "")
    cfg.precision = CUP2D_PRECISION_FP64;
  else if (precision == "fp32")
    cfg.precision = CUP2D_PRECISION_FP32;
  else
    throw Error(CUP2D_ERR_ARG, "precision '" + precision + "' (fp64, fp32)");
  cfg.refine_sweeps = cfg.precision == CUP2D_PRECISION_FP32 && kind != CUP2D_POISSON_FREESPACE
                          ? refineSweeps : 0;
  check(nullptr, cup2d_create(&cfg, &ctx), "cup2d_create");
}

void SimulationData::addShape(Shape s) {
  int id = -1;
  check(ctx, cup2d_add_shape(ctx, &s.desc, &id), "cup2d_add_shape");
  s.obstacleID = id;
  shapes.push_back(s);
}

void SimulationData::resetAll() {
  time = 0;
  step = 0;
  uinfx = 0;
  uinfy = 0;
  nextDumpTime = 0;
}

bool SimulationData::bDump() const {
  const bool timeDump = dumpTime > 0 && time >= nextDumpTime;
  const bool stepDump = dumpFreq > 0 && (step % dumpFreq) == 0;
  return stepDump || timeDump;
}

void SimulationData::registerDump() { nextDumpTime += dumpTime; }

void SimulationData::dumpAll(const std::string& name) {
  const size_t n = (size_t)NX() * NY();
  std::vector<double> v(2 * n), c(n), p(n), r(n);
  download(vel, v.data(), false);
  download(chi, c.data(), false);
  download(pres, p.data(), false);
  download(invRho, r.data(), false);
  char num[16];
  std::snprintf(num, sizeof(num), "%07d", step);
  const double h = getH();
  const double* glue[3] = {v.data(), v.data() + n, c.data()};
  const double* ps[1] = {p.data()};
  const double* rs[1] = {r.data()};
  if (write_h5_xmf(path4serialization, "velChi_" + name + num, time, step, NX(), NY(), h, 3, glue,
                   "Vector") ||
      write_h5_xmf(path4serialization, "pres_" + name + num, time, step, NX(), NY(), h, 1, ps,
                   "Scalar") ||
      write_h5_xmf(path4serialization, "invRho_" + name + num, time, step, NX(), NY(), h, 1, rs,
                   "Scalar"))
    throw Error(CUP2D_ERR_ARG, "dumpAll: cannot write into " + path4serialization);
}

bool SimulationData::bOver() const {
  const bool timeEnd = endTime > 0 && time >= endTime;
  const bool stepEnd = nsteps > 0 && step > nsteps;
  return timeEnd || stepEnd;
}

double SimulationData::maxRelSpeed() const {
  double m = 0;
  for (const auto& s : shapes) m = std::max(m, s.getMaxVel());
  return m;
}

double SimulationData::getH() const {
  return extent / (double)std::max(NX(), NY());
}

void SimulationData::pullScalars() {
  double uinf[2];
  check(ctx, cup2d_get_scalars(ctx, &time, &dt, &step, uinf, nullptr), "cup2d_get_scalars");
  uinfx = uinf[0];
  uinfy = uinf[1];
}

void SimulationData::pushScalars() {
  check(ctx, cup2d_set_scalars(ctx, time, step, uinfx, uinfy), "cup2d_set_scalars");
}

void SimulationData::refreshShapes() {
  for (auto& s : shapes)
    check(ctx, cup2d_get_shape(ctx, s.obstacleID, &s.state), "cup2d_get_shape");
}

void SimulationData::upload(const Field& f, const double* host, bool blocks) {
  check(ctx,
        cup2d_upload_field(ctx, f.id, blocks ? CUP2D_LAYOUT_BLOCKS : CUP2D_LAYOUT_DENSE, host),
        "cup2d_upload_field");
}

void SimulationData::download(const Field& f, double* host, bool blocks) const {
  check(ctx,
        cup2d_download_field(ctx, f.id, blocks ? CUP2D_LAYOUT_BLOCKS : CUP2D_LAYOUT_DENSE, host),
        "cup2d_download_field");
}

}  // namespace cup2d
