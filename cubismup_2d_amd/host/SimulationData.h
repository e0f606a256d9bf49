// SimulationData.h -- host-side state of one simulation on the GPU.
//
// Mirrors the members of SimulationData (SimulationData.h:17-116) that the
// per-timestep operators read: grid size, extent, nu, CFL, lambda, fadeLen,
// poissonType, time, dt, step, U_inf and the shapes.  The grids themselves
// live on the device inside the cup2d_ctx; `vel`, `chi`, `pres`, `tmp`,
// `tmpV`, `uDef` and `invRho` are field handles that name them.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "cup2d.h"

namespace cup2d {

// Error raised where the reference aborts or exits (Helpers.cpp:205-285,
// PutObjectsOnGrid.cpp:213-218, Simulation.cpp:163-171): the status code of
// the failing C-ABI call and its message.
struct Error : std::runtime_error {
  int status;
  Error(int st, const std::string& msg) : std::runtime_error(msg), status(st) {}
};

// Throws cup2d::Error when a C-ABI call fails.
void check(cup2d_ctx* ctx, int status, const char* what);

// A field of the grid (the role of the reference's ScalarGrid/VectorGrid
// pointers, SimulationData.h:36-44).
struct Field {
  int id;
  int ncomp;
};

// One body: the Shape parameters parsed by Shape::Shape and Disk::Disk
// (Shape.cpp:331-345, ShapesSimple.h:19-21) and its state after the last
// step (Shape.h:22-62), refreshed by SimulationData::refreshShapes.
struct Shape {
  std::string name;
  int obstacleID = 0;
  cup2d_shape_desc desc{};
  cup2d_shape_state state{};
  double getMaxVel() const;  // Shape::getMaxVel, Shape.cpp:20: sqrt(u^2 + v^2)
};

class SimulationData {
 public:
  // grid: Simulation::parseRuntime (Simulation.cpp:66-104)
  int bpdx = 0, bpdy = 0;
  double extent = 1;
  double extents[2] = {1, 1};
  int nsteps = 0;
  double endTime = 0;
  std::string poissonType;
  double CFL = 0.1, lambda = 1e4, dlm = 0, nu = 1e-2;
  double fadeLen = 0.01;                 // -fadeLen, fraction of the extent
  double fadeLenX = 0.01, fadeLenY = 0.01;  // fadeLen * extent (Simulation.cpp:97-98)
  int verbose = 1, iterativePenalization = 0, muteAll = 0;
  std::string precision = "fp64";  // -precision fp64|fp32
  int refineSweeps = 1;            // -refineSweeps: FP64 sweeps per FP32 Poisson solve
  std::string path2file = "./";  // -file: directory of the text logs
  // field dumps (SimulationData.h:62-67): every dumpFreq steps (-fdump) and
  // every dumpTime of simulated time (-tdump), into path4serialization
  // (-serialization)
  int dumpFreq = 0;
  double dumpTime = 0, nextDumpTime = 0;
  std::string path4serialization = "./";

  // state (SimulationData.h:41-45)
  double time = 0, dt = 0, uinfx = 0, uinfy = 0;
  int step = 0;

  std::vector<Shape> shapes;
  cup2d_ctx* ctx = nullptr;
  void* stream = nullptr;  // hipStream_t to enqueue on; nullptr: the library's own

  const Field vel{CUP2D_VEL, 2}, chi{CUP2D_CHI, 1}, pres{CUP2D_PRES, 1}, tmp{CUP2D_TMP, 1},
      tmpV{CUP2D_TMPV, 2}, uDef{CUP2D_UDEF, 2}, invRho{CUP2D_INVRHO, 1},
      vFluid{CUP2D_VFLUID, 2};

  SimulationData() = default;
  SimulationData(const SimulationData&) = delete;
  SimulationData& operator=(const SimulationData&) = delete;
  ~SimulationData();

  // SimulationData::allocateGrid (SimulationData.cpp:17-47): creates the
  // device context; the Poisson kind is fixed here.
  void allocateGrid();
  // Registers the parsed shapes with the device (Simulation::createShapes).
  void addShape(Shape s);
  // SimulationData::resetAll (SimulationData.cpp:99-109).
  void resetAll();
  bool bOver() const;        // SimulationData.cpp:172-177
  bool bDump() const;        // SimulationData.cpp:179-185
  void registerDump();       // SimulationData.cpp:111-114
  // SimulationData::dumpAll (SimulationData.cpp:207-239): "velChi_" (u, v,
  // chi), "pres_" and "invRho_" + name + the step in seven digits, as HDF5 +
  // XDMF (DumpHDF5.h)
  void dumpAll(const std::string& name);
  double maxRelSpeed() const;  // SimulationData.cpp:142-148
  double getH() const;         // h = extent / max(NX, NY)
  int NX() const { return bpdx * CUP2D_BS; }
  int NY() const { return bpdy * CUP2D_BS; }

  // Host <-> device state: the device owns time, dt, step and U_inf during a
  // step; these copy them across.
  void pullScalars();
  void pushScalars();
  void refreshShapes();

  // Field exchange in the reference's block layout (blockID order, components
  // interleaved) or dense (comp, NY, NX).
  void upload(const Field& f, const double* host, bool blocks = true);
  void download(const Field& f, double* host, bool blocks = true) const;
};

}  // namespace cup2d
