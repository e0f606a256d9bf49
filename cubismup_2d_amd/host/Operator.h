// Operator.h -- the reference's operator interface (Operator.h:14-27) over
// the device context, and the per-timestep operators of the constant-density
// pipeline (Simulation.cpp:196-221).
//
// Each operator() enqueues the same work as the reference operator on the
// context's stream and waits for it, so host code that calls them one by one
// sees the reference's sequential semantics.  Device-side errors (a body
// leaving the domain, a full obstacle pool) surface as cup2d::Error.
#pragma once

#include <memory>
#include <string>

#include "SimulationData.h"

namespace cup2d {

class Operator {
 protected:
  SimulationData& sim;

 public:
  explicit Operator(SimulationData& s) : sim(s) {}
  virtual ~Operator() {}
  virtual void operator()(const double dt) = 0;
  virtual std::string getName() = 0;
};

// PoissonSolver (PoissonSolver.h:14-41): solves lap(p) = rhs with the RHS in
// sim.tmp and the solution written to sim.pres.  solve() includes cub2rhs's
// boundary fade (PoissonSolver.cpp:61-141) and sol2cub; both stay on the
// device.  The source is modified in place by the fade, as in the reference.
class PoissonSolver {
 protected:
  SimulationData& sim;

 public:
  const size_t totNy, totNx;
  int iter = 0;
  explicit PoissonSolver(SimulationData& s)
      : sim(s), totNy((size_t)s.NY()), totNx((size_t)s.NX()) {}
  virtual ~PoissonSolver() {}
  // BSRC = sim.tmp, BDST = sim.pres (PressureSingle.cpp:356).
  virtual void solve(const Field& BSRC, const Field& BDST) = 0;
  virtual std::string getName() = 0;
  // PoissonSolver::makeSolver (PoissonSolver.cpp:23-59) for the poissonType
  // values with an FFT solver: "periodic", "freespace", "cosine" or "" (the
  // default), and the CUDA names "cpu_periodic", "cpu_freespace".  Other
  // types throw cup2d::Error.
  static std::unique_ptr<PoissonSolver> makeSolver(SimulationData& sim);
};

// FFTW_periodic (FFTW_periodic.h:94-166) on the device.
class HIP_periodic : public PoissonSolver {
 public:
  using PoissonSolver::PoissonSolver;
  void solve(const Field& BSRC, const Field& BDST) override;
  std::string getName() override { return "HIP_periodic"; }
};

// FFTW_dirichlet (FFTW_dirichlet.h:23-148), the reference's default solver
// (poissonType "" or "cosine"), on the device.
class HIP_cosine : public PoissonSolver {
 public:
  using PoissonSolver::PoissonSolver;
  void solve(const Field& BSRC, const Field& BDST) override;
  std::string getName() override { return "HIP_cosine"; }
};

// FFTW_freespace (FFTW_freespace.cpp:14-138) on the device.
class HIP_freespace : public PoissonSolver {
 public:
  using PoissonSolver::PoissonSolver;
  void solve(const Field& BSRC, const Field& BDST) override;
  std::string getName() override { return "HIP_freespace"; }
};

// findMaxU (Helpers.cpp:136-176).
class findMaxU {
  SimulationData& sim;

 public:
  explicit findMaxU(SimulationData& s) : sim(s) {}
  double run() const;
};

// PutObjectsOnGrid (PutObjectsOnGrid.cpp:186-241).
class PutObjectsOnGrid : public Operator {
 public:
  using Operator::Operator;
  void operator()(const double dt) override;
  std::string getName() override { return "PutObjectsOnGrid"; }
};

// advDiff (advDiff.cpp:36-118).
class advDiff : public Operator {
 public:
  using Operator::Operator;
  void operator()(const double dt) override;
  std::string getName() override { return "advDiff"; }
};

// PressureSingle (PressureSingle.cpp:346-380): owns its Poisson solver.
class PressureSingle : public Operator {
  std::unique_ptr<PoissonSolver> pressureSolver;

  void updatePressureRHS(const double dt) const;   // PressureSingle.cpp:110-214
  void pressureCorrection(const double dt) const;  // PressureSingle.cpp:216-241
  void integrateMomenta(const double dt) const;    // :23-67 + Shape::updateVelocity + :243-344
  void penalize(const double dt) const;            // PressureSingle.cpp:69-108

 public:
  explicit PressureSingle(SimulationData& s);
  void operator()(const double dt) override;
  std::string getName() override { return "PressureSingle"; }
};

// PressureIterator_unif (PressureIterator_unif.cpp:216-289), selected by
// -iterativePenalization 1 (Simulation.cpp:215-216).  The Poisson solver is
// the context's FFT solver; iterations and relDF of the last call are kept.
class PressureIterator_unif : public Operator {
 public:
  int iterations = 0;
  double relDF = 0;
  using Operator::Operator;
  void operator()(const double dt) override;
  std::string getName() override { return "PressureIterator_unif"; }
};

// ComputeForces (ComputeForces.cpp:16-111).  Unless -muteAll, appends each
// body's forces and powers to forceValues_<id>.dat and powerValues_<id>.dat
// in -file's directory, with the reference's headers and columns
// (Shape.cpp:305-328).
class ComputeForces : public Operator {
 public:
  using Operator::Operator;
  // log lines of the bodies' current force state, stamped with sim.time or
  // with the given time (the device-loop driver logs after a run of steps)
  static void writeLogs(SimulationData& sim);
  static void writeLogs(SimulationData& sim, double time);
  void operator()(const double dt) override;
  std::string getName() override { return "ComputeForces"; }
};

}  // namespace cup2d
