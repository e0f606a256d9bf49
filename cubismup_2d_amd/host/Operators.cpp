// Operators.cpp -- the pipeline operators and Poisson solvers over the C-ABI.
#include "Operator.h"

namespace cup2d {

// ---------------------------------------------------------------- Poisson
std::unique_ptr<PoissonSolver> PoissonSolver::makeSolver(SimulationData& sim) {
  const std::string& t = sim.poissonType;
  if (t == "periodic" || t == "cpu_periodic") return std::make_unique<HIP_periodic>(sim);
  if (t == "freespace" || t == "cpu_freespace") return std::make_unique<HIP_freespace>(sim);
  throw Error(CUP2D_ERR_ARG, "PoissonSolver::makeSolver: no device solver for '" + t + "'");
}

static void solve_device(SimulationData& sim, const Field& src, const Field& dst) {
  if (src.id != CUP2D_TMP || dst.id != CUP2D_PRES)
    throw Error(CUP2D_ERR_ARG, "PoissonSolver::solve: source must be tmp and destination pres");
  check(sim.ctx, cup2d_poisson_solve(sim.ctx), "cup2d_poisson_solve");
}

void HIP_periodic::solve(const Field& BSRC, const Field& BDST) {
  solve_device(sim, BSRC, BDST);
  iter = 1;
}

void HIP_freespace::solve(const Field& BSRC, const Field& BDST) {
  solve_device(sim, BSRC, BDST);
  iter = 1;
}

// ---------------------------------------------------------------- helpers
double findMaxU::run() const {
  double m = 0;
  check(sim.ctx, cup2d_find_max_u(sim.ctx, &m), "cup2d_find_max_u");
  return m;
}

// ---------------------------------------------------------------- pipeline
void PutObjectsOnGrid::operator()(const double dt) {
  check(sim.ctx, cup2d_put_objects(sim.ctx, dt), "PutObjectsOnGrid");
}

void advDiff::operator()(const double dt) {
  check(sim.ctx, cup2d_advdiff(sim.ctx, dt), "advDiff");
}

PressureSingle::PressureSingle(SimulationData& s)
    : Operator(s), pressureSolver(PoissonSolver::makeSolver(s)) {}

void PressureSingle::updatePressureRHS(const double dt) const {
  check(sim.ctx, cup2d_pressure_rhs(sim.ctx, dt), "PressureSingle::updatePressureRHS");
}

void PressureSingle::pressureCorrection(const double dt) const {
  check(sim.ctx, cup2d_pressure_correction(sim.ctx, dt), "PressureSingle::pressureCorrection");
}

void PressureSingle::integrateMomenta(const double dt) const {
  check(sim.ctx, cup2d_body_update(sim.ctx, dt), "PressureSingle::integrateMomenta");
}

void PressureSingle::penalize(const double dt) const {
  check(sim.ctx, cup2d_penalize(sim.ctx, dt), "PressureSingle::penalize");
}

// PressureSingle::operator() (PressureSingle.cpp:346-373): RHS, solve,
// correction, body momenta + velocities + collisions, penalization.
void PressureSingle::operator()(const double dt) {
  updatePressureRHS(dt);
  pressureSolver->solve(sim.tmp, sim.pres);
  pressureCorrection(dt);
  integrateMomenta(dt);
  penalize(dt);
}

void PressureIterator_unif::operator()(const double dt) {
  check(sim.ctx, cup2d_pressure_iterator(sim.ctx, dt, &iterations, &relDF),
        "PressureIterator_unif");
}

void ComputeForces::operator()(const double) {
  check(sim.ctx, cup2d_compute_forces(sim.ctx), "ComputeForces");
}

}  // namespace cup2d
