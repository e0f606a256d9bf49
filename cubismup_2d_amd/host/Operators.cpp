// Operators.cpp -- the pipeline operators and Poisson solvers over the C-ABI.
#include "Operator.h"

#include <fstream>

namespace cup2d {

// ---------------------------------------------------------------- Poisson
std::unique_ptr<PoissonSolver> PoissonSolver::makeSolver(SimulationData& sim) {
  const std::string& t = sim.poissonType;
  if (t == "periodic" || t == "cpu_periodic") return std::make_unique<HIP_periodic>(sim);
  if (t == "freespace" || t == "cpu_freespace") return std::make_unique<HIP_freespace>(sim);
  if (t == "" || t == "cosine" || t == "cpu_cosine") return std::make_unique<HIP_cosine>(sim);
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

void HIP_cosine::solve(const Field& BSRC, const Field& BDST) {
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
  if (!sim.muteAll) {
    sim.refreshShapes();
    writeLogs(sim);
  }
}

void ComputeForces::writeLogs(SimulationData& sim) { writeLogs(sim, sim.time); }

void ComputeForces::writeLogs(SimulationData& sim, double time) {
  for (const Shape& sh : sim.shapes) {
    const cup2d_shape_state& t = sh.state;
    const std::string id = std::to_string(sh.obstacleID);
    std::ofstream f(sim.path2file + "/forceValues_" + id + ".dat", std::ios::app);
    std::ofstream p(sim.path2file + "/powerValues_" + id + ".dat", std::ios::app);
    if (!f || !p) throw Error(CUP2D_ERR_ARG, "cannot write the force logs in " + sim.path2file);
    f.seekp(0, std::ios::end);
    if (sim.step == 0 || f.tellp() == 0) {
      f << "time Fx Fy FxPres FyPres FxVisc FyVisc tau tauPres tauVisc"
           " drag thrust perimeter circulation area_penal mass_penal"
           " forcex_penal forcey_penal torque_penal\n";
      p << "time Pthrust Pdrag PoutBnd Pout defPowerBnd defPower"
           " EffPDefBnd EffPDef\n";
    }
    f << time << " " << t.forcex << " " << t.forcey << " " << t.forcex_P << " " << t.forcey_P
      << " " << t.forcex_V << " " << t.forcey_V << " " << t.torque << " " << t.torque_P << " "
      << t.torque_V << " " << t.drag << " " << t.thrust << " " << t.perimeter << " "
      << t.circulation << "\n";
    p << time << " " << t.Pthrust << " " << t.Pdrag << " " << t.PoutBnd << " " << t.Pout
      << " " << t.defPowerBnd << " " << t.defPower << " " << t.EffPDefBnd << " " << t.EffPDef
      << "\n";
  }
}

}  // namespace cup2d
