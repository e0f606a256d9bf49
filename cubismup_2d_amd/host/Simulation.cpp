// Simulation.cpp -- the reference's driver loop (Simulation.cpp:49-305) over
// the device operators.
#include "Simulation.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <sstream>

namespace cup2d {

// ---------------------------------------------------------- ArgumentParser
static std::string strip_dash(const std::string& k) {
  size_t i = 0;
  while (i < k.size() && k[i] == '-') ++i;
  return k.substr(i);
}

ArgumentParser::ArgumentParser(int argc, char** argv) {
  for (int i = 1; i < argc; ++i) {
    std::string k = argv[i];
    if (k.size() < 2 || k[0] != '-') continue;
    // a value may itself start with '-' (negative numbers); any following
    // token that is not a "-name" key is the value
    std::string v = "true";
    if (i + 1 < argc) {
      const std::string nx = argv[i + 1];
      const bool is_key = nx.size() > 1 && nx[0] == '-' && std::isalpha((unsigned char)nx[1]);
      if (!is_key) {
        v = nx;
        ++i;
      }
    }
    kv[strip_dash(k)] = v;
  }
}

std::string ArgumentParser::asString(const std::string& key, const std::string& def) const {
  auto it = kv.find(strip_dash(key));
  return it == kv.end() ? def : it->second;
}

double ArgumentParser::asDouble(const std::string& key, double def) const {
  auto it = kv.find(strip_dash(key));
  if (it == kv.end()) return def;
  char* end = nullptr;
  const double v = std::strtod(it->second.c_str(), &end);
  if (end == it->second.c_str()) throw Error(CUP2D_ERR_ARG, "bad number for -" + key);
  return v;
}

int ArgumentParser::asInt(const std::string& key, int def) const {
  auto it = kv.find(strip_dash(key));
  if (it == kv.end()) return def;
  return (int)asDouble(key, def);
}

bool ArgumentParser::asBool(const std::string& key, bool def) const {
  auto it = kv.find(strip_dash(key));
  if (it == kv.end()) return def;
  const std::string& v = it->second;
  if (v == "true" || v == "True" || v == "1") return true;
  if (v == "false" || v == "False" || v == "0") return false;
  return asDouble(key, 0) != 0;
}

double ArgumentParser::required(const std::string& key) const {
  if (!has(strip_dash(key))) throw Error(CUP2D_ERR_ARG, "missing required option -" + strip_dash(key));
  return asDouble(key, 0);
}

// ---------------------------------------------------------- shapes
Shape parseShape(const std::string& line_in, const double extents[2]) {
  std::string line = line_in;
  std::replace(line.begin(), line.end(), '_', ' ');
  std::istringstream ls(line);
  std::string name;
  ls >> name;
  std::map<std::string, std::string> kv;
  std::string tok;
  while (ls >> tok) {
    const size_t eq = tok.find('=');
    if (eq == std::string::npos) continue;
    kv[strip_dash(tok.substr(0, eq))] = tok.substr(eq + 1);
  }
  ArgumentParser p(kv);
  if (name != "disk" && name != "carlingfish")
    throw Error(CUP2D_ERR_ARG,
                "shape '" + name + "' is not supported on the device (disk, carlingfish)");
  Shape s;
  s.name = name;
  cup2d_shape_desc& d = s.desc;
  d.kind = name == "disk" ? 0 : 1;
  d.xpos = p.asDouble("xpos", 0.5 * extents[0]);
  d.ypos = p.asDouble("ypos", 0.5 * extents[1]);
  d.angle_deg = p.asDouble("angle", 0);
  d.rhoS = p.asDouble("rhoS", 1);
  const bool bFixed = p.asBool("bFixed", false), bForced = p.asBool("bForced", false);
  d.bFixedx = p.asBool("bFixedx", bFixed);
  d.bFixedy = p.asBool("bFixedy", bFixed);
  d.bForcedx = p.asBool("bForcedx", bForced);
  d.bForcedy = p.asBool("bForcedy", bForced);
  d.bBlockAng = p.asBool("bBlockAng", d.bForcedx || d.bForcedy);
  d.xvel = p.asDouble("xvel", 0);
  d.yvel = p.asDouble("yvel", 0);
  d.angvel = p.asDouble("angvel", 0);
  d.timeForced = p.asDouble("timeForced", std::numeric_limits<double>::max());
  d.radius = p.asDouble("radius", 0.1);
  d.tAccel = p.asDouble("tAccel", -1);
  if (d.kind == 1) {
    // Fish::Fish and CarlingFish::CarlingFish (Fish.h:27-29, CarlingFish.cpp:90-95)
    d.length = p.asDouble("L", 0.1);
    d.Tperiod = p.asDouble("T", 1);
    d.phaseShift = p.asDouble("phi", 0);
    d.amplitudeFactor = p.asDouble("amplitudeFactor", 1.0);
    // midline envelope and half-width law: the reference's CarlingFish by
    // default; envelope=quadratic gives L (envA0 + envA2 (s/L)^2) with
    // wavelength waveLength*L, profile=naca a NACA 4-digit half-thickness
    const std::string env = p.asString("envelope", "carling");
    const std::string prof = p.asString("profile", "carling");
    if (env != "carling" && env != "quadratic")
      throw Error(CUP2D_ERR_ARG, "envelope must be carling or quadratic");
    if (prof != "carling" && prof != "naca")
      throw Error(CUP2D_ERR_ARG, "profile must be carling or naca");
    d.envelope = env == "carling" ? 0 : 1;
    d.env_a0 = p.asDouble("envA0", 0.02);
    d.env_a2 = p.asDouble("envA2", 0.08);
    d.waveLength = p.asDouble("waveLength", 1.0);
    d.profile = prof == "carling" ? 0 : 1;
    d.naca_t = p.asDouble("nacaT", 0.08);
  }
  return s;
}

// the -shapes argument: descriptors separated by newlines or commas
// (Simulation.cpp:106-173)
static std::vector<Shape> parseShapes(const std::string& arg, const double extents[2]) {
  std::vector<Shape> out;
  std::stringstream descriptors(arg);
  std::string lines;
  while (std::getline(descriptors, lines)) {
    std::stringstream ls(lines);
    std::string line;
    while (std::getline(ls, line, ',')) {
      std::string t = line;
      std::replace(t.begin(), t.end(), '_', ' ');
      std::istringstream probe(t);
      std::string objectName;
      probe >> objectName;
      if (objectName.empty() || objectName[0] == '#') continue;
      out.push_back(parseShape(line, extents));
    }
  }
  if (out.empty()) throw Error(CUP2D_ERR_ARG, "FATAL - Did not create any obstacles.");
  return out;
}

// ---------------------------------------------------------- Simulation
Simulation::Simulation(int argc, char** argv) : parser(argc, argv) {}
Simulation::Simulation(const ArgumentParser& p) : parser(p) {}

void Simulation::parseRuntime() {
  sim.bpdx = (int)parser.required("bpdx");
  sim.bpdy = (int)parser.required("bpdy");
  sim.extent = parser.asDouble("extent", 1);
  sim.nsteps = parser.asInt("nsteps", 0);
  sim.endTime = parser.asDouble("tend", 0);
  sim.poissonType = parser.asString("poissonType", "");
  sim.CFL = parser.asDouble("CFL", .1);
  sim.lambda = parser.asDouble("lambda", 1e3 / sim.CFL);
  sim.dlm = parser.asDouble("dlm", 0);
  sim.nu = parser.asDouble("nu", 1e-2);
  sim.fadeLen = parser.asDouble("fadeLen", 0.01);
  sim.verbose = parser.asInt("verbose", 1);
  sim.iterativePenalization = parser.asInt("iterativePenalization", 0);
  sim.precision = parser.asString("precision", "fp64");
  sim.dumpFreq = parser.asInt("fdump", 0);  // Simulation.cpp:84-85
  sim.dumpTime = parser.asDouble("tdump", 0);
  sim.path4serialization = parser.asString("serialization", "./");
  sim.refineSweeps = parser.asInt("refineSweeps", 1);
  sim.muteAll = parser.asInt("muteAll", 0);
  sim.path2file = parser.asString("file", "./");
  if (sim.muteAll) sim.verbose = 0;
  deviceLoop = parser.asBool("deviceLoop", false);
  if (sim.dlm > 0) throw Error(CUP2D_ERR_ARG, "-dlm > 0 has no device implementation");
  {  // shape descriptors are checked before any device allocation
    const double mx = std::max(sim.bpdx, sim.bpdy);
    const double ext[2] = {sim.extent * sim.bpdx / mx, sim.extent * sim.bpdy / mx};
    (void)parseShapes(parser.asString("shapes", ""), ext);
  }
  sim.allocateGrid();
}

void Simulation::createShapes() {
  for (const Shape& s : parseShapes(parser.asString("shapes", ""), sim.extents)) sim.addShape(s);
}

void Simulation::init() {
  parseRuntime();
  createShapes();
  pipeline.clear();
  pipeline.push_back(std::make_unique<PutObjectsOnGrid>(sim));
  pipeline.push_back(std::make_unique<advDiff>(sim));
  if (sim.iterativePenalization)
    pipeline.push_back(std::make_unique<PressureIterator_unif>(sim));
  else
    pipeline.push_back(std::make_unique<PressureSingle>(sim));
  pipeline.push_back(std::make_unique<ComputeForces>(sim));
  if (sim.verbose) {
    std::printf("Operator ordering:\n");
    for (auto& op : pipeline) std::printf("\t%s\n", op->getName().c_str());
  }
  reset();
  sim.dt = 0;
}

// Simulation::reset (Simulation.cpp:232-241): resetAll, IC, PutObjectsOnGrid(0)
// and ApplyObjVel, performed on the device by cup2d_reset.
void Simulation::reset() {
  sim.resetAll();
  check(sim.ctx, cup2d_reset(sim.ctx), "Simulation::reset");
  sim.pullScalars();
  sim.refreshShapes();
  if (sim.dumpFreq > 0 || sim.dumpTime > 0) sim.dumpAll("IC");  // Simulation.cpp:229
}

double Simulation::calcMaxTimestep() {
  check(sim.ctx, cup2d_calc_max_timestep(sim.ctx, &sim.dt), "calcMaxTimestep");
  if (sim.verbose) {
    double maxU = 0;
    sim.pullScalars();
    check(sim.ctx, cup2d_find_max_u(sim.ctx, &maxU), "findMaxU");
    std::printf("step:%d, time:%f, dt=%f, uinf:[%f %f], maxU:%f\n", sim.step, sim.time, sim.dt,
                sim.uinfx, sim.uinfy, maxU);
  }
  return sim.dt;
}

bool Simulation::advance(const double dt) {
  if (!(dt > 2.2e-16)) throw Error(CUP2D_ERR_ARG, "advance: dt must be > 2.2e-16");
  const bool bDump = sim.bDump();  // Simulation.cpp:282
  sim.dt = dt;
  for (auto& op : pipeline) (*op)(sim.dt);
  sim.pullScalars();  // U_inf was set on the device by PutObjectsOnGrid
  sim.time += sim.dt;
  sim.step++;
  sim.pushScalars();
  sim.refreshShapes();
  if (bDump) {  // Simulation.cpp:292-297
    sim.registerDump();
    sim.dumpAll("avemaria_");
  }
  return sim.bOver();
}

void Simulation::simulate() {
  // the reference loops until bOver(); with neither -nsteps nor -tend that
  // never happens, which is refused here
  if (sim.nsteps <= 0 && sim.endTime <= 0)
    throw Error(CUP2D_ERR_ARG, "simulate: set -nsteps or -tend");
  if (!deviceLoop) {
    while (true) {
      const double dt = calcMaxTimestep();
      if (advance(dt)) break;
    }
    return;
  }
  // whole steps enqueued back to back with dt computed on the device; the
  // stop condition is checked every `chunk` steps (the step count is exact,
  // -tend may overshoot by up to chunk-1 steps); the force logs get one line
  // per run of steps instead of one per step
  const int chunk = 50;
  while (!sim.bOver()) {
    int n = chunk;
    if (sim.nsteps > 0) n = std::min(n, sim.nsteps + 1 - sim.step);
    check(sim.ctx, cup2d_simulate(sim.ctx, n), "cup2d_simulate");
    check(sim.ctx, cup2d_synchronize(sim.ctx), "cup2d_synchronize");
    sim.pullScalars();
    sim.refreshShapes();
    // forces of the last step of the run, computed at its start time
    if (!sim.muteAll) ComputeForces::writeLogs(sim, sim.time - sim.dt);
    if (sim.verbose)
      std::printf("step:%d, time:%f, dt=%f, uinf:[%f %f]\n", sim.step, sim.time, sim.dt,
                  sim.uinfx, sim.uinfy);
  }
}

}  // namespace cup2d
