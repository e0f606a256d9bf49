// Simulation.h -- the reference's Simulation driver (Simulation.h:16-40,
// Simulation.cpp:49-305) over the device operators.
#pragma once

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "Operator.h"

namespace cup2d {

// Command-line options in the reference's "-key value" form
// (cubism::ArgumentParser as used in Simulation.cpp:66-104).
class ArgumentParser {
  std::map<std::string, std::string> kv;

 public:
  ArgumentParser(int argc, char** argv);
  explicit ArgumentParser(const std::map<std::string, std::string>& m) : kv(m) {}
  bool has(const std::string& key) const { return kv.count(key) != 0; }
  std::string asString(const std::string& key, const std::string& def) const;
  double asDouble(const std::string& key, double def) const;
  int asInt(const std::string& key, int def) const;
  bool asBool(const std::string& key, bool def) const;
  // strict mode: throws cup2d::Error when the key is missing
  double required(const std::string& key) const;
};

// Parses one shape line such as "disk radius=0.05 xpos=0.25 bForced=1" (the
// '_'-separated -shapes syntax after Simulation.cpp:112's replacement) into a
// Shape with the reference's defaults (Shape.cpp:331-345, ShapesSimple.h:19-21).
Shape parseShape(const std::string& line, const double extents[2]);

class Simulation {
 public:
  SimulationData sim;

 protected:
  ArgumentParser parser;
  std::vector<std::unique_ptr<Operator>> pipeline;
  // Run whole steps with dt computed on the device (no host round trip per
  // operator); "-deviceLoop 1".  Results are identical to the operator path.
  bool deviceLoop = false;

  void createShapes();
  void parseRuntime();

 public:
  Simulation(int argc, char** argv);
  explicit Simulation(const ArgumentParser& p);

  void reset();
  void init();
  void simulate();
  double calcMaxTimestep();
  bool advance(const double DT);

  const std::vector<Shape>& getShapes() { return sim.shapes; }
  const std::vector<std::unique_ptr<Operator>>& getPipeline() const { return pipeline; }
};

}  // namespace cup2d
