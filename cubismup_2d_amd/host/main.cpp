// main.cpp -- command-line driver with the reference's options (main.cpp,
// Simulation.cpp:66-173), e.g.
//   cup2d_run -bpdx 16 -bpdy 8 -CFL 0.1 -nu 0.00002 -poissonType freespace
//             -nsteps 499 -shapes "disk_radius=0.05_xpos=0.25_ypos=0.25_bForced=1_bFixed=1_xvel=0.2"
// Extra options: -deviceLoop 1 (whole steps on the device, dt computed
// there), -dumpFinal <path> (raw float64: vel (2,NY,NX) then pres (NY,NX)).
// Prints one final JSON line with time, step and each body's forces.
#include <cstdio>
#include <vector>

#include "Simulation.h"

int main(int argc, char** argv) {
  try {
    cup2d::Simulation s(argc, argv);
    s.init();
    s.simulate();
    cup2d::ArgumentParser p(argc, argv);
    const std::string dump = p.asString("dumpFinal", "");
    if (!dump.empty()) {
      const size_t n = (size_t)s.sim.NX() * s.sim.NY();
      std::vector<double> buf(3 * n);
      s.sim.download(s.sim.vel, buf.data(), false);
      s.sim.download(s.sim.pres, buf.data() + 2 * n, false);
      FILE* f = std::fopen(dump.c_str(), "wb");
      if (!f || std::fwrite(buf.data(), sizeof(double), buf.size(), f) != buf.size()) {
        std::fprintf(stderr, "cannot write %s\n", dump.c_str());
        return 1;
      }
      std::fclose(f);
    }
    std::printf("{\"time\": %.17g, \"step\": %d, \"dt\": %.17g, \"shapes\": [", s.sim.time,
                s.sim.step, s.sim.dt);
    for (size_t i = 0; i < s.sim.shapes.size(); ++i) {
      const cup2d_shape_state& st = s.sim.shapes[i].state;
      std::printf("%s{\"drag\": %.17g, \"thrust\": %.17g, \"forcex\": %.17g, \"forcey\": %.17g, "
                  "\"u\": %.17g, \"v\": %.17g, \"omega\": %.17g, \"x\": %.17g, \"y\": %.17g}",
                  i ? ", " : "", st.drag, st.thrust, st.forcex, st.forcey, st.u, st.v, st.omega,
                  st.center[0], st.center[1]);
    }
    std::printf("]}\n");
    return 0;
  } catch (const cup2d::Error& e) {
    std::fprintf(stderr, "cup2d_run: %s (status %d)\n", e.what(), e.status);
    return 2;
  }
}
