// DumpHDF5.h -- HDF5 + XDMF field dumps (SimulationData::dumpAll,
// SimulationData.cpp:49-97,207-239).  Built into libcup2d_io.so, which links
// the HDF5 library; the device library does not depend on it.
#pragma once

#include <string>

namespace cup2d {
// Writes <dir>/<stem>.h5 (dataset "data", float32 [NY][NX][nch], attributes
// "time" and "step") and <dir>/<stem>.xmf (the cell-centred attribute on the
// uniform mesh of spacing h).  channels[c] are dense (NY, NX) planes.
// attribute_type: "Scalar" or "Vector".  Returns 0 or a negative code.
int write_h5_xmf(const std::string& dir, const std::string& stem, double time, int step, int NX,
                 int NY, double h, int nch, const double* const* channels,
                 const char* attribute_type);
}  // namespace cup2d

extern "C" int cup2d_io_write_h5(const char* dir, const char* stem, double time, int step, int NX,
                                 int NY, double h, int nch, const double* const* channels,
                                 const char* attribute_type);
