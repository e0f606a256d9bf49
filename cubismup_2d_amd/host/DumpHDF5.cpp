// DumpHDF5.cpp -- HDF5 + XDMF field dumps, the files SimulationData::dumpAll
// writes (SimulationData.cpp:49-97,207-239: "velChi_" u, v, chi glued as one
// three-channel vector field, StreamerGlue in Definitions.h:218-239;
// "pres_" and "invRho_" scalars; float data; the step number zero-padded to
// seven digits).  The reference writes them through Cubism's DumpHDF5, which
// is not part of its tree; this writer keeps its shape: one dataset "data"
// of float32 [NY][NX][channels] per file, the time and step as attributes,
// and an XDMF description of the cell-centred field on the uniform mesh
// (origin 0, spacing h) that ParaView and VisIt read.
#include "DumpHDF5.h"

#include <hdf5.h>

#include <cstdio>
#include <vector>

namespace cup2d {

int write_h5_xmf(const std::string& dir, const std::string& stem, double time, int step, int NX,
                 int NY, double h, int nch, const double* const* channels,
                 const char* attribute_type) {
  if (NX <= 0 || NY <= 0 || nch <= 0 || !channels) return -1;
  const std::string base = (dir.empty() ? std::string(".") : dir) + "/" + stem;
  const size_t n = (size_t)NX * NY;
  std::vector<float> buf(n * nch);
  for (size_t i = 0; i < n; ++i)
    for (int c = 0; c < nch; ++c) buf[i * nch + c] = (float)channels[c][i];

  const hid_t file = H5Fcreate((base + ".h5").c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (file < 0) return -2;
  const hsize_t dims[3] = {(hsize_t)NY, (hsize_t)NX, (hsize_t)nch};
  const hid_t space = H5Screate_simple(3, dims, nullptr);
  const hid_t dset = H5Dcreate2(file, "data", H5T_NATIVE_FLOAT, space, H5P_DEFAULT, H5P_DEFAULT,
                                H5P_DEFAULT);
  herr_t st = dset >= 0 ? H5Dwrite(dset, H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT,
                                   buf.data())
                        : -1;
  const hid_t aspace = H5Screate(H5S_SCALAR);
  if (st >= 0) {
    const hid_t at = H5Acreate2(dset, "time", H5T_NATIVE_DOUBLE, aspace, H5P_DEFAULT, H5P_DEFAULT);
    st = at >= 0 ? H5Awrite(at, H5T_NATIVE_DOUBLE, &time) : -1;
    if (at >= 0) H5Aclose(at);
  }
  if (st >= 0) {
    const hid_t as = H5Acreate2(dset, "step", H5T_NATIVE_INT, aspace, H5P_DEFAULT, H5P_DEFAULT);
    st = as >= 0 ? H5Awrite(as, H5T_NATIVE_INT, &step) : -1;
    if (as >= 0) H5Aclose(as);
  }
  H5Sclose(aspace);
  if (dset >= 0) H5Dclose(dset);
  H5Sclose(space);
  H5Fclose(file);
  if (st < 0) return -3;

  FILE* x = std::fopen((base + ".xmf").c_str(), "w");
  if (!x) return -4;
  std::fprintf(x, "<?xml version=\"1.0\" ?>\n<!DOCTYPE Xdmf SYSTEM \"Xdmf.dtd\" []>\n");
  std::fprintf(x, "<Xdmf Version=\"2.0\">\n <Domain>\n  <Grid GridType=\"Uniform\">\n");
  std::fprintf(x, "   <Time Value=\"%.17e\"/>\n", time);
  std::fprintf(x, "   <Topology TopologyType=\"2DCoRectMesh\" Dimensions=\"%d %d\"/>\n", NY + 1,
               NX + 1);
  std::fprintf(x, "   <Geometry GeometryType=\"ORIGIN_DXDY\">\n");
  std::fprintf(x, "    <DataItem Dimensions=\"2\" NumberType=\"Float\" Precision=\"8\" "
                  "Format=\"XML\">0 0</DataItem>\n");
  std::fprintf(x, "    <DataItem Dimensions=\"2\" NumberType=\"Float\" Precision=\"8\" "
                  "Format=\"XML\">%.17e %.17e</DataItem>\n", h, h);
  std::fprintf(x, "   </Geometry>\n");
  std::fprintf(x, "   <Attribute Name=\"data\" AttributeType=\"%s\" Center=\"Cell\">\n",
               attribute_type);
  std::fprintf(x, "    <DataItem Dimensions=\"%d %d %d\" NumberType=\"Float\" Precision=\"4\" "
                  "Format=\"HDF\">%s.h5:/data</DataItem>\n", NY, NX, nch, stem.c_str());
  std::fprintf(x, "   </Attribute>\n  </Grid>\n </Domain>\n</Xdmf>\n");
  std::fclose(x);
  return 0;
}

}  // namespace cup2d

extern "C" int cup2d_io_write_h5(const char* dir, const char* stem, double time, int step, int NX,
                                 int NY, double h, int nch, const double* const* channels,
                                 const char* attribute_type) {
  if (!dir || !stem || !attribute_type) return -1;
  return cup2d::write_h5_xmf(dir, stem, time, step, NX, NY, h, nch, channels, attribute_type);
}
