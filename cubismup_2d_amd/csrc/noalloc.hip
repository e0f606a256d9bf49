// noalloc.hip -- findMaxU over caller-owned device planes with no allocation.
//
// cup2d_find_max_u reduces the context's velocity through a partials buffer
// that cup2d_create allocates.  cup2d_find_max_u_noalloc does the same
// reduction (findMaxU::run, Helpers.cpp:136-176: the max over cells of
// |u+Uinf|, |v+Vinf|, |u|, |v|, NaN kept) for callers that cannot allocate
// device memory (code under stream capture, tight serving loops): the
// per-workgroup partial maxima live in the file-scope array `scratch` below,
// whose device address is looked up once per device.  Nothing here calls
// hipMalloc or hipFree.
//
// Every call on one device shares that device's `scratch`; two calls that
// overlap on one device (different streams, no ordering between them) race
// on it.  Calls ordered on one stream are safe: each call's final kernel
// reads only the partials its own first kernel wrote.
#include <mutex>

#include "device_util.hpp"

namespace cup2d {

void set_ctxfree_error(const char* msg);

// partial maxima: one per workgroup of the first kernel
constexpr int NOALLOC_MAX_PARTS = 4096;
constexpr int NOALLOC_NT = 256;
// the largest plane accepted (cells): 2^28, e.g. 16384 x 16384
constexpr long long NOALLOC_MAX_CELLS = 1LL << 28;

static __device__ __align__(256) double scratch[NOALLOC_MAX_PARTS];

constexpr int NOALLOC_MAX_DEVICES = 64;
static double* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};
static std::mutex g_scratch_mu;

// Device address of `scratch` on the current device (looked up once per
// device with that device current); nullptr on failure.
static double* scratch_ptr(int* dev_out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_MAX_DEVICES) return nullptr;
  *dev_out = dev;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  return g_scratch_ptr[dev];
}

__global__ __launch_bounds__(NOALLOC_NT) void k_max_u_parts(const double* __restrict__ u,
                                                             const double* __restrict__ v,
                                                             long long n, double UINF,
                                                             double VINF,
                                                             double* __restrict__ part) {
  __shared__ double lds[NOALLOC_NT / 64];
  double m = 0;
  const long long stride = (long long)gridDim.x * NOALLOC_NT;
  for (long long i = (long long)blockIdx.x * NOALLOC_NT + threadIdx.x; i < n; i += stride) {
    const double a = u[i], b = v[i];
    m = nmax(m, fabs(a + UINF));
    m = nmax(m, fabs(b + VINF));
    m = nmax(m, fabs(a));
    m = nmax(m, fabs(b));
  }
  m = block_max<NOALLOC_NT>(m, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = m;
}

__global__ __launch_bounds__(NOALLOC_NT) void k_max_u_final(const double* __restrict__ part,
                                                             int nparts,
                                                             double* __restrict__ out) {
  __shared__ double lds[NOALLOC_NT / 64];
  const double m = block_max<NOALLOC_NT>(max_array<NOALLOC_NT>(part, nparts), lds);
  if (threadIdx.x == 0) *out = m;
}

}  // namespace cup2d

using namespace cup2d;

extern "C" int cup2d_find_max_u_noalloc(const double* u, const double* v, long long ncells,
                                        double uinfx, double uinfy, double* maxU,
                                        void* stream) {
  if (!u || !v || !maxU || ncells <= 0) {
    set_ctxfree_error("find_max_u_noalloc: null plane or output, or ncells <= 0");
    return CUP2D_ERR_ARG;
  }
  if (ncells > NOALLOC_MAX_CELLS) {
    set_ctxfree_error("find_max_u_noalloc: ncells exceeds the maximum of 2^28 cells");
    return CUP2D_ERR_ARG;
  }
  int dev = -1;
  double* part = scratch_ptr(&dev);
  if (!part) {
    set_ctxfree_error("find_max_u_noalloc: no device address for the scratch array");
    return CUP2D_ERR_HIP;
  }
  long long nwg = (ncells + 8LL * NOALLOC_NT - 1) / (8LL * NOALLOC_NT);  // ~8 cells per thread
  if (nwg > NOALLOC_MAX_PARTS) nwg = NOALLOC_MAX_PARTS;
  if (nwg < 1) nwg = 1;
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_max_u_parts, dim3((unsigned)nwg), dim3(NOALLOC_NT), 0, s, u, v, ncells,
                     uinfx, uinfy, part);
  if (hipGetLastError() != hipSuccess) {
    set_ctxfree_error("find_max_u_noalloc: launch failed");
    return CUP2D_ERR_HIP;
  }
  hipLaunchKernelGGL(k_max_u_final, dim3(1), dim3(NOALLOC_NT), 0, s, part, (int)nwg, maxU);
  if (hipGetLastError() != hipSuccess) {
    set_ctxfree_error("find_max_u_noalloc: launch failed");
    return CUP2D_ERR_HIP;
  }
  return CUP2D_OK;
}
