// stencils.hip -- full-grid stencil kernels of the CubismUP_2D step on gfx950.
//
// Grid layout: dense row-major planes (see internal.hpp).  Each stencil
// kernel runs one workgroup over a strip of TX columns x RY rows; every
// thread owns one column and marches up the rows with a three-row register
// window, so each plane is read once from HBM (left/right neighbours come
// from the same cache lines).  Domain faces use the zero-gradient ghost of
// BlockLabOpen (Definitions.h:125-157), multiplied by advDiff's inflow fade
// where the reference fades the lab (advDiff.cpp:50-55,68-71).
//
// All arithmetic follows the reference expression order and is compiled with
// -ffp-contract=off, so every cell value is the reference's IEEE result.
#include "device_util.hpp"

namespace cup2d {

constexpr int BC_KILL_FAC = 100;  // Definitions.h:38
constexpr double EPSD = 2.220446049250313e-16;

int grid_workgroups(int NX, int NY);

int launch_check(cup2d_ctx* c, const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    c->err = std::string(what) + ": " + hipGetErrorString(e);
    return CUP2D_ERR_HIP;
  }
  return CUP2D_OK;
}

// Launch shape of the full-grid stencil kernels (chosen by a sweep over
// RY in {8,16,32}, TX in {128,256}, unroll in {2,4,8}): rows per workgroup,
// widest column strip, unroll depth of the row march.
constexpr int STENCIL_RY = 8, STENCIL_TX = 256;
#define CUP2D_UNROLL_ROWS _Pragma("unroll 4")

static inline int strip_width(int NX) {
  for (int t = STENCIL_TX; t >= 32; t >>= 1)
    if (NX % t == 0) return t;
  return 32;
}
// The lanes of a strip workgroup's wave whose left / right neighbour column
// lies outside the wave (the other wave, or the next strip).
struct WaveEdges {
  bool w, e;
  __device__ WaveEdges() {
    const int lane = threadIdx.x & 63;
    const int last = (blockDim.x < 64 ? (int)blockDim.x : 64) - 1;
    w = lane == 0;
    e = lane == last;
  }
};

static constexpr int RYS_ = STENCIL_RY;
// workgroups of the stencil kernels (advDiff, RHS, projection share a grid)
int grid_workgroups(int NX, int NY) { return (NX / strip_width(NX)) * (NY / RYS_); }

// ------------------------------------------------------------ findMaxU
// Helpers.cpp:136-176: max over cells of |u+Uinf|, |v+Vinf|, |u|, |v|.
// Writes one partial per workgroup into part[0..gridDim.x) and clears the
// nclear obstacle-block partials that follow.
template <typename T>
__global__ __launch_bounds__(256) void k_find_max_u(const T* __restrict__ u,
                                                    const T* __restrict__ v, size_t n,
                                                    const StepState* __restrict__ st,
                                                    double* __restrict__ part, int nclear) {
  __shared__ double lds[8];
  const double UINF = st->uinf[0], VINF = st->uinf[1];
  double m = 0;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
    const double a = u[i], b = v[i];
    m = nmax(m, fabs(a + UINF));
    m = nmax(m, fabs(b + VINF));
    m = nmax(m, fabs(a));
    m = nmax(m, fabs(b));
  }
  m = block_max<256>(m, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = m;
  if (blockIdx.x == 0)
    for (int i = threadIdx.x; i < nclear; i += 256) part[gridDim.x + i] = 0;
}

// Simulation::calcMaxTimestep (Simulation.cpp:254-277).  rampCFL for
// step<100 is tabulated on the host with the same libm expression.
__global__ __launch_bounds__(256) void k_dt(const double* __restrict__ part, int nparts, Grid g,
                                            const cup2d_shape_state* __restrict__ ss,
                                            const double* __restrict__ ramp,
                                            StepState* __restrict__ st, int set_dt) {
  __shared__ double lds[8];
  const double maxU = block_max<256>(max_array<256>(part, nparts), lds);
  if (threadIdx.x == 0) {
    st->maxU = maxU;
    if (!finite_max(maxU) && st->error == 0) st->error = CUP2D_ERR_NONFINITE;
    if (set_dt) {
      const double h = g.h;
      const double dtFourier = h * h / g.nu, dtCFL = maxU < 2.2e-16 ? 1 : h / maxU;
      double maxUb = 0;
      for (int k = 0; k < g.nshapes; ++k)
        maxUb = dmax(maxUb, sqrt(ss[k].u * ss[k].u + ss[k].v * ss[k].v));
      const double dtBody = maxUb < 2.2e-16 ? 1 : h / maxUb;
      const double mn = dmin(dmin(dtCFL, dtFourier), dtBody);
      double dt = g.CFL * mn;
      if (st->step < 100) dt = ramp[st->step] * mn;
      st->dt = dt;
    }
  }
}

int enqueue_find_max_u_partials(cup2d_ctx* c) {
  TimerScope ts(c, T_MAXU);
  CUP2D_REAL_DISPATCH(c, hipLaunchKernelGGL(k_find_max_u<T>, dim3(c->nwg), dim3(256), 0,
                                            c->stream, tp<T>(c->u), tp<T>(c->v),
                                            (size_t)c->g.NX * c->g.NY, c->st, c->maxpart,
                                            c->maxpart_n - c->nwg));
  return launch_check(c, "find_max_u");
}

static int find_max_u_and_dt(cup2d_ctx* c, int set_dt) {
  int rc = enqueue_find_max_u_partials(c);
  if (rc) return rc;
  TimerScope ts(c, T_DT);
  hipLaunchKernelGGL(k_dt, dim3(1), dim3(256), 0, c->stream, c->maxpart, c->nwg, c->g, c->ss,
                     c->ramp, c->st, set_dt);
  return launch_check(c, "calc_max_timestep");
}
int enqueue_find_max_u(cup2d_ctx* c) { return find_max_u_and_dt(c, 0); }
int enqueue_dt(cup2d_ctx* c) { return find_max_u_and_dt(c, 1); }

__global__ void k_set_dt(StepState* st, double dt) { st->dt = dt; }
int enqueue_set_dt(cup2d_ctx* c, double dt) {
  hipLaunchKernelGGL(k_set_dt, dim3(1), dim3(1), 0, c->stream, c->st, dt);
  return launch_check(c, "set_dt");
}

__global__ void k_step_end(StepState* st) {
  st->time += st->dt;
  st->step += 1;
}
int enqueue_step_end(cup2d_ctx* c) {
  hipLaunchKernelGGL(k_step_end, dim3(1), dim3(1), 0, c->stream, c->st);
  return launch_check(c, "step_end");
}

// Tile (bx, by) of this workgroup in the stencil grid.  Workgroup b of a
// launch runs on XCD b % 8; numbering the tiles so that each XCD gets one
// contiguous band of tiles keeps the halo rows and columns that neighbouring
// tiles share inside one L2.
__device__ inline void stencil_tile(int* bx, int* by) {
  const int gx = gridDim.x, n = gx * gridDim.y;
  int id = blockIdx.x + gx * blockIdx.y;
  if (n % 8 == 0) id = (id % 8) * (n / 8) + id / 8;
  *bx = id % gx;
  *by = id / gx;
}

// ------------------------------------------------------------- advDiff
// advDiff.cpp:14-34 (stencil), :47-55 (factors), :68-79 (faded lab).  Each
// workgroup also writes its share of the boundary flux
// IF = sum_E u - sum_W u + sum_N v - sum_S v (advDiff.cpp:93-102).
// T: the velocity planes' type; the stencil is evaluated in T (FP32 at
// FP32 precision), the factors are formed in double and rounded once.
template <int RY, typename T>
__device__ inline void advdiff_strip(const T* __restrict__ u, const T* __restrict__ v,
                                     T* __restrict__ uo, T* __restrict__ vo,
                                     const Grid& g, const StepState* __restrict__ st,
                                     double* __restrict__ fluxpart, int bx, int by, int gx, int gy,
                                     double* lds) {
  const int NX = g.NX, NY = g.NY;
  const int x = bx * blockDim.x + threadIdx.x;
  const int y0 = by * RY;
  const double dt = st->dt, h = g.h;
  const double dUINF0 = st->uinf[0], dUINF1 = st->uinf[1];
  const T UINF0 = (T)dUINF0, UINF1 = (T)dUINF1;
  const T dfac = (T)((g.nu / h) * (dt / h)), afac = (T)(-0.5 * dt / h);
  const double norUinf = dmax(dmax(fabs(dUINF0), fabs(dUINF1)), EPSD);
  const double tW = dmax(dUINF0, 0.0) / norUinf, tS = dmax(dUINF1, 0.0) / norUinf;
  const double tE = dmin(dUINF0, 0.0) / norUinf, tN = dmin(dUINF1, 0.0) / norUinf;
  const T fadeW = (T)(1 - tW * tW / BC_KILL_FAC), fadeS = (T)(1 - tS * tS / BC_KILL_FAC);
  const T fadeE = (T)(1 - tE * tE / BC_KILL_FAC), fadeN = (T)(1 - tN * tN / BC_KILL_FAC);

  // All loads of the strip are issued first (clamped indices, so every load
  // is unconditional), then the rows are computed with selects of the
  // boundary values: the wave keeps the whole strip's loads in flight.
  // The x neighbours come from the adjacent lanes (lane_from_left/right);
  // only the wave's edge lanes load theirs (uH, vH: the west column for
  // lane 0, the east one for the last lane).
  const size_t i0 = (size_t)y0 * NX + x;
  const bool hasW = x > 0, hasE = x + 1 < NX;
  const size_t dW = hasW ? 1 : 0, dE = hasE ? 1 : 0;
  const WaveEdges we;
  T uC[RY + 2], vC[RY + 2], uH[RY], vH[RY];
  {
    const size_t is = y0 > 0 ? i0 - NX : i0;
    uC[0] = u[is];
    vC[0] = v[is];
#pragma unroll
    for (int iy = 0; iy <= RY; ++iy) {
      const size_t i = y0 + iy < NY ? i0 + (size_t)iy * NX : i0 + (size_t)(iy - 1) * NX;
      uC[iy + 1] = u[i];
      vC[iy + 1] = v[i];
    }
    if (we.w || we.e) {
      const size_t dh = we.w ? (size_t)0 - dW : dE;
#pragma unroll
      for (int iy = 0; iy < RY; ++iy) {
        const size_t i = i0 + (size_t)iy * NX + dh;
        uH[iy] = u[i];
        vH[iy] = v[i];
      }
    }
  }
  T uc = uC[1], vc = vC[1];
  T us = y0 == 0 ? uc * fadeS : uC[0];
  T vs = y0 == 0 ? vc * fadeS : vC[0];
  double flux = 0;
#pragma unroll
  for (int iy = 0; iy < RY; ++iy) {
    const int y = y0 + iy;
    const bool hasN = y + 1 < NY;
    const T un = hasN ? uC[iy + 2] : uc * fadeN, vn = hasN ? vC[iy + 2] : vc * fadeN;
    // (the shifts run before the selects: every lane takes part in them)
    const T sul = lane_from_left(uc), svl = lane_from_left(vc);
    const T sur = lane_from_right(uc), svr = lane_from_right(vc);
    const T ul = we.w ? uH[iy] : sul, vl = we.w ? vH[iy] : svl;
    const T ur = we.e ? uH[iy] : sur, vr = we.e ? vH[iy] : svr;
    const T uw = hasW ? ul : uc * fadeW, vw = hasW ? vl : vc * fadeW;
    const T ue = hasE ? ur : uc * fadeE, ve = hasE ? vr : vc * fadeE;
    const T dUadv = (uc + UINF0) * (ue - uw) + (vc + UINF1) * (un - us);
    const T dUdif = ue + un + uw + us - 4 * uc;
    const T dVadv = (uc + UINF0) * (ve - vw) + (vc + UINF1) * (vn - vs);
    const T dVdif = ve + vn + vw + vs - 4 * vc;
    const T u1 = uc + (afac * dUadv + dfac * dUdif);
    const T v1 = vc + (afac * dVadv + dfac * dVdif);
    const size_t idx = i0 + (size_t)iy * NX;
    uo[idx] = u1;
    vo[idx] = v1;
    // selects, not branches, so the row loop stays one basic block
    flux = x == 0 ? flux - u1 : flux;
    flux = x == NX - 1 ? flux + u1 : flux;
    flux = y == 0 ? flux - v1 : flux;
    flux = y == NY - 1 ? flux + v1 : flux;
    us = uc; vs = vc;
    uc = un; vc = vn;
  }
  const int e = edge_index(bx, by, gx, gy);
  if (e < 0) return;
  flux = block_sum_dyn(flux, lds);
  if (threadIdx.x == 0) fluxpart[e] = flux;
}

// Strips of RY rows x blockDim.x columns.  nstrip_x x nstrip_y strips; with a
// grid smaller than that each workgroup takes the strips id, id + gridDim.x,
// ... (the grid is a multiple of 8, so a workgroup keeps its XCD), which
// leaves workgroup slots free for the obstacle kernels of the side stream.
template <int RY, typename T>
__global__ __launch_bounds__(256) void k_advdiff(const T* __restrict__ u,
                                                 const T* __restrict__ v,
                                                 T* __restrict__ uo, T* __restrict__ vo,
                                                 Grid g, const StepState* __restrict__ st,
                                                 double* __restrict__ fluxpart, int gx, int gy) {
  __shared__ double lds[8];
  const int n = gx * gy;
  for (int id = blockIdx.x; id < n; id += gridDim.x) {
    const int t = (n % 8 == 0) ? (id % 8) * (n / 8) + id / 8 : id;
    advdiff_strip<RY, T>(u, v, uo, vo, g, st, fluxpart, t % gx, t / gx, gx, gy, lds);
  }
}

// one workgroup: corr from the boundary workgroups' partials
__global__ __launch_bounds__(256) void k_flux_reduce(const double* __restrict__ fluxpart, int gx,
                                                     int gy, Grid g, StepState* __restrict__ st) {
  __shared__ double lds[8];
  const double corr = flux_corr(fluxpart, edge_count(gx, gy), g, lds);
  if (threadIdx.x == 0) st->flux_corr = corr;
}

// advDiff.cpp:107-115 applied to the stored velocity (operator-level API).
template <typename T>
__global__ __launch_bounds__(256) void k_flux_fix(T* __restrict__ u, T* __restrict__ v, Grid g,
                                                  const StepState* __restrict__ st) {
  const int NX = g.NX, NY = g.NY;
  const T corr = (T)st->flux_corr;
  const int n = 2 * NY + 2 * NX;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    if (i < NY) u[(size_t)i * NX] += corr;
    else if (i < 2 * NY) u[(size_t)(i - NY) * NX + NX - 1] -= corr;
    else if (i < 2 * NY + NX) v[i - 2 * NY] += corr;
    else v[(size_t)(NY - 1) * NX + (i - 2 * NY - NX)] -= corr;
  }
}

static constexpr int RYS = STENCIL_RY;  // rows per workgroup of the stencil kernels

int enqueue_advdiff(cup2d_ctx* c, int defer_fix) {
  return enqueue_advdiff_on(c, defer_fix, c->stream);
}

int enqueue_advdiff_on(cup2d_ctx* c, int defer_fix, hipStream_t s) {
  const int tx = strip_width(c->g.NX);
  const int gx = c->g.NX / tx, gy = c->g.NY / RYS;
  {
    const int nwg = gx * gy;
    KernelTimer kt(c, T_ADVDIFF);
    CUP2D_REAL_DISPATCH(c, CUP2D_TIMED_LAUNCH(c, (k_advdiff<RYS, T>), dim3(nwg), dim3(tx), 0, s,
                                              tp<T>(c->u), tp<T>(c->v), tp<T>(c->tu),
                                              tp<T>(c->tv), c->g, (const StepState*)c->st,
                                              c->fluxpart, gx, gy));
  }
  std::swap(c->u, c->tu);  // advDiff.cpp:81-87 copy tmpV -> vel
  std::swap(c->v, c->tv);
  c->flux_deferred = defer_fix != 0;
  if (defer_fix) {
    // the correction once, for every reader of the deferred fix
    hipLaunchKernelGGL(k_flux_reduce, dim3(1), dim3(256), 0, s, c->fluxpart, gx, gy, c->g,
                       c->st);
  } else {
    TimerScope ts(c, T_FLUXFIX, s);
    hipLaunchKernelGGL(k_flux_reduce, dim3(1), dim3(256), 0, s, c->fluxpart, gx, gy, c->g,
                       c->st);
    const int n = 2 * c->g.NX + 2 * c->g.NY;
    CUP2D_REAL_DISPATCH(c, hipLaunchKernelGGL(k_flux_fix<T>, dim3((n + 255) / 256), dim3(256), 0,
                                              s, tp<T>(c->u), tp<T>(c->v), c->g, c->st));
  }
  return launch_check(c, "advdiff");
}

// --------------------------------------------------------- pressure RHS
// PressureSingle.cpp:110-128: r = (h/2dt) (u_E - u_W + v_N - v_S), zero-
// gradient ghosts.  With a deferred advDiff flux correction (FIX), boundary
// values are read as u_W + corr, u_E - corr, v_S + corr, v_N - corr, exactly
// the values advDiff.cpp:107-115 would have stored.
// The correction is computed from the advDiff edge partials by the edge
// workgroups, the only ones with boundary cells; workgroup (0,0) stores it
// for the projection.
template <int RY, bool FIX, typename T>
__global__ __launch_bounds__(256) void k_rhs(const T* __restrict__ u, const T* __restrict__ v,
                                             T* __restrict__ r, Grid g,
                                             StepState* __restrict__ st,
                                             const double* __restrict__ fluxpart) {
  const int NX = g.NX, NY = g.NY;
  int bx, by;
  stencil_tile(&bx, &by);
  const int x = bx * blockDim.x + threadIdx.x;
  const int y0 = by * RY;
  const T facDiv = (T)(0.5 * g.h / st->dt);
  T corr = 0;
  if (FIX && edge_index(bx, by, gridDim.x, gridDim.y) >= 0) {
    __shared__ double lds[8];
    const double dcorr = flux_corr(fluxpart, edge_count(gridDim.x, gridDim.y), g, lds);
    corr = (T)dcorr;
    if (bx == 0 && by == 0 && threadIdx.x == 0) st->flux_corr = dcorr;
  }
  size_t idx = (size_t)y0 * NX + x;
  T vc = v[idx];
  if (FIX && y0 == 0) vc = vc + corr;
  const T vs0 = (y0 == 0) ? vc : v[idx - NX];  // y0-1 is never a boundary row
  T vs = vs0;
  CUP2D_UNROLL_ROWS
  for (int iy = 0; iy < RY; ++iy, idx += NX) {
    const int y = y0 + iy;
    T vn;
    if (y + 1 < NY) {
      vn = v[idx + NX];
      if (FIX && y + 1 == NY - 1) vn = vn - corr;
    } else {
      vn = vc;
    }
    T uc = u[idx];
    if (FIX && x == 0) uc = uc + corr;
    if (FIX && x == NX - 1) uc = uc - corr;
    T uw, ue;
    if (x > 0) {
      uw = u[idx - 1];
      if (FIX && x - 1 == 0) uw = uw + corr;
    } else {
      uw = uc;
    }
    if (x + 1 < NX) {
      ue = u[idx + 1];
      if (FIX && x + 1 == NX - 1) ue = ue - corr;
    } else {
      ue = uc;
    }
    const T divVx = ue - uw;
    const T divVy = vn - vs;
    r[idx] = facDiv * (divVx + divVy);
    vs = vc;
    vc = vn;
  }
}

// stencil grid of k_advdiff / k_rhs / k_project and its edge workgroups
int stencil_edge_count(const cup2d_ctx* c) {
  const int tx = strip_width(c->g.NX);
  return edge_count(c->g.NX / tx, c->g.NY / RYS);
}

int enqueue_pressure_rhs(cup2d_ctx* c) {
  const int tx = strip_width(c->g.NX);
  TimerScope ts(c, T_RHS);
  const dim3 grid(c->g.NX / tx, c->g.NY / RYS);
  if (c->flux_deferred)
    CUP2D_REAL_DISPATCH(c, hipLaunchKernelGGL((k_rhs<RYS, true, T>), grid, dim3(tx), 0,
                                              c->stream, tp<T>(c->u), tp<T>(c->v), tp<T>(c->rhs),
                                              c->g, c->st, c->fluxpart));
  else
    CUP2D_REAL_DISPATCH(c, hipLaunchKernelGGL((k_rhs<RYS, false, T>), grid, dim3(tx), 0,
                                              c->stream, tp<T>(c->u), tp<T>(c->v), tp<T>(c->rhs),
                                              c->g, c->st, c->fluxpart));
  return launch_check(c, "pressure_rhs");
}


// ------------------------------------------------------- cub2rhs fade
// PoissonSolver.cpp:81-98: the sums of the faded RHS r*fa^2 over the band
// blocks, of its positive parts and of its negative parts; one workgroup per
// band block writes three partials.  k_fade_reduce forms the correction
// factor from them (a redone early row pass forms it itself), and the
// forward row FFT rewrites the band cells.
template <typename T>
__global__ __launch_bounds__(256) void k_fade_sums(const T* __restrict__ r,
                                                   const int* __restrict__ blocks, int nblocks,
                                                   Grid g, double* __restrict__ part,
                                                   const T* __restrict__ u,
                                                   const T* __restrict__ v,
                                                   const unsigned char* __restrict__ obmap,
                                                   const StepState* __restrict__ st) {
  __shared__ double lds[12];
  const int j = blockIdx.x;
  if (j >= nblocks) return;
  // u != nullptr: the RHS of cells outside obstacle blocks is formed from the
  // advDiff output (see k_rows_fwd) with the deferred flux correction, the
  // obstacle blocks' from the RHS plane.  obmap == nullptr: every block from
  // the advDiff output (the obstacle blocks' partials are then replaced once
  // their RHS is final, by the obstacle RHS kernel).
  T corr = 0, facDiv = 0;
  if (u) {
    corr = (T)st->flux_corr;
    facDiv = (T)(0.5 * g.h / st->dt);
  }
  const int b = blocks[j];
  const bool from_plane = !u || (obmap && obmap[b]);
  double a[3];
  fade_block_sums<T>(g, b, r, u, v, from_plane, corr, facDiv, a, lds);
  if (threadIdx.x == 0) {
    part[3 * j] = a[0];
    part[3 * j + 1] = a[1];
    part[3 * j + 2] = a[2];
  }
}

int fade_partials(const cup2d_ctx* c) { return c->band_nblocks; }

// the correction factor of the fade (PoissonSolver.cpp:105-121) from the
// partials of k_fade_sums, once for every reader
__global__ __launch_bounds__(256) void k_fade_reduce(const double* __restrict__ part, int npart,
                                                     StepState* __restrict__ st) {
  __shared__ double lds[64];
  double sums[3];
  block_sum_table<256, 3>(part, npart, 3, sums, lds);
  if (threadIdx.x == 0) st->fade_corr = fade_corr_factor(sums[0], sums[1], sums[2]);
}

// The fade sums of every band block from the advDiff output, on stream s
// right after advDiff (the whole-step RHS); the obstacle RHS kernel replaces
// the partials of obstacle blocks, and enqueue_fade_sums then only reduces.
int enqueue_fade_sums_early(cup2d_ctx* c, hipStream_t s) {
  c->fade_early = false;
  if (c->band_nblocks == 0) return CUP2D_OK;
  CUP2D_REAL_DISPATCH(c, hipLaunchKernelGGL(k_fade_sums<T>, dim3(fade_partials(c)), dim3(256), 0,
                                            s, tp<T>(c->rhs), c->band_blocks, c->band_nblocks,
                                            c->g, c->red, tp<T>(c->u), tp<T>(c->v), nullptr,
                                            c->st));
  c->fade_early = true;
  return launch_check(c, "cub2rhs fade (early)");
}

int enqueue_fade_reduce_on(cup2d_ctx* c, hipStream_t s) {
  if (c->band_nblocks == 0) return CUP2D_OK;
  hipLaunchKernelGGL(k_fade_reduce, dim3(1), dim3(256), 0, s, c->red, fade_partials(c), c->st);
  return launch_check(c, "cub2rhs fade reduce");
}

int enqueue_fade_sums(cup2d_ctx* c, int compute) {
  if (c->band_nblocks == 0) return CUP2D_OK;
  TimerScope ts(c, T_FADE);
  if (!c->fade_early || !compute)
    CUP2D_REAL_DISPATCH(c, hipLaunchKernelGGL(k_fade_sums<T>, dim3(fade_partials(c)), dim3(256),
                                              0, c->stream, tp<T>(c->rhs), c->band_blocks,
                                              c->band_nblocks, c->g, c->red,
                                              compute ? tp<T>(c->u) : nullptr, tp<T>(c->v),
                                              c->obmap, c->st));
  c->fade_early = false;
  hipLaunchKernelGGL(k_fade_reduce, dim3(1), dim3(256), 0, c->stream, c->red, fade_partials(c),
                     c->st);
  return launch_check(c, "cub2rhs fade");
}

// ------------------------------------------------- pressure correction
// PressureSingle.cpp:216-241: u -= (dt/2h)(p_E - p_W), v -= (dt/2h)(p_N - p_S).
// Applies a deferred advDiff flux correction (see k_rhs) and writes the
// findMaxU partial (Helpers.cpp:160-175) of the cells outside obstacle
// blocks; penalization reports the obstacle blocks after it changes them.
// T: the velocity planes' type; the pressure is double.  At FP32 precision
// the gradient of the (FP64-refined) pressure is formed in double and the
// update is rounded to float once.
template <int RY, typename T>
__global__ __launch_bounds__(256) void k_project(T* __restrict__ u, T* __restrict__ v,
                                                 const double* __restrict__ p, Grid g,
                                                 const StepState* __restrict__ st, int fix,
                                                 const unsigned char* __restrict__ obmap,
                                                 double* __restrict__ maxpart) {
  __shared__ double lds[8];
  const int NX = g.NX, NY = g.NY;
  int bx, by;
  stencil_tile(&bx, &by);
  const int x = bx * blockDim.x + threadIdx.x;
  const int y0 = by * RY;
  const double pFac = -0.5 * st->dt / g.h;
  const double UINF = st->uinf[0], VINF = st->uinf[1];
  const double corr = fix ? st->flux_corr : 0.0;
  // All loads of the strip first (clamped indices), then the rows with
  // selects of the boundary values; the RY rows lie in one block row.
  static_assert(BS % RY == 0, "a strip lies in one block row");
  // The x neighbours of p come from the adjacent lanes; only the wave's
  // edge lanes load theirs (pH: west for lane 0, east for the last lane).
  const size_t i0 = (size_t)y0 * NX + x;
  const bool hasW = x > 0, hasE = x + 1 < NX;
  const size_t dW = hasW ? 1 : 0, dE = hasE ? 1 : 0;
  const WaveEdges we;
  double pC[RY + 2], pH[RY];
  T uu[RY], vv[RY];
  pC[0] = p[y0 > 0 ? i0 - NX : i0];
#pragma unroll
  for (int iy = 0; iy <= RY; ++iy)
    pC[iy + 1] = p[y0 + iy < NY ? i0 + (size_t)iy * NX : i0 + (size_t)(iy - 1) * NX];
#pragma unroll
  for (int iy = 0; iy < RY; ++iy) {
    const size_t i = i0 + (size_t)iy * NX;
    uu[iy] = u[i];
    vv[iy] = v[i];
  }
  if (we.w || we.e) {
    const size_t dh = we.w ? (size_t)0 - dW : dE;
#pragma unroll
    for (int iy = 0; iy < RY; ++iy) pH[iy] = p[i0 + (size_t)iy * NX + dh];
  }
  const unsigned char inob = obmap[(y0 / BS) * g.bpdx + x / BS];  // always allocated
  const bool count = (maxpart != nullptr) & (inob == 0);
  double pc = pC[1];
  double ps = y0 == 0 ? pc : pC[0];
  double m = 0;
#pragma unroll
  for (int iy = 0; iy < RY; ++iy) {
    const int y = y0 + iy;
    const double pn = y + 1 < NY ? pC[iy + 2] : pc;
    // (the shifts run before the selects: every lane takes part in them)
    const double spl = lane_from_left(pc), spr = lane_from_right(pc);
    const double pl = we.w ? pH[iy] : spl;
    const double pr = we.e ? pH[iy] : spr;
    const double pw = hasW ? pl : pc;
    const double pe = hasE ? pr : pc;
    // selects, not branches, so the row loop stays one basic block
    T ux = uu[iy], vy = vv[iy];
    const T tcorr = (T)corr;
    ux = fix && x == 0 ? ux + tcorr : (fix && x == NX - 1 ? ux - tcorr : ux);
    vy = fix && y == 0 ? vy + tcorr : (fix && y == NY - 1 ? vy - tcorr : vy);
    const T u1 = (T)((double)ux + pFac * (pe - pw));
    const T v1 = (T)((double)vy + pFac * (pn - ps));
    const size_t idx = i0 + (size_t)iy * NX;
    u[idx] = u1;
    v[idx] = v1;
    double mc = nmax(m, fabs((double)u1 + UINF));
    mc = nmax(mc, fabs((double)v1 + VINF));
    mc = nmax(mc, fabs((double)u1));
    mc = nmax(mc, fabs((double)v1));
    m = count ? mc : m;
    ps = pc;
    pc = pn;
  }
  if (maxpart) {
    m = block_max_dyn(m, lds);
    if (threadIdx.x == 0) maxpart[by * gridDim.x + bx] = m;
  }
}

int enqueue_pressure_correction(cup2d_ctx* c, int want_max) {
  const int tx = strip_width(c->g.NX);
  KernelTimer kt(c, T_PROJECT);
  CUP2D_REAL_DISPATCH(c, CUP2D_TIMED_LAUNCH(c, (k_project<RYS, T>), dim3(c->g.NX / tx,
                                                                         c->g.NY / RYS),
                                            dim3(tx), 0, c->stream, tp<T>(c->u), tp<T>(c->v),
                                            (const double*)c->pres, c->g,
                                            (const StepState*)c->st, c->flux_deferred ? 1 : 0,
                                            (const unsigned char*)c->obmap,
                                            want_max ? c->maxpart : nullptr));
  c->flux_deferred = false;
  return launch_check(c, "pressure_correction");
}

// ------------------------------------------------- PressureIterator_unif
// operator() start (PressureIterator_unif.cpp:219-226): tmpV = vel.
__global__ void k_copy2(const double* __restrict__ a, const double* __restrict__ b,
                        double* __restrict__ ca, double* __restrict__ cb, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    ca[i] = a[i];
    cb[i] = b[i];
  }
}
int enqueue_copy_vel_to_tmpv(cup2d_ctx* c) {
  hipLaunchKernelGGL(k_copy2, dim3(2048), dim3(256), 0, c->stream, c->u, c->v, c->tu, c->tv,
                     (size_t)c->g.NX * c->g.NY);
  return launch_check(c, "copy vel to tmpV");
}

// pressureCorrection (PressureIterator_unif.cpp:51-75): vFluid = tmpV + pFac grad p.
template <int RY>
__global__ __launch_bounds__(256) void k_iter_project(const double* __restrict__ tu,
                                                      const double* __restrict__ tv,
                                                      const double* __restrict__ p,
                                                      double* __restrict__ fu,
                                                      double* __restrict__ fv, Grid g,
                                                      const StepState* __restrict__ st) {
  const int NX = g.NX, NY = g.NY;
  int bx, by;
  stencil_tile(&bx, &by);
  const int x = bx * blockDim.x + threadIdx.x;
  const int y0 = by * RY;
  const double pFac = -0.5 * st->dt / g.h;
  size_t idx = (size_t)y0 * NX + x;
  double pc = p[idx];
  double ps = (y0 == 0) ? pc : p[idx - NX];
  CUP2D_UNROLL_ROWS
  for (int iy = 0; iy < RY; ++iy, idx += NX) {
    const int y = y0 + iy;
    const double pn = (y + 1 < NY) ? p[idx + NX] : pc;
    const double pw = (x > 0) ? p[idx - 1] : pc;
    const double pe = (x + 1 < NX) ? p[idx + 1] : pc;
    fu[idx] = tu[idx] + pFac * (pe - pw);
    fv[idx] = tv[idx] + pFac * (pn - ps);
    ps = pc;
    pc = pn;
  }
}
int enqueue_iter_project(cup2d_ctx* c) {
  const int tx = strip_width(c->g.NX);
  TimerScope ts(c, T_PROJECT);
  hipLaunchKernelGGL(k_iter_project<RYS>, dim3(c->g.NX / tx, c->g.NY / RYS), dim3(tx), 0,
                     c->stream, c->tu, c->tv, c->pres, c->vfu, c->vfv, c->g, c->st);
  return launch_check(c, "iterator projection");
}

// updatePressureRHS (PressureIterator_unif.cpp:22-49):
// r = (h/2dt) (div vel - CHI div uDef), zero-gradient ghosts.
template <int RY>
__global__ __launch_bounds__(256) void k_iter_rhs(const double* __restrict__ u,
                                                  const double* __restrict__ v,
                                                  const double* __restrict__ udx,
                                                  const double* __restrict__ udy,
                                                  const double* __restrict__ chi,
                                                  double* __restrict__ r, Grid g,
                                                  const StepState* __restrict__ st) {
  const int NX = g.NX, NY = g.NY;
  int bx, by;
  stencil_tile(&bx, &by);
  const int x = bx * blockDim.x + threadIdx.x;
  const int y0 = by * RY;
  const double facDiv = 0.5 * g.h / st->dt;
  CUP2D_UNROLL_ROWS
  for (int iy = 0; iy < RY; ++iy) {
    const int y = y0 + iy;
    const size_t idx = (size_t)y * NX + x;
    const size_t ie = x + 1 < NX ? idx + 1 : idx, iw = x > 0 ? idx - 1 : idx;
    const size_t in = y + 1 < NY ? idx + NX : idx, is = y > 0 ? idx - NX : idx;
    const double divVx = u[ie] - u[iw];
    const double divVy = v[in] - v[is];
    const double divUSx = udx[ie] - udx[iw];
    const double divUSy = udy[in] - udy[is];
    r[idx] = facDiv * (divVx + divVy - chi[idx] * (divUSx + divUSy));
  }
}
int enqueue_iter_rhs(cup2d_ctx* c) {
  const int tx = strip_width(c->g.NX);
  TimerScope ts(c, T_RHS);
  hipLaunchKernelGGL(k_iter_rhs<RYS>, dim3(c->g.NX / tx, c->g.NY / RYS), dim3(tx), 0, c->stream,
                     c->u, c->v, c->udx, c->udy, c->chi, c->rhs, c->g, c->st);
  return launch_check(c, "iterator rhs");
}

// ------------------------------------------------------ IC / ApplyObjVel
// IC::operator() (Helpers.cpp:38-82) plus the PutObjectsOnGrid clear values
// (chi=0, udef=0, invRho=1; PutObjectsOnGrid.cpp:191-197).
__global__ void k_ic(double* u, double* v, double* tu, double* tv, double* pres, double* rhs,
                     double* chi, double* udx, double* udy, double* invrho, double* vfu,
                     double* vfv, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    u[i] = 0; v[i] = 0; tu[i] = 0; tv[i] = 0; pres[i] = 0; rhs[i] = 0;
    chi[i] = 0; udx[i] = 0; udy[i] = 0; invrho[i] = 1; vfu[i] = 0; vfv[i] = 0;
  }
}
__global__ void k_ic_state(StepState* st) {
  st->time = 0; st->dt = 0; st->step = 0; st->uinf[0] = 0; st->uinf[1] = 0;
  st->error = 0;
  st->maxU = 0; st->fade_corr = 0; st->flux_corr = 0;
  st->dt_done = 0;
}
int enqueue_ic(cup2d_ctx* c) {
  hipLaunchKernelGGL(k_ic, dim3(2048), dim3(256), 0, c->stream, c->u, c->v, c->tu, c->tv,
                     c->pres, c->rhs, c->chi, c->udx, c->udy, c->invrho, c->vfu, c->vfv,
                     (size_t)c->g.NX * c->g.NY);
  hipLaunchKernelGGL(k_ic_state, dim3(1), dim3(1), 0, c->stream, c->st);
  return launch_check(c, "ic");
}

// Helpers.cpp:290-307: u = u(1-chi) + udef*chi.
template <typename T>
__global__ void k_apply_obj_vel(T* u, T* v, const double* chi, const double* udx,
                                const double* udy, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const double X = chi[i];
    u[i] = (T)((double)u[i] * (1 - X) + udx[i] * X);
    v[i] = (T)((double)v[i] * (1 - X) + udy[i] * X);
  }
}
int enqueue_apply_obj_vel(cup2d_ctx* c) {
  CUP2D_REAL_DISPATCH(c, hipLaunchKernelGGL(k_apply_obj_vel<T>, dim3(2048), dim3(256), 0,
                                            c->stream, tp<T>(c->u), tp<T>(c->v), c->chi, c->udx,
                                            c->udy, (size_t)c->g.NX * c->g.NY));
  return launch_check(c, "apply_obj_vel");
}

// ------------------------------------------------------ layout converters
// Block layout: [(by*bpdx+bx)][iy][ix][comp]; dense: plane[(by*32+iy)*NX + bx*32+ix].
// T: the planes' type (float for the velocity and RHS planes at FP32
// precision); the staging buffer is double.
template <typename T>
__global__ void k_blocks_to_dense(const double* __restrict__ s, T* a, T* b, int nc, Grid g,
                                  size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const size_t blk = i / (BS * BS);
    const int k = (int)(i % (BS * BS)), ix = k % BS, iy = k / BS;
    const int bx = (int)(blk % g.bpdx), by = (int)(blk / g.bpdx);
    const size_t d = (size_t)(by * BS + iy) * g.NX + bx * BS + ix;
    a[d] = (T)s[i * nc];
    if (nc > 1) b[d] = (T)s[i * nc + 1];
  }
}
template <typename T>
__global__ void k_dense_to_blocks(const T* __restrict__ a, const T* __restrict__ b, double* s,
                                  int nc, Grid g, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const size_t blk = i / (BS * BS);
    const int k = (int)(i % (BS * BS)), ix = k % BS, iy = k / BS;
    const int bx = (int)(blk % g.bpdx), by = (int)(blk / g.bpdx);
    const size_t d = (size_t)(by * BS + iy) * g.NX + bx * BS + ix;
    s[i * nc] = (double)a[d];
    if (nc > 1) s[i * nc + 1] = (double)b[d];
  }
}
int enqueue_blocks_to_dense(cup2d_ctx* c, const double* stage, double* a, double* b, int nc,
                            bool f32) {
  const size_t n = (size_t)c->g.NX * c->g.NY;
  if (f32)
    hipLaunchKernelGGL(k_blocks_to_dense<float>, dim3(2048), dim3(256), 0, c->stream, stage,
                       tp<float>(a), b ? tp<float>(b) : nullptr, nc, c->g, n);
  else
    hipLaunchKernelGGL(k_blocks_to_dense<double>, dim3(2048), dim3(256), 0, c->stream, stage, a,
                       b, nc, c->g, n);
  return launch_check(c, "blocks_to_dense");
}
int enqueue_dense_to_blocks(cup2d_ctx* c, const double* a, const double* b, double* stage,
                            int nc, bool f32) {
  const size_t n = (size_t)c->g.NX * c->g.NY;
  if (f32)
    hipLaunchKernelGGL(k_dense_to_blocks<float>, dim3(2048), dim3(256), 0, c->stream,
                       tp<float>(a), b ? tp<float>(b) : nullptr, stage, nc, c->g, n);
  else
    hipLaunchKernelGGL(k_dense_to_blocks<double>, dim3(2048), dim3(256), 0, c->stream, a, b,
                       stage, nc, c->g, n);
  return launch_check(c, "dense_to_blocks");
}

// dense double staging <-> a float plane (FP32 precision)
__global__ void k_stage_to_f32(const double* __restrict__ s, float* __restrict__ a, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    a[i] = (float)s[i];
}
__global__ void k_f32_to_stage(const float* __restrict__ a, double* __restrict__ s, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    s[i] = (double)a[i];
}
int enqueue_stage_to_f32(cup2d_ctx* c, const double* stage, double* plane) {
  hipLaunchKernelGGL(k_stage_to_f32, dim3(2048), dim3(256), 0, c->stream, stage, tp<float>(plane),
                     (size_t)c->g.NX * c->g.NY);
  return launch_check(c, "stage_to_f32");
}
int enqueue_f32_to_stage(cup2d_ctx* c, const double* plane, double* stage) {
  hipLaunchKernelGGL(k_f32_to_stage, dim3(2048), dim3(256), 0, c->stream, tp<float>(plane), stage,
                     (size_t)c->g.NX * c->g.NY);
  return launch_check(c, "f32_to_stage");
}

}  // namespace cup2d
