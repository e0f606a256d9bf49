// adjoint.hip -- the discrete adjoint of the whole step for a forced, fixed
// disk, with respect to its imposed surface velocity.
//
// Objective: J = (1/N) sum_n drag_n, drag_n = ComputeForces' drag of the
// disk after step n (ComputeForces.cpp:91-93: sum over surface points of
// max(0, -F.u_hat)).  Parameters: the Fourier coefficients of the disk's
// surface velocity (cup2d_set_surface_velocity), which enter the step only
// through the penalization target (PressureSingle.cpp:96-107).  dt of each
// step is held at the forward run's value (the CFL dt is a non-smooth
// function of max|u|), and the body is forced and fixed, so its velocity,
// position and chi are the same every step.
//
// The forward step (operator path, FP64, explicit penalization):
//   w     = advDiff(u_n)                        advDiff.cpp:14-79
//   wf    = (I - e e^T / nb) w                   boundary-flux correction,
//                                               advDiff.cpp:89-116 (e: +1 on
//                                               the W/S faces, -1 on E/N)
//   r     = facDiv div(wf)                       PressureSingle.cpp:112-128
//   r     = obstacle correction of r             PressureSingle.cpp:130-213
//   r     = cub2rhs fade of r                    PoissonSolver.cpp:81-121
//   p     = S r                                  FFT Poisson solve
//   z     = wf + pFac grad p                     PressureSingle.cpp:216-241
//   u_n+1 = (1 - chi) z + chi (U_s + udef)       PressureSingle.cpp:69-108
//   drag  = D(u_n+1, p)                          ComputeForces.cpp:50-93
// Its adjoint, run backwards one step at a time (kernels below, in the order
// the host calls them): drag -> penalization (and the parameter gradient) ->
// projection -> S (the solve is symmetric: even Green's function, real even
// symbols) -> fade -> obstacle correction -> divergence -> flux correction
// (a symmetric projector) -> advDiff linearized at u_n.  The fade and the
// obstacle correction rescale signed parts by ratios of sums; their adjoints
// hold the sign pattern (active set) of the forward values and carry the
// derivative of the ratio, so the adjoint is exact wherever no value sits on
// a sign change, as it is for the drag's max(0, .) per surface point.
//
// Vectors in the adjoint are dense planes like the fields.  The drag's
// surface stencils scatter with atomic adds (a few hundred surface points);
// every other kernel gathers.
#include "device_util.hpp"

namespace cup2d {

namespace {
constexpr int ANT = 256;
constexpr double EPSD = 2.220446049250313e-16;
constexpr int BC_KILL = 100;  // Definitions.h:38

__device__ inline size_t cidx(const Grid& g, int x, int y) {
  x = x < 0 ? 0 : (x >= g.NX ? g.NX - 1 : x);
  y = y < 0 ? 0 : (y >= g.NY ? g.NY - 1 : y);
  return (size_t)y * g.NX + x;
}

struct ASlot {
  const double *chi, *sd, *sx, *sy;
};
__device__ inline ASlot aslot(const Obstacles& ob, int s) {
  const size_t o = (size_t)s * BS * BS;
  return ASlot{ob.chi + o, ob.sdelta + o, ob.sdx + o, ob.sdy + o};
}

// r of one cell from the flux-corrected advDiff output (k_rhs, no deferred fix)
__device__ inline double rhs_raw(const Grid& g, const double* __restrict__ u,
                                 const double* __restrict__ v, int x, int y, double facDiv) {
  const size_t i = (size_t)y * g.NX + x;
  const double uc = u[i], vc = v[i];
  const double ue = x + 1 < g.NX ? u[i + 1] : uc, uw = x > 0 ? u[i - 1] : uc;
  const double vn = y + 1 < g.NY ? v[i + g.NX] : vc, vs = y > 0 ? v[i - g.NX] : vc;
  const double divVx = ue - uw;
  const double divVy = vn - vs;
  return facDiv * (divVx + divVy);
}

// (G^T a)_j for the central difference with zero-gradient ends,
// (G p)_i = p_min(i+1,N-1) - p_max(i-1,0), along one direction: stride s,
// index k of N.
__device__ inline double gradT(const double* __restrict__ a, size_t j, int k, int N, size_t s) {
  double r = 0;
  if (k >= 1) r += a[j - s];
  if (k <= N - 2) r -= a[j + s];
  if (k == N - 1) r += a[j];
  if (k == 0) r -= a[j];
  return r;
}
}  // namespace

// drag of the step (1/N weight): surface points with F.u_hat < 0, scattered
// onto the adjoints of u, v and p.  One workgroup per slot.
__global__ __launch_bounds__(ANT) void k_adj_forces(Grid g, Obstacles ob, int shape,
                                                    const double* __restrict__ u,
                                                    const double* __restrict__ v,
                                                    const double* __restrict__ p, double vu0,
                                                    double vu1, double wJ, double* ub,
                                                    double* vb, double* pb) {
  const int s = blockIdx.x;
  if (s >= ob.cnt[0] || ob.slot_shape[s] != shape) return;
  const int b = ob.slot_block[s], bx = b % g.bpdx, by = b / g.bpdx;
  const ASlot o = aslot(ob, s);
  const double NUoH = g.nu / g.h;
  for (int i = threadIdx.x; i < BS * BS; i += ANT) {
    if (!(o.sd[i] > 0)) continue;
    const int gx = bx * BS + i % BS, gy = by * BS + i / BS;
    const double D11 = NUoH * (u[cidx(g, gx + 1, gy)] - u[cidx(g, gx - 1, gy)]);
    const double D22 = NUoH * (v[cidx(g, gx, gy + 1)] - v[cidx(g, gx, gy - 1)]);
    const double D12 = NUoH / 2 *
                       (u[cidx(g, gx, gy + 1)] - u[cidx(g, gx, gy - 1)] +
                        v[cidx(g, gx + 1, gy)] - v[cidx(g, gx - 1, gy)]);
    const double nX = o.sx[i], nY = o.sy[i];
    const double Pc = p[(size_t)gy * g.NX + gx];
    const double fXT = D11 * nX + D12 * nY - Pc * nX;
    const double fYT = D12 * nX + D22 * nY - Pc * nY;
    const double fp = fXT * vu0 + fYT * vu1;
    if (!(fp < 0)) continue;  // drag = max(0, -fp)
    const double gX = -wJ * vu0, gY = -wJ * vu1;
    const double gD11 = gX * nX, gD12 = gX * nY + gY * nX, gD22 = gY * nY;
    const double gP = -(gX * nX + gY * nY);
    const double c12 = NUoH / 2 * gD12;
    atomicAdd(ub + cidx(g, gx + 1, gy), NUoH * gD11);
    atomicAdd(ub + cidx(g, gx - 1, gy), -NUoH * gD11);
    atomicAdd(vb + cidx(g, gx, gy + 1), NUoH * gD22);
    atomicAdd(vb + cidx(g, gx, gy - 1), -NUoH * gD22);
    atomicAdd(ub + cidx(g, gx, gy + 1), c12);
    atomicAdd(ub + cidx(g, gx, gy - 1), -c12);
    atomicAdd(vb + cidx(g, gx + 1, gy), c12);
    atomicAdd(vb + cidx(g, gx - 1, gy), -c12);
    atomicAdd(pb + (size_t)gy * g.NX + gx, gP);
  }
}

// explicit penalization u <- (1 - chi) u + chi (U_s + udef): the adjoint of
// the penalized cells scales by 1 - chi; the surface-velocity coefficients
// receive chi (ub, vb) . d(udef)/d(coef), per workgroup into part[wg][4M].
// One workgroup per distinct obstacle block.
// Modes per accumulation pass of k_adj_penalize: its 4 x ADJ_MCH sums stay
// in registers (4 x SURF_MAX_MODES indexed by a run-time mode would live in
// scratch memory).
constexpr int ADJ_MCH = 8;
constexpr int ADJ_CPT = BS * BS / ANT;  // cells per thread of a block
static_assert(BS * BS % ANT == 0, "whole cells per thread");

__global__ __launch_bounds__(ANT) void k_adj_penalize(Grid g, Obstacles ob, int shape,
                                                      const double* __restrict__ chi,
                                                      const cup2d_shape_state* __restrict__ ss,
                                                      SurfVel sv, double* ub, double* vb,
                                                      double* __restrict__ part) {
  __shared__ double lds[(ANT / 64) * 4 * ADJ_MCH];
  const int M = sv.M;
  double* out = part + (size_t)blockIdx.x * 4 * M;
  const int nub = ob.cnt[1];
  int s = -1;  // this block's slot of the shape (a slot is one block and shape)
  if ((int)blockIdx.x < nub) {
    const int s0 = ob.ub_start[blockIdx.x], cnt = ob.ub_count[blockIdx.x];
    for (int q = 0; q < cnt; ++q)
      if (ob.slot_shape[ob.ub_slots[s0 + q]] == shape) s = ob.ub_slots[s0 + q];
  }
  if (s < 0) {  // no cell of the shape: zero partials
    for (int k = threadIdx.x; k < 4 * M; k += ANT) out[k] = 0;
    return;
  }
  // per cell: the weights of the normal and tangential parts and the unit
  // direction z from the centre of mass, whose powers z^(m+1) carry mode m;
  // the adjoint velocity is scaled by (1 - chi) on the way
  const int b = ob.ub_block[blockIdx.x], bx = b % g.bpdx, by = b / g.bpdx;
  const ASlot o = aslot(ob, s);
  const double Cx = ss[shape].centerOfMass[0], Cy = ss[shape].centerOfMass[1];
  double gn[ADJ_CPT], gt[ADJ_CPT], zr[ADJ_CPT], zi[ADJ_CPT], wr[ADJ_CPT], wi[ADJ_CPT];
#pragma unroll
  for (int j = 0; j < ADJ_CPT; ++j) {
    const int i = threadIdx.x + j * ANT;
    const int ix = i % BS, iy = i / BS;
    const size_t d = (size_t)(by * BS + iy) * g.NX + bx * BS + ix;
    const double X = o.chi[i];
    gn[j] = gt[j] = 0;
    zr[j] = 1;
    zi[j] = 0;
    wr[j] = 1;
    wi[j] = 0;
    if (chi[d] > X || X <= 0) continue;
    const double a = ub[d], c = vb[d];
    ub[d] = (1 - X) * a;
    vb[d] = (1 - X) * c;
    const double p0 = cell_pos(bx, ix, g.hb, g.h) - Cx, p1 = cell_pos(by, iy, g.hb, g.h) - Cy;
    const double r = sqrt(p0 * p0 + p1 * p1);
    zr[j] = r > 0 ? p0 / r : 1.0;
    zi[j] = r > 0 ? p1 / r : 0.0;
    gn[j] = X * (a * zr[j] + c * zi[j]);
    gt[j] = X * (-a * zi[j] + c * zr[j]);
  }
  for (int m0 = 0; m0 < M; m0 += ADJ_MCH) {
    double acc[4 * ADJ_MCH];
#pragma unroll
    for (int k = 0; k < 4 * ADJ_MCH; ++k) acc[k] = 0;
#pragma unroll
    for (int j = 0; j < ADJ_CPT; ++j) {
#pragma unroll
      for (int mm = 0; mm < ADJ_MCH; ++mm) {
        if (m0 + mm >= M) continue;  // uniform; keeps acc's indices static
        const double tr = wr[j] * zr[j] - wi[j] * zi[j];
        wi[j] = wr[j] * zi[j] + wi[j] * zr[j];
        wr[j] = tr;
        acc[mm] += gn[j] * wr[j];
        acc[ADJ_MCH + mm] += gn[j] * wi[j];
        acc[2 * ADJ_MCH + mm] += gt[j] * wr[j];
        acc[3 * ADJ_MCH + mm] += gt[j] * wi[j];
      }
    }
    block_sum_multi<ANT, 4 * ADJ_MCH>(acc, lds);
    if (threadIdx.x == 0) {
#pragma unroll
      for (int mm = 0; mm < ADJ_MCH; ++mm) {
        if (m0 + mm >= M) continue;
#pragma unroll
        for (int q = 0; q < 4; ++q) out[q * M + m0 + mm] = acc[q * ADJ_MCH + mm];
      }
    }
  }
}

// the parameter gradient += sum over workgroups of the partials: one
// workgroup per parameter, the sum in a fixed order (block_sum_array)
__global__ __launch_bounds__(ANT) void k_adj_theta(const double* __restrict__ part, int nwg,
                                                   int n4, double* __restrict__ theta) {
  __shared__ double lds[ANT / 64];
  const int k = blockIdx.x;
  const double s = block_sum_array<ANT>(part + k, nwg, n4, lds);
  if (threadIdx.x == 0) theta[k] += s;
}

// projection z = wf + pFac grad p: pb += pFac (Gx^T ub + Gy^T vb)
__global__ __launch_bounds__(ANT) void k_adj_project(Grid g, double pFac,
                                                     const double* __restrict__ ub,
                                                     const double* __restrict__ vb,
                                                     double* __restrict__ pb) {
  const size_t n = (size_t)g.NX * g.NY;
  for (size_t j = (size_t)blockIdx.x * ANT + threadIdx.x; j < n; j += (size_t)gridDim.x * ANT) {
    const int x = (int)(j % g.NX), y = (int)(j / g.NX);
    pb[j] += pFac * (gradT(ub, j, x, g.NX, 1) + gradT(vb, j, y, g.NY, g.NX));
  }
}

// cub2rhs fade, pass 1: per band block, the forward sums of f = r fa^2
// (total, positive and negative parts) and the adjoint sums
// sum rb max(f,0) and sum rb (-min(f,0)); r recomputed from wf.
__global__ __launch_bounds__(ANT) void k_adj_fade1(Grid g, const int* __restrict__ blocks,
                                                   int nblocks, double facDiv,
                                                   const double* __restrict__ wu,
                                                   const double* __restrict__ wv,
                                                   const double* __restrict__ rb,
                                                   double* __restrict__ part) {
  __shared__ double lds[(ANT / 64) * 5];
  const int j = blockIdx.x;
  if (j >= nblocks) return;
  const int b = blocks[j], bx = b % g.bpdx, by = b / g.bpdx;
  double a[5] = {0, 0, 0, 0, 0};
  for (int i = threadIdx.x; i < BS * BS; i += ANT) {
    const int ix = i % BS, iy = i / BS, x = bx * BS + ix, y = by * BS + iy;
    const double fa = fade_arg(g, cell_pos(bx, ix, g.hb, g.h), cell_pos(by, iy, g.hb, g.h));
    const double f = rhs_raw(g, wu, wv, x, y, facDiv) * (fa * fa);
    const double rbar = rb[(size_t)y * g.NX + x];
    a[0] += f;
    if (f > 0) { a[1] += f; a[3] += rbar * f; }
    if (f < 0) { a[2] -= f; a[4] -= rbar * f; }
  }
  block_sum_multi<ANT, 5>(a, lds);
  if (threadIdx.x == 0)
    for (int k = 0; k < 5; ++k) part[5 * j + k] = a[k];
}

// cub2rhs fade, pass 2: every workgroup sums the partials (same order), forms
// corr = Sr / (Sr > 0 ? P : N) and applies the transposed fade to its block.
__global__ __launch_bounds__(ANT) void k_adj_fade2(Grid g, const int* __restrict__ blocks,
                                                   int nblocks, double facDiv,
                                                   const double* __restrict__ wu,
                                                   const double* __restrict__ wv,
                                                   const double* __restrict__ part,
                                                   double* __restrict__ rb) {
  __shared__ double lds[(ANT / 64) * 5];
  const int j = blockIdx.x;
  if (j >= nblocks) return;
  double tot[5];
  block_sum_table<ANT, 5>(part, nblocks, 5, tot, lds);  // the same order in every workgroup
  const double Sr = tot[0], P = dmax(tot[1], EPSD), Nn = dmax(tot[2], EPSD);
  const double corr = Sr > 0 ? Sr / P : Sr / Nn;
  const double Cbar = corr > 0 ? tot[3] : tot[4];
  // a denominator held at epsilon (fade_corr_factor's clamp) has no derivative
  const double kP = tot[1] < EPSD ? 0.0 : 1.0, kN = tot[2] < EPSD ? 0.0 : 1.0;
  const int b = blocks[j], bx = b % g.bpdx, by = b / g.bpdx;
  for (int i = threadIdx.x; i < BS * BS; i += ANT) {
    const int ix = i % BS, iy = i / BS, x = bx * BS + ix, y = by * BS + iy;
    const double fa = fade_arg(g, cell_pos(bx, ix, g.hb, g.h), cell_pos(by, iy, g.hb, g.h));
    const double fa2 = fa * fa;
    const double f = rhs_raw(g, wu, wv, x, y, facDiv) * fa2;
    const size_t d = (size_t)y * g.NX + x;
    const double direct = (1 - fa2) + fa2 * (corr > 0 ? corr * (f > 0) : -corr * (f < 0));
    const double dcorr = Sr > 0 ? (1 - kP * corr * (f > 0)) / P : (1 + kN * corr * (f < 0)) / Nn;
    rb[d] = rb[d] * direct + Cbar * dcorr * fa2;
  }
}

// obstacle RHS correction, pass 1: per distinct block, the forward sums of
// the shape (S = sum -chi r, positive and negative perimeter parts of
// T = r (1 - chi)) and the adjoint sums for both signs of corr:
// sum rb (-T) [perimeter, T > 0] and sum rb T [perimeter, T < 0].
__global__ __launch_bounds__(ANT) void k_adj_obst1(Grid g, Obstacles ob, int shape,
                                                   double facDiv,
                                                   const double* __restrict__ chi,
                                                   const double* __restrict__ wu,
                                                   const double* __restrict__ wv,
                                                   const double* __restrict__ rb,
                                                   double* __restrict__ part) {
  __shared__ double lds[(ANT / 64) * 5];
  double a[5] = {0, 0, 0, 0, 0};
  const int nub = ob.cnt[1];
  if ((int)blockIdx.x < nub) {
    const int b = ob.ub_block[blockIdx.x], bx = b % g.bpdx, by = b / g.bpdx;
    const int s0 = ob.ub_start[blockIdx.x], cnt = ob.ub_count[blockIdx.x];
    for (int q = 0; q < cnt; ++q) {
      const int s = ob.ub_slots[s0 + q];
      if (ob.slot_shape[s] != shape) continue;
      const ASlot o = aslot(ob, s);
      for (int i = threadIdx.x; i < BS * BS; i += ANT) {
        const int x = bx * BS + i % BS, y = by * BS + i / BS;
        const size_t d = (size_t)y * g.NX + x;
        if (chi[d] > o.chi[i]) continue;
        const double gx = chi[cidx(g, x + 1, y)] - chi[cidx(g, x - 1, y)];
        const double gy = chi[cidx(g, x, y + 1)] - chi[cidx(g, x, y - 1)];
        const bool perim = gx * gx + gy * gy > 0;
        const double rr = rhs_raw(g, wu, wv, x, y, facDiv);
        const double src = -o.chi[i] * rr;
        const double T = rr + src;
        a[0] += src;
        if (perim && T > 0) { a[1] += T; a[3] -= rb[d] * T; }
        if (perim && T < 0) { a[2] -= T; a[4] += rb[d] * T; }
      }
    }
  }
  block_sum_multi<ANT, 5>(a, lds);
  if (threadIdx.x == 0)
    for (int k = 0; k < 5; ++k) part[5 * blockIdx.x + k] = a[k];
}

// obstacle RHS correction, pass 2: totals (in order), corr = S / D, and the
// transposed correction and source term on the block's cells.
__global__ __launch_bounds__(ANT) void k_adj_obst2(Grid g, Obstacles ob, int shape,
                                                   double facDiv,
                                                   const double* __restrict__ chi,
                                                   const double* __restrict__ wu,
                                                   const double* __restrict__ wv,
                                                   const double* __restrict__ part,
                                                   double* __restrict__ rb) {
  __shared__ double lds[(ANT / 64) * 5];
  const int nub = ob.cnt[1];
  if ((int)blockIdx.x >= nub) return;
  double tot[5];
  block_sum_table<ANT, 5>(part, nub, 5, tot, lds);  // the same order in every workgroup
  const double S = tot[0];
  const double Draw = S > 0 ? tot[1] : tot[2];
  const double D = dmax(Draw, EPSD);
  const double kD = Draw < EPSD ? 0.0 : 1.0;  // a clamped denominator has no derivative
  const double corr = S / D;
  const double Cbar = corr > 0 ? tot[3] : (corr < 0 ? tot[4] : 0.0);
  const int b = ob.ub_block[blockIdx.x], bx = b % g.bpdx, by = b / g.bpdx;
  const int s0 = ob.ub_start[blockIdx.x], cnt = ob.ub_count[blockIdx.x];
  for (int q = 0; q < cnt; ++q) {
    const int s = ob.ub_slots[s0 + q];
    if (ob.slot_shape[s] != shape) continue;
    const ASlot o = aslot(ob, s);
    for (int i = threadIdx.x; i < BS * BS; i += ANT) {
      const int x = bx * BS + i % BS, y = by * BS + i / BS;
      const size_t d = (size_t)y * g.NX + x;
      if (chi[d] > o.chi[i]) continue;
      const double gx = chi[cidx(g, x + 1, y)] - chi[cidx(g, x - 1, y)];
      const double gy = chi[cidx(g, x, y + 1)] - chi[cidx(g, x, y - 1)];
      const bool perim = gx * gx + gy * gy > 0;
      const double X = o.chi[i];
      const double rr = rhs_raw(g, wu, wv, x, y, facDiv);
      const double T = rr + (-X * rr);
      double sigma = 0;
      if (perim && T > 0 && corr > 0) sigma = 1;
      else if (perim && T < 0 && corr < 0) sigma = -1;
      double dcorr = 0;  // d corr / d T through the denominator
      if (S > 0 && perim && T > 0) dcorr = -kD * corr / D;
      if (!(S > 0) && perim && T < 0) dcorr = kD * corr / D;
      const double Tbar = rb[d] * (1 - sigma * corr) + Cbar * dcorr;
      rb[d] = Tbar * (1 - X) + (Cbar / D) * (-X);
    }
  }
}

// divergence r = facDiv div(wf): wb = ub + facDiv Gx^T rb, vb' = vb + facDiv Gy^T rb
__global__ __launch_bounds__(ANT) void k_adj_div(Grid g, double facDiv,
                                                 const double* __restrict__ ub,
                                                 const double* __restrict__ vb,
                                                 const double* __restrict__ rb,
                                                 double* __restrict__ wbu,
                                                 double* __restrict__ wbv) {
  const size_t n = (size_t)g.NX * g.NY;
  for (size_t j = (size_t)blockIdx.x * ANT + threadIdx.x; j < n; j += (size_t)gridDim.x * ANT) {
    const int x = (int)(j % g.NX), y = (int)(j / g.NX);
    wbu[j] = ub[j] + facDiv * gradT(rb, j, x, g.NX, 1);
    wbv[j] = vb[j] + facDiv * gradT(rb, j, y, g.NY, g.NX);
  }
}

// e^T wb over the domain faces (e: +1 on the W and S faces, -1 on E and N),
// one workgroup, into *out
__global__ __launch_bounds__(ANT) void k_adj_flux_sum(Grid g, const double* __restrict__ wbu,
                                                      const double* __restrict__ wbv,
                                                      double* __restrict__ out) {
  __shared__ double lds[ANT / 64];
  double s = 0;
  for (int i = threadIdx.x; i < g.NY; i += ANT)
    s += wbu[(size_t)i * g.NX] - wbu[(size_t)i * g.NX + g.NX - 1];
  for (int i = threadIdx.x; i < g.NX; i += ANT)
    s += wbv[i] - wbv[(size_t)(g.NY - 1) * g.NX + i];
  s = block_sum<ANT>(s, lds);
  if (threadIdx.x == 0) *out = s;
}

namespace {
// the adjoint of the flux-corrected advDiff output, read through the
// projector: wb - e (e^T wb) / nb
__device__ inline double wbar_u(const Grid& g, const double* __restrict__ wbu, int x, int y,
                                double corr) {
  double a = wbu[(size_t)y * g.NX + x];
  if (x == 0) a -= corr;
  if (x == g.NX - 1) a += corr;
  return a;
}
__device__ inline double wbar_v(const Grid& g, const double* __restrict__ wbv, int x, int y,
                                double corr) {
  double a = wbv[(size_t)y * g.NX + x];
  if (y == 0) a -= corr;
  if (y == g.NY - 1) a += corr;
  return a;
}
}  // namespace

// advDiff (advDiff.cpp:14-79) linearized at (u, v) and transposed, as a
// gather: every forward output cell i that reads cell j contributes
// d out_i / d in_j times its adjoint.  Face ghosts are the boundary cell times
// the inflow fade (advDiff.cpp:50-55,68-71).  corr = (e^T wb) / nb.
__global__ __launch_bounds__(ANT) void k_adj_advdiff(Grid g, double dt, double UINF0,
                                                     double UINF1,
                                                     const double* __restrict__ u,
                                                     const double* __restrict__ v,
                                                     const double* __restrict__ wbu,
                                                     const double* __restrict__ wbv,
                                                     const double* __restrict__ fluxsum,
                                                     double* __restrict__ ub,
                                                     double* __restrict__ vb) {
  const int NX = g.NX, NY = g.NY;
  const double h = g.h;
  const double dfac = (g.nu / h) * (dt / h), afac = -0.5 * dt / h;
  const double norUinf = dmax(dmax(fabs(UINF0), fabs(UINF1)), EPSD);
  const double tW = dmax(UINF0, 0.0) / norUinf, tS = dmax(UINF1, 0.0) / norUinf;
  const double tE = dmin(UINF0, 0.0) / norUinf, tN = dmin(UINF1, 0.0) / norUinf;
  const double fadeW = 1 - tW * tW / BC_KILL, fadeS = 1 - tS * tS / BC_KILL;
  const double fadeE = 1 - tE * tE / BC_KILL, fadeN = 1 - tN * tN / BC_KILL;
  const double corr = *fluxsum / (2.0 * NY + 2.0 * NX);
  const size_t n = (size_t)NX * NY;
  for (size_t j = (size_t)blockIdx.x * ANT + threadIdx.x; j < n; j += (size_t)gridDim.x * ANT) {
    const int x = (int)(j % NX), y = (int)(j / NX);
    const bool hasW = x > 0, hasE = x + 1 < NX, hasS = y > 0, hasN = y + 1 < NY;
    const double uc = u[j], vc = v[j];
    // the forward neighbours of cell j (with ghosts)
    const double ue = hasE ? u[j + 1] : uc * fadeE, uw = hasW ? u[j - 1] : uc * fadeW;
    const double un = hasN ? u[j + NX] : uc * fadeN, us = hasS ? u[j - NX] : uc * fadeS;
    const double ve = hasE ? v[j + 1] : vc * fadeE, vw = hasW ? v[j - 1] : vc * fadeW;
    const double vn = hasN ? v[j + NX] : vc * fadeN, vs = hasS ? v[j - NX] : vc * fadeS;
    const double au = wbar_u(g, wbu, x, y, corr), av = wbar_v(g, wbv, x, y, corr);
    const double cu = uc + UINF0, cv = vc + UINF1;
    // own outputs: d u*_j / d u_j, d v*_j / d u_j, and the ghost terms
    double gu = au * (1 + afac * (ue - uw) - 4 * dfac) + av * afac * (ve - vw);
    double gv = av * (1 + afac * (vn - vs) - 4 * dfac) + au * afac * (un - us);
    if (!hasE) { gu += au * (afac * cu + dfac) * fadeE; gv += av * (afac * cu + dfac) * fadeE; }
    if (!hasW) { gu += au * (-afac * cu + dfac) * fadeW; gv += av * (-afac * cu + dfac) * fadeW; }
    if (!hasN) { gu += au * (afac * cv + dfac) * fadeN; gv += av * (afac * cv + dfac) * fadeN; }
    if (!hasS) { gu += au * (-afac * cv + dfac) * fadeS; gv += av * (-afac * cv + dfac) * fadeS; }
    // neighbours' outputs that read cell j as their E, W, N or S value
    if (hasW) {  // j is the E neighbour of j-1
      const double c = afac * (u[j - 1] + UINF0) + dfac;
      gu += wbar_u(g, wbu, x - 1, y, corr) * c;
      gv += wbar_v(g, wbv, x - 1, y, corr) * c;
    }
    if (hasE) {  // j is the W neighbour of j+1
      const double c = -afac * (u[j + 1] + UINF0) + dfac;
      gu += wbar_u(g, wbu, x + 1, y, corr) * c;
      gv += wbar_v(g, wbv, x + 1, y, corr) * c;
    }
    if (hasS) {  // j is the N neighbour of j-NX
      const double c = afac * (v[j - NX] + UINF1) + dfac;
      gu += wbar_u(g, wbu, x, y - 1, corr) * c;
      gv += wbar_v(g, wbv, x, y - 1, corr) * c;
    }
    if (hasN) {  // j is the S neighbour of j+NX
      const double c = -afac * (v[j + NX] + UINF1) + dfac;
      gu += wbar_u(g, wbu, x, y + 1, corr) * c;
      gv += wbar_v(g, wbv, x, y + 1, corr) * c;
    }
    ub[j] = gu;
    vb[j] = gv;
  }
}

// objective accumulation: J += drag of the shape / N
__global__ void k_adj_acc_drag(const cup2d_shape_state* __restrict__ ss, int shape, double w,
                               double* J) {
  if (threadIdx.x == 0) *J += w * ss[shape].drag;
}

__global__ __launch_bounds__(ANT) void k_adj_fill(double* a, size_t n, double val) {
  for (size_t j = (size_t)blockIdx.x * ANT + threadIdx.x; j < n; j += (size_t)gridDim.x * ANT)
    a[j] = val;
}

}  // namespace cup2d

namespace cup2d {

int enqueue_adjoint_fill(cup2d_ctx* c, double* a, size_t n, double val) {
  hipLaunchKernelGGL(k_adj_fill, dim3(1024), dim3(ANT), 0, c->stream, a, n, val);
  return launch_check(c, "adjoint fill");
}

int enqueue_adjoint_acc_drag(cup2d_ctx* c, int shape, double w, double* J) {
  hipLaunchKernelGGL(k_adj_acc_drag, dim3(1), dim3(64), 0, c->stream, c->ss, shape, w, J);
  return launch_check(c, "adjoint drag");
}

// One step of the adjoint, backwards: on entry (a.ub, a.vb) is the adjoint
// of the step's output velocity; on exit that of its input velocity, and the
// surface-velocity gradient a.theta has this step's share added.
int enqueue_adjoint_step(cup2d_ctx* c, const AdjStep& a) {
  const Grid& g = c->g;
  const size_t n = (size_t)g.NX * g.NY;
  const int G = c->ob.cap > 0 ? c->ob.cap : 1;
  const double facDiv = 0.5 * g.h / a.dt, pFac = -0.5 * a.dt / g.h;
  hipStream_t s = c->stream;
  hipLaunchKernelGGL(k_adj_fill, dim3(1024), dim3(ANT), 0, s, a.pb, n, 0.0);
  hipLaunchKernelGGL(k_adj_forces, dim3(G), dim3(ANT), 0, s, g, c->ob, a.shape, a.u1, a.v1, a.p1,
                     a.vu0, a.vu1, a.wJ, a.ub, a.vb, a.pb);
  hipLaunchKernelGGL(k_adj_penalize, dim3(G), dim3(ANT), 0, s, g, c->ob, a.shape, c->chi, c->ss,
                     surf_of(c), a.ub, a.vb, a.theta_part);
  if (c->surf_modes > 0)
    hipLaunchKernelGGL(k_adj_theta, dim3(4 * c->surf_modes), dim3(ANT), 0, s, a.theta_part, G,
                       4 * c->surf_modes, a.theta);
  hipLaunchKernelGGL(k_adj_project, dim3(2048), dim3(ANT), 0, s, g, pFac, a.ub, a.vb, a.pb);
  int rc = launch_check(c, "adjoint: forces, penalization, projection");
  if (rc) return rc;
  rc = enqueue_poisson_fft(c, a.pb, a.rb, 0);  // the solve is symmetric
  if (rc) return rc;
  if (c->band_nblocks > 0) {
    hipLaunchKernelGGL(k_adj_fade1, dim3(c->band_nblocks), dim3(ANT), 0, s, g, c->band_blocks,
                       c->band_nblocks, facDiv, a.wu, a.wv, a.rb, a.part);
    hipLaunchKernelGGL(k_adj_fade2, dim3(c->band_nblocks), dim3(ANT), 0, s, g, c->band_blocks,
                       c->band_nblocks, facDiv, a.wu, a.wv, a.part, a.rb);
  }
  hipLaunchKernelGGL(k_adj_obst1, dim3(G), dim3(ANT), 0, s, g, c->ob, a.shape, facDiv, c->chi,
                     a.wu, a.wv, a.rb, a.part);
  hipLaunchKernelGGL(k_adj_obst2, dim3(G), dim3(ANT), 0, s, g, c->ob, a.shape, facDiv, c->chi,
                     a.wu, a.wv, a.part, a.rb);
  hipLaunchKernelGGL(k_adj_div, dim3(2048), dim3(ANT), 0, s, g, facDiv, a.ub, a.vb, a.rb, a.wbu,
                     a.wbv);
  hipLaunchKernelGGL(k_adj_flux_sum, dim3(1), dim3(ANT), 0, s, g, a.wbu, a.wbv, a.fluxsum);
  hipLaunchKernelGGL(k_adj_advdiff, dim3(2048), dim3(ANT), 0, s, g, a.dt, a.uinf0, a.uinf1, a.u0,
                     a.v0, a.wbu, a.wbv, a.fluxsum, a.ub, a.vb);
  return launch_check(c, "adjoint: fade, obstacle RHS, divergence, advDiff");
}

}  // namespace cup2d
