// fish.hip -- swimmers (CarlingFish) on the device: midline kinematics and
// internal-frame moments (Fish::create steps 1-5, Fish.cpp:60-150) and the
// per-block shape construction PutFishOnBlocks (FishLibrary.cpp:403-638).
//
// Arithmetic follows the reference expression by expression (-ffp-contract=off);
// trigonometry uses det_sincos and the midline integrals use a fixed
// 256-lane summation order, so a CPU restatement that does the same produces
// identical bits.  PutFishOnBlocks is a scatter in the reference (each
// midline point writes into the cells around it); here every cell gathers the
// points in the reference's loop order, which gives the same minimum-distance
// winner and the same sums without atomics.
#include <cmath>
#include <vector>

#include "device_util.hpp"
#include "fish.hpp"
#include "internal.hpp"

namespace cup2d {
namespace {

constexpr int NT = 256;
constexpr double EPSD = 2.220446049250313e-16;  // std::numeric_limits<double>::epsilon()

// AmplitudeFish::midlineLatPos / midlineLatVel (CarlingFish.cpp:300-308), or
// the envelope L (a0 + a2 (s/L)^2) with wavelength waveLength*L.
__device__ inline double lat_pos(const FishDev& F, double s, double t) {
  if (F.envelope == 0) {
    const double arg = 2 * M_PI * (s / F.length - t / F.Tperiod + F.phaseShift);
    return 4. / 33. * (s + 0.03125 * F.length) * det_sin(arg);
  }
  const double x = s / F.length;
  const double arg = 2 * M_PI * (x / F.waveLength - t / F.Tperiod + F.phaseShift);
  return F.length * (F.env_a0 + F.env_a2 * x * x) * det_sin(arg);
}
__device__ inline double lat_vel(const FishDev& F, double s, double t) {
  if (F.envelope == 0) {
    const double arg = 2 * M_PI * (s / F.length - t / F.Tperiod + F.phaseShift);
    return -4. / 33. * (s + 0.03125 * F.length) * (2 * M_PI / F.Tperiod) * det_cos(arg);
  }
  const double x = s / F.length;
  const double arg = 2 * M_PI * (x / F.waveLength - t / F.Tperiod + F.phaseShift);
  return -F.length * (F.env_a0 + F.env_a2 * x * x) * (2 * M_PI / F.Tperiod) * det_cos(arg);
}

// FishData::_d_ds (FishLibrary.h:102-107)
__device__ inline double d_ds(const double* rS, int Nm, int idx, const double* vals) {
  if (idx == 0) return (vals[idx + 1] - vals[idx]) / (rS[idx + 1] - rS[idx]);
  if (idx == Nm - 1) return (vals[idx] - vals[idx - 1]) / (rS[idx] - rS[idx - 1]);
  return ((vals[idx + 1] - vals[idx]) / (rS[idx + 1] - rS[idx]) +
          (vals[idx] - vals[idx - 1]) / (rS[idx] - rS[idx - 1])) / 2;
}

struct Mid {
  double *rX, *rY, *vX, *vY, *norX, *norY, *vNorX, *vNorY;
};
__device__ inline Mid mid_of(const FishDev& F) {
  const int n = F.Nm;
  return Mid{F.mid, F.mid + n, F.mid + 2 * n, F.mid + 3 * n,
             F.mid + 4 * n, F.mid + 5 * n, F.mid + 6 * n, F.mid + 7 * n};
}

// trapezoid weight and _integrationFac1..3 (FishLibrary.h:108-118, w^3 as w*w*w)
__device__ inline void facs(const FishDev& F, const Mid& M, int i, double* ds, double* fac1,
                            double* fac2, double* fac3) {
  const int Nm = F.Nm;
  const double* rS = F.rS;
  *ds = (i == 0) ? rS[1] - rS[0] : ((i == Nm - 1) ? rS[Nm - 1] - rS[Nm - 2] : rS[i + 1] - rS[i - 1]);
  const double w = F.width[i], w3 = w * w * w;
  const double dnorXi = d_ds(rS, Nm, i, M.norX), dnorYi = d_ds(rS, Nm, i, M.norY);
  *fac1 = 2 * w;
  *fac2 = 2 * w3 * (dnorXi * M.norY[i] - dnorYi * M.norX[i]) / 3;
  *fac3 = 2 * w3 / 3;
}

// pairwise combination of NQ rows of NT lane partials: lane l += lane l + 128,
// 64, ..., 1 (the order of the CPU restatement's lane sum)
template <int NQ>
__device__ inline void lane_tree(double* red) {
  const int t = threadIdx.x;
  for (int st = NT / 2; st >= 1; st >>= 1) {
    __syncthreads();
    if (t < st)
#pragma unroll
      for (int q = 0; q < NQ; ++q) red[q * NT + t] = red[q * NT + t] + red[q * NT + t + st];
  }
  __syncthreads();
}

}  // namespace

// One workgroup per swimmer, after the bodies have moved:
// Fish::updatePosition's theta_internal update (Fish.cpp:220-225),
// AmplitudeFish::computeMidline (CarlingFish.cpp:332-369),
// integrateLinearMomentum, changeToCoMFrameLinear, integrateAngularMomentum,
// changeToCoMFrameAngular (FishLibrary.cpp:86-176) and the AreaSegment boxes
// in the computational frame (Fish.cpp:118-150, FishLibrary.cpp:295-342).
__global__ __launch_bounds__(NT) void k_fish_midline(const FishDev* __restrict__ fd,
                                                     const cup2d_shape_state* __restrict__ ss,
                                                     const StepState* __restrict__ st,
                                                     double h) {
  __shared__ double red[5 * NT];
  __shared__ double scan[2][FISH_MAX_NM];
  const FishDev F = fd[blockIdx.x];
  const int Nm = F.Nm, t = threadIdx.x;
  const Mid M = mid_of(F);
  const double* rS = F.rS;
  const double time = st->time, dt = st->dt;
  const double theta = F.st[6] - dt * F.st[7];
  const double T = F.Tperiod;
  const double rampFac = time < T ? det_sin(0.5 * M_PI * time / T) : 1.0;
  const double rampFacVel = time < T ? 0.5 * M_PI / T * det_cos(0.5 * M_PI * time / T) : 0.0;
  // 1. midline
  for (int i = t; i < Nm; i += NT) {
    const double lp = lat_pos(F, rS[i], time);
    M.rY[i] = rampFac * lp;
    M.vY[i] = rampFac * lat_vel(F, rS[i], time) + rampFacVel * lp;
  }
  __syncthreads();
  for (int i = 1 + t; i < Nm; i += NT) {
    const double dy = M.rY[i] - M.rY[i - 1], ds = rS[i] - rS[i - 1];
    const double dx = sqrt(ds * ds - dy * dy);
    const double dVy = M.vY[i] - M.vY[i - 1];
    const double dVx = -dy / dx * dVy;
    scan[0][i] = dx;
    scan[1][i] = dVx;
    M.norX[i - 1] = -dy / ds;
    M.norY[i - 1] = dx / ds;
    M.vNorX[i - 1] = -dVy / ds;
    M.vNorY[i - 1] = dVx / ds;
  }
  __syncthreads();
  // running sums rX[i] += rX[i-1] in index order, one lane per array
  if (t == 0 || t == 64) {
    const int a = t == 0 ? 0 : 1;
    double* out = a == 0 ? M.rX : M.vX;
    double acc = 0;
    out[0] = 0;
    for (int i = 1; i < Nm; ++i) {
      acc = scan[a][i] + acc;
      out[i] = acc;
    }
  }
  if (t == 128) {
    M.norX[Nm - 1] = M.norX[Nm - 2];
    M.norY[Nm - 1] = M.norY[Nm - 2];
    M.vNorX[Nm - 1] = M.vNorX[Nm - 2];
    M.vNorY[Nm - 1] = M.vNorY[Nm - 2];
  }
  __syncthreads();
  // 2. integrateLinearMomentum
  {
    double a[5] = {0, 0, 0, 0, 0};
    for (int i = t; i < Nm; i += NT) {
      double ds, fac1, fac2, fac3;
      facs(F, M, i, &ds, &fac1, &fac2, &fac3);
      a[0] += fac1 * ds / 2;
      a[1] += (M.rX[i] * fac1 + M.norX[i] * fac2) * ds / 2;
      a[2] += (M.rY[i] * fac1 + M.norY[i] * fac2) * ds / 2;
      a[3] += (M.vX[i] * fac1 + M.vNorX[i] * fac2) * ds / 2;
      a[4] += (M.vY[i] * fac1 + M.vNorY[i] * fac2) * ds / 2;
    }
#pragma unroll
    for (int q = 0; q < 5; ++q) red[q * NT + t] = a[q];
  }
  lane_tree<5>(red);
  const double area = red[0];
  const double C0 = red[NT] / area, C1 = red[2 * NT] / area;
  const double V0 = red[3 * NT] / area, V1 = red[4 * NT] / area;
  __syncthreads();
  // 3. changeToCoMFrameLinear
  for (int i = t; i < Nm; i += NT) {
    M.rX[i] -= C0; M.rY[i] -= C1;
    M.vX[i] -= V0; M.vY[i] -= V1;
  }
  __syncthreads();
  // integrateAngularMomentum
  {
    double am = 0, jj = 0;
    for (int i = t; i < Nm; i += NT) {
      double ds, fac1, fac2, fac3;
      facs(F, M, i, &ds, &fac1, &fac2, &fac3);
      const double rX = M.rX[i], rY = M.rY[i], vX = M.vX[i], vY = M.vY[i];
      const double norX = M.norX[i], norY = M.norY[i], vNorX = M.vNorX[i], vNorY = M.vNorY[i];
      const double tmp_M = (rX * vY - rY * vX) * fac1 +
                           (rX * vNorY - rY * vNorX + vY * norX - vX * norY) * fac2 +
                           (norX * vNorY - norY * vNorX) * fac3;
      const double tmp_J = (rX * rX + rY * rY) * fac1 + 2 * (rX * norX + rY * norY) * fac2 + fac3;
      am += tmp_M * ds / 2;
      jj += tmp_J * ds / 2;
    }
    red[t] = am;
    red[NT + t] = jj;
  }
  lane_tree<2>(red);
  const double J = red[NT];
  const double angvel = red[0] / J;
  __syncthreads();
  // changeToCoMFrameAngular(theta_internal, angvel_internal)
  {
    double sn, cs;
    det_sincos(theta, &sn, &cs);
    const double R00 = cs, R01 = -sn, R10 = sn, R11 = cs;
    for (int i = t; i < Nm; i += NT) {
      double vX = M.vX[i] + angvel * M.rY[i];
      double vY = M.vY[i] - angvel * M.rX[i];
      const double px = M.rX[i], py = M.rY[i];
      M.rX[i] = R00 * px + R01 * py;
      M.rY[i] = R10 * px + R11 * py;
      M.vX[i] = R00 * vX + R01 * vY;
      M.vY[i] = R10 * vX + R11 * vY;
    }
  }
  __syncthreads();
  // _computeMidlineNormals (FishLibrary.cpp:66-84)
  for (int i = t; i < Nm - 1; i += NT) {
    const double ds = rS[i + 1] - rS[i];
    const double tX = M.rX[i + 1] - M.rX[i], tY = M.rY[i + 1] - M.rY[i];
    const double tVX = M.vX[i + 1] - M.vX[i], tVY = M.vY[i + 1] - M.vY[i];
    M.norX[i] = -tY / ds;
    M.norY[i] = tX / ds;
    M.vNorX[i] = -tVY / ds;
    M.vNorY[i] = tVX / ds;
  }
  __syncthreads();
  if (t == 0) {
    M.norX[Nm - 1] = M.norX[Nm - 2];
    M.norY[Nm - 1] = M.norY[Nm - 2];
    M.vNorX[Nm - 1] = M.vNorX[Nm - 2];
    M.vNorY[Nm - 1] = M.vNorY[Nm - 2];
  }
  __syncthreads();
  // 4-5. AreaSegments in the computational frame of (center, orientation)
  const cup2d_shape_state& S = ss[F.shape];
  double sn, cs;
  det_sincos(S.orientation, &sn, &cs);
  const double R[2][2] = {{cs, -sn}, {sn, cs}};
  double ub[4] = {1e300, -1e300, 1e300, -1e300};
  for (int i = t; i < F.Nseg; i += NT) {
    const int next_idx = (i + 1) * (Nm - 1) / F.Nseg, idx = i * (Nm - 1) / F.Nseg;
    double bbox[2][2] = {{1e9, -1e9}, {1e9, -1e9}};
    for (int s = idx; s <= next_idx; ++s) {
      const double w = F.width[s];
      const double xB0 = M.rX[s] - M.norX[s] * w, xB1 = M.rX[s] + M.norX[s] * w;
      const double yB0 = M.rY[s] - M.norY[s] * w, yB1 = M.rY[s] + M.norY[s] * w;
      const double maxX = dmax(xB0, xB1), minX = dmin(xB0, xB1);
      const double maxY = dmax(yB0, yB1), minY = dmin(yB0, yB1);
      bbox[0][0] = dmin(bbox[0][0], minX);
      bbox[0][1] = dmax(bbox[0][1], maxX);
      bbox[1][0] = dmin(bbox[1][0], minY);
      bbox[1][1] = dmax(bbox[1][1], maxY);
    }
    const double safe = 2 * h;
    const double w[2] = {(bbox[0][1] - bbox[0][0]) / 2 + safe, (bbox[1][1] - bbox[1][0]) / 2 + safe};
    const double p[2] = {(bbox[0][1] + bbox[0][0]) / 2, (bbox[1][1] + bbox[1][0]) / 2};
    const double nx[2] = {1, 0}, ny[2] = {0, 1};
    double c[2], nI[2], nJ[2];
    for (int d = 0; d < 2; ++d) {
      c[d] = R[d][0] * p[0] + R[d][1] * p[1];
      nI[d] = R[d][0] * nx[0] + R[d][1] * nx[1];
      nJ[d] = R[d][0] * ny[0] + R[d][1] * ny[1];
    }
    c[0] += S.center[0];
    c[1] += S.center[1];
    const double magI = sqrt(nI[0] * nI[0] + nI[1] * nI[1]);
    const double magJ = sqrt(nJ[0] * nJ[0] + nJ[1] * nJ[1]);
    const double invMagI = 1 / magI, invMagJ = 1 / magJ;
    for (int d = 0; d < 2; ++d) {
      nI[d] = fabs(nI[d]) * invMagI;
      nJ[d] = fabs(nJ[d]) * invMagJ;
    }
    const double wX[2] = {w[0] * nI[0], w[0] * nI[1]};
    const double wY[2] = {w[1] * nJ[0], w[1] * nJ[1]};
    double* out = F.seg + (size_t)i * FISH_SEGD;
    out[SG_SAFE] = safe;
    for (int d = 0; d < 2; ++d) {
      out[SG_LAB + 2 * d] = c[d] - wX[d] - wY[d];
      out[SG_LAB + 2 * d + 1] = c[d] + wX[d] + wY[d];
      out[SG_OBJ + 2 * d] = c[d] - w[d];
      out[SG_OBJ + 2 * d + 1] = c[d] + w[d];
      out[SG_NI + d] = nI[d];
      out[SG_NJ + d] = nJ[d];
    }
    ub[0] = dmin(ub[0], out[SG_LAB]);
    ub[1] = dmax(ub[1], out[SG_LAB + 1]);
    ub[2] = dmin(ub[2], out[SG_LAB + 2]);
    ub[3] = dmax(ub[3], out[SG_LAB + 3]);
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) red[q * NT + t] = ub[q];
  for (int s = NT / 2; s >= 1; s >>= 1) {
    __syncthreads();
    if (t < s) {
      red[t] = dmin(red[t], red[t + s]);
      red[NT + t] = dmax(red[NT + t], red[NT + t + s]);
      red[2 * NT + t] = dmin(red[2 * NT + t], red[2 * NT + t + s]);
      red[3 * NT + t] = dmax(red[3 * NT + t], red[3 * NT + t + s]);
    }
  }
  __syncthreads();
  if (t == 0) {
    for (int q = 0; q < 4; ++q) F.box[q] = red[q * NT];
    F.st[0] = area; F.st[1] = J;
    F.st[2] = C0; F.st[3] = C1;
    F.st[4] = V0; F.st[5] = V1;
    F.st[6] = theta; F.st[7] = angvel;
  }
}

namespace {

// frame of PutFishOnBlocks (FishLibrary.h:170-204)
struct Frame {
  double pos0, pos1, R00, R01, R10, R11;
  __device__ inline void to_comp(double x[2]) const {  // changeToComputationalFrame
    const double p0 = x[0], p1 = x[1];
    x[0] = R00 * p0 + R01 * p1;
    x[1] = R10 * p0 + R11 * p1;
    x[0] += pos0;
    x[1] += pos1;
  }
  __device__ inline void vel_to_comp(double x[2]) const {  // changeVelocityToComputationalFrame
    const double p0 = x[0], p1 = x[1];
    x[0] = R00 * p0 + R01 * p1;
    x[1] = R10 * p0 + R11 * p1;
  }
  __device__ inline void from_comp(double x[2]) const {  // changeFromComputationalFrame
    const double p0 = x[0] - pos0, p1 = x[1] - pos1;
    x[0] = R00 * p0 + R10 * p1;
    x[1] = R01 * p0 + R11 * p1;
  }
};

__device__ inline double dist_sq(const double a[2], const double b[2]) {
  return (a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]);
}

// surface point (ss, signp) of constructSurface (FishLibrary.cpp:470-486)
struct SurfPt {
  double myP[2], pP[2], pM[2], ud[2];
  int iap[2];
  int valid;
};
__device__ inline void surf_point(const FishDev& F, const Mid& M, const Frame& Fr, int ss,
                                  int signp, double org0, double org1, double invh, SurfPt& P) {
  const double* width = F.width;
  P.myP[0] = M.rX[ss + 0] + width[ss + 0] * signp * M.norX[ss + 0];
  P.myP[1] = M.rY[ss + 0] + width[ss + 0] * signp * M.norY[ss + 0];
  Fr.to_comp(P.myP);
  P.iap[0] = (int)floor((P.myP[0] - org0) * invh);
  P.iap[1] = (int)floor((P.myP[1] - org1) * invh);
  P.pP[0] = M.rX[ss + 1] + width[ss + 1] * signp * M.norX[ss + 1];
  P.pP[1] = M.rY[ss + 1] + width[ss + 1] * signp * M.norY[ss + 1];
  Fr.to_comp(P.pP);
  P.pM[0] = M.rX[ss - 1] + width[ss - 1] * signp * M.norX[ss - 1];
  P.pM[1] = M.rY[ss - 1] + width[ss - 1] * signp * M.norY[ss - 1];
  Fr.to_comp(P.pM);
  P.ud[0] = M.vX[ss + 0] + width[ss + 0] * signp * M.vNorX[ss + 0];
  P.ud[1] = M.vY[ss + 0] + width[ss + 0] * signp * M.vNorY[ss + 0];
  Fr.vel_to_comp(P.ud);
}

constexpr int PTS_PER_SEG = 18;               // 9 midline points x 2 sides
constexpr int SEG_PER_CHUNK = NT / PTS_PER_SEG;  // 14 segments of surface points
constexpr int ENT_PER_SEG = 9;
constexpr int SEG_PER_ICHUNK = NT / ENT_PER_SEG;  // 28 segments of internal slices
constexpr int FISH_MAX_SEG = FISH_MAX_NM / 8;

}  // namespace

// PutFishOnBlocks::operator() (FishLibrary.cpp:403-417) for every obstacle
// slot that belongs to a swimmer: one workgroup per slot, four cells of one
// column per thread.
//   constructSurface (:438-573): the cell's signed squared distance, chi weight
//     W and W*udef come from the first surface point, in loop order, with the
//     smallest squared distance among those whose 4x4 support covers the cell;
//   constructInternl (:575-638): hat-kernel P2M of the internal slices, summed
//     in loop order;
//   signedDistanceSqrt (:419-436): udef /= chi, signed square root.
// Writes the slot's dist, udef and rho (chi is set by the Towers pass).
__global__ __launch_bounds__(NT) void k_fish_place(Grid g, const ShapeParams* __restrict__ sp,
                                                   const cup2d_shape_state* __restrict__ ss,
                                                   const FishDev* __restrict__ fd,
                                                   Obstacles ob) {
  __shared__ int sl[FISH_MAX_SEG];
  __shared__ int s_tmp[NT];
  __shared__ SurfPt pts[NT];
  __shared__ double ent[NT][5];  // internal slice: m0 m1 n0 n1 (index units), Nw
  __shared__ int ent_ss[NT];
  const int s = blockIdx.x;
  if (s >= ob.cnt[0]) return;
  const int k = ob.slot_shape[s];
  const int f = sp[k].fish;
  if (f < 0) return;
  const FishDev F = fd[f];
  const Mid M = mid_of(F);
  const int Nm = F.Nm, t = threadIdx.x;
  const int b = ob.slot_block[s], bx = b % g.bpdx, by = b / g.bpdx;
  const double h = g.h, invh = 1.0 / h;
  const double org0 = cell_pos(bx, 0, g.hb, h), org1 = cell_pos(by, 0, g.hb, h);
  // segments meeting this block, in segment order (Fish.cpp:158-173)
  int nsl;
  {
    double start[2], end[2];
    block_corners(g, bx, by, start, end);
    int cnt = 0;
    const int per = (F.Nseg + NT - 1) / NT;  // consecutive segments per thread
    const int i0 = t * per, i1 = i0 + per < F.Nseg ? i0 + per : F.Nseg;
    for (int i = i0; i < i1; ++i) cnt += seg_intersects(F.seg + (size_t)i * FISH_SEGD, start, end);
    s_tmp[t] = cnt;
    __syncthreads();
    for (int off = 1; off < NT; off <<= 1) {
      const int v = t >= off ? s_tmp[t - off] : 0;
      __syncthreads();
      s_tmp[t] += v;
      __syncthreads();
    }
    int pos = t == 0 ? 0 : s_tmp[t - 1];
    nsl = s_tmp[NT - 1];
    for (int i = i0; i < i1; ++i)
      if (seg_intersects(F.seg + (size_t)i * FISH_SEGD, start, end)) sl[pos++] = i;
    __syncthreads();
  }
  const cup2d_shape_state& S = ss[k];
  Frame Fr;
  {
    double sn, cs;
    det_sincos(S.orientation, &sn, &cs);
    Fr.pos0 = S.center[0];
    Fr.pos1 = S.center[1];
    Fr.R00 = cs; Fr.R01 = -sn; Fr.R10 = sn; Fr.R11 = cs;
  }
  const int cx = t % BS, cy0 = (t / BS) * 4;
  double px[4][2];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    px[c][0] = cell_pos(bx, cx, g.hb, h);
    px[c][1] = cell_pos(by, cy0 + c, g.hb, h);
  }
  // ---- constructSurface: first point with the smallest squared distance
  double best[4] = {1, 1, 1, 1};  // |dist| starts at 1 (dist = -1)
  int win[4] = {-1, -1, -1, -1};
  for (int j0 = 0; j0 < nsl; j0 += SEG_PER_CHUNK) {
    const int nj = nsl - j0 < SEG_PER_CHUNK ? nsl - j0 : SEG_PER_CHUNK;
    if (t < nj * PTS_PER_SEG) {
      const int j = j0 + t / PTS_PER_SEG, r = t % PTS_PER_SEG;
      const int s0 = 8 * sl[j], s1 = 8 * sl[j] + 8;
      const int first = s0 > 1 ? s0 : 1, last = s1 < Nm - 2 ? s1 : Nm - 2;
      const int sidx = first + r / 2, signp = (r & 1) ? 1 : -1;
      SurfPt P;
      if (sidx <= last) {
        surf_point(F, M, Fr, sidx, signp, org0, org1, invh, P);
        P.valid = 1;
      } else {
        P.valid = 0;
      }
      pts[t] = P;
    }
    __syncthreads();
    const int np = nj * PTS_PER_SEG;
    for (int p = 0; p < np; ++p) {
      if (!pts[p].valid) continue;
      const int a0 = pts[p].iap[0], a1 = pts[p].iap[1];
      if (cx < a0 - 1 || cx > a0 + 2 || cy0 + 3 < a1 - 1 || cy0 > a1 + 2) continue;
      const double m0 = pts[p].myP[0], m1 = pts[p].myP[1];
      const double P0 = pts[p].pP[0], P1 = pts[p].pP[1], Q0 = pts[p].pM[0], Q1 = pts[p].pM[1];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int cy = cy0 + c;
        if (cy < a1 - 1 || cy > a1 + 2) continue;
        const double my[2] = {m0, m1}, pp[2] = {P0, P1}, pm[2] = {Q0, Q1};
        const double dist0 = dist_sq(px[c], my), distP = dist_sq(px[c], pp), distM = dist_sq(px[c], pm);
        const double d1 = dmin(dmin(dist0, distP), distM);
        if (d1 < best[c]) {
          best[c] = d1;
          win[c] = (j0 * PTS_PER_SEG) + p;
        }
      }
    }
    __syncthreads();
  }
  double dist[4], chi[4], udx[4], udy[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    dist[c] = -1; chi[c] = 0; udx[c] = 0; udy[c] = 0;
    if (win[c] < 0) continue;
    const int j = win[c] / PTS_PER_SEG, r = win[c] % PTS_PER_SEG;
    const int s0 = 8 * sl[j];
    const int first = s0 > 1 ? s0 : 1;
    const int sidx = first + r / 2, signp = (r & 1) ? 1 : -1;
    SurfPt P;
    surf_point(F, M, Fr, sidx, signp, org0, org1, invh, P);
    double p[2] = {px[c][0], px[c][1]};
    const double dist0 = dist_sq(p, P.myP), distP = dist_sq(p, P.pP), distM = dist_sq(p, P.pM);
    Fr.from_comp(p);
    int close_s = sidx, secnd_s = sidx + (distP < distM ? 1 : -1);
    double dist1 = dist0;
    if (distP < dist0 || distM < dist0) {
      dist1 = distP < distM ? distP : distM;
      close_s = secnd_s;
      secnd_s = sidx;
    }
    const double* rX = M.rX;
    const double* rY = M.rY;
    const double* width = F.width;
    const double dSsq = (rX[close_s] - rX[secnd_s]) * (rX[close_s] - rX[secnd_s]) +
                        (rY[close_s] - rY[secnd_s]) * (rY[close_s] - rY[secnd_s]);
    const double cnt2ML = width[close_s] * width[close_s];
    const double nxt2ML = width[secnd_s] * width[secnd_s];
    const double safeW = dmax(width[close_s], width[secnd_s]) + 2 * h;
    const double xMidl[2] = {rX[close_s], rY[close_s]};
    const double grd2ML = dist_sq(p, xMidl);
    const double diffH = fabs(width[close_s] - width[secnd_s]);
    double sign2d;
    if (dSsq > diffH * diffH || grd2ML > safeW * safeW) {
      sign2d = grd2ML > cnt2ML ? -1 : 1;
    } else {
      const double corr = 2 * sqrt(cnt2ML * nxt2ML);
      const double Rsq = (cnt2ML + nxt2ML - corr + dSsq) * (cnt2ML + nxt2ML + corr + dSsq) / 4 / dSsq;
      const double maxAx = dmax(cnt2ML, nxt2ML);
      const int idAx1 = cnt2ML > nxt2ML ? close_s : secnd_s;
      const int idAx2 = idAx1 == close_s ? secnd_s : close_s;
      const double d = sqrt((Rsq - maxAx) / dSsq);
      const double xCentr[2] = {rX[idAx1] + (rX[idAx1] - rX[idAx2]) * d,
                                rY[idAx1] + (rY[idAx1] - rY[idAx2]) * d};
      const double grd2Core = dist_sq(p, xCentr);
      sign2d = grd2Core > Rsq ? -1 : 1;
    }
    const double W = 1 - sqrt(dist1) * (invh / 3);
    udx[c] = W * P.ud[0];
    udy[c] = W * P.ud[1];
    dist[c] = sign2d * dist1;
    chi[c] = W;
  }
  // ---- constructInternl: hat P2M of the slices at offsets iw*h
  for (int j0 = 0; j0 < nsl; j0 += SEG_PER_ICHUNK) {
    const int nj = nsl - j0 < SEG_PER_ICHUNK ? nsl - j0 : SEG_PER_ICHUNK;
    __syncthreads();
    if (t < nj * ENT_PER_SEG) {
      const int j = j0 + t / ENT_PER_SEG, r = t % ENT_PER_SEG;
      const int s0 = 8 * sl[j], s1 = 8 * sl[j] + 8;
      const int first = s0 > 1 ? s0 : 1, last = s1 < Nm - 2 ? s1 : Nm - 2;
      const int sidx = first + r;
      if (sidx <= last) {
        double xp[2] = {M.rX[sidx], M.rY[sidx]};
        Fr.to_comp(xp);
        ent[t][0] = (xp[0] - org0) * invh;
        ent[t][1] = (xp[1] - org1) * invh;
        ent[t][2] = Fr.R00 * M.norX[sidx] + Fr.R01 * M.norY[sidx];
        ent[t][3] = Fr.R10 * M.norX[sidx] + Fr.R11 * M.norY[sidx];
        ent[t][4] = floor(F.width[sidx] / h);
        ent_ss[t] = sidx;
      } else {
        ent_ss[t] = -1;
      }
    }
    __syncthreads();
    const int ne = nj * ENT_PER_SEG;
    for (int e = 0; e < ne; ++e) {
      const int sidx = ent_ss[e];
      if (sidx < 0) continue;
      const double m0 = ent[e][0], m1 = ent[e][1], n0 = ent[e][2], n1 = ent[e][3];
      const int Nw = (int)ent[e][4];
      const double d0 = cx - m0;
      // cells within sqrt(2) of the slice line, in index units
      if (fabs((cy0 - m1) * n0 - d0 * n1) > 1.5 && fabs((cy0 + 3 - m1) * n0 - d0 * n1) > 1.5 &&
          ((cy0 - m1) * n0 - d0 * n1) * ((cy0 + 3 - m1) * n0 - d0 * n1) > 0)
        continue;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int cy = cy0 + c;
        const double d1 = cy - m1;
        const double perp = d1 * n0 - d0 * n1;
        if (fabs(perp) > 1.5) continue;
        const double tp = d0 * n0 + d1 * n1;
        int lo = (int)ceil(tp - 1.5), hi = (int)floor(tp + 1.5);
        lo = lo > -Nw + 1 ? lo : -Nw + 1;
        hi = hi < Nw - 1 ? hi : Nw - 1;
        for (int iw = lo; iw <= hi; ++iw) {
          const double offsetW = iw * h;
          double xp[2] = {M.rX[sidx] + offsetW * M.norX[sidx], M.rY[sidx] + offsetW * M.norY[sidx]};
          Fr.to_comp(xp);
          xp[0] = (xp[0] - org0) * invh;
          xp[1] = (xp[1] - org1) * invh;
          const double ap0 = floor(xp[0]), ap1 = floor(xp[1]);
          const int iap0 = (int)ap0, iap1 = (int)ap1;
          const int sx = cx - iap0, sy = cy - iap1;
          if (sx < 0 || sx > 1 || sy < 0 || sy > 1) continue;
          double ud[2] = {M.vX[sidx] + offsetW * M.vNorX[sidx], M.vY[sidx] + offsetW * M.vNorY[sidx]};
          Fr.vel_to_comp(ud);
          const double wx = 1 - fabs(xp[0] - (sx == 0 ? ap0 : ap0 + 1));
          const double wy = 1 - fabs(xp[1] - (sy == 0 ? ap1 : ap1 + 1));
          const double wxwy = wy * wx;
          udx[c] += wxwy * ud[0];
          udy[c] += wxwy * ud[1];
          chi[c] += wxwy;
          if (fabs(dist[c] + 1) < EPSD) dist[c] = 1;
        }
      }
    }
  }
  // ---- signedDistanceSqrt
  const size_t base = (size_t)s * BS * BS;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int i = (cy0 + c) * BS + cx;
    const double normfac = chi[c] > EPSD ? chi[c] : 1;
    ob.udef[base * 2 + i] = udx[c] / normfac;
    ob.udef[base * 2 + BS * BS + i] = udy[c] / normfac;
    ob.dist[base + i] = dist[c] >= 0 ? sqrt(dist[c]) : -sqrt(-dist[c]);
    ob.rho[base + i] = 1;
    ob.chi[base + i] = 0;
  }
}

// ------------------------------------------------------------------- host
// FishData::FishData (FishLibrary.cpp:14-38, FishLibrary.h:44-60) and the
// half-width table (AmplitudeFish::_width, CarlingFish.cpp:322-329, or the
// NACA 4-digit half-thickness 5 t L (0.2969 sqrt(x) - 0.1260 x - 0.3516 x^2 +
// 0.2843 x^3 - 0.1036 x^4)).
int fish_discretize(const cup2d_shape_desc* d, double h, std::vector<double>& rS,
                    std::vector<double>& width) {
  const double length = d->length;
  const double fracRefined = 0.1, fracMid = 1 - 2 * fracRefined;
  const double dSmid_tgt = h / std::sqrt(2);
  const double dSrefine_tgt = 0.125 * h;
  const int Nmid = (int)std::ceil(length * fracMid / dSmid_tgt / 8) * 8;
  const double dSmid = length * fracMid / Nmid;
  const int Nend = (int)std::ceil(fracRefined * length * 2 / (dSmid + dSrefine_tgt) / 4) * 4;
  const double dSref = fracRefined * length * 2 / Nend - dSmid;
  const int Nm = Nmid + 2 * Nend + 1;
  if (!(dSref > 0) || Nend < 2 || Nm < 17 || Nm > FISH_MAX_NM) return -1;
  rS.assign(Nm, 0.0);
  width.assign(Nm, 0.0);
  rS[0] = 0;
  int k = 0;
  for (int i = 0; i < Nend; ++i, k++)
    rS[k + 1] = rS[k] + dSref + (dSmid - dSref) * i / ((double)Nend - 1.);
  for (int i = 0; i < Nmid; ++i, k++) rS[k + 1] = rS[k] + dSmid;
  for (int i = 0; i < Nend; ++i, k++)
    rS[k + 1] = rS[k] + dSref + (dSmid - dSref) * (Nend - i - 1) / ((double)Nend - 1.);
  rS[k] = dmin(rS[k], length);
  const double L = length;
  for (int i = 0; i < Nm; ++i) {
    const double s = rS[i];
    double w = 0;
    if (s < 0 || s > L) {
      w = 0;
    } else if (d->profile == 0) {
      const double sb = .04 * length, st = .95 * length, wt = .01 * length, wh = .04 * length;
      w = s < sb ? std::sqrt(2 * wh * s - s * s)
                 : (s < st ? wh - (wh - wt) * ((s - sb) / (st - sb)) : (wt * (L - s) / (L - st)));
    } else {
      const double x = s / L;
      w = 5 * d->naca_t * L *
          (0.2969 * std::sqrt(x) - 0.1260 * x - 0.3516 * x * x + 0.2843 * x * x * x -
           0.1036 * x * x * x * x);
      w = w < 0 ? 0 : w;
    }
    width[i] = w;
  }
  return Nm;
}

int fish_add(cup2d_ctx* c, const cup2d_shape_desc* d, int shape) {
  if ((int)c->hfish.size() >= MAX_FISH) {
    c->err = "too many swimmers";
    return CUP2D_ERR_ARG;
  }
  std::vector<double> rS, width;
  const int Nm = fish_discretize(d, c->g.h, rS, width);
  if (Nm < 0) {
    c->err = "swimmer discretisation failed (length too small or too large for the grid)";
    return CUP2D_ERR_ARG;
  }
  FishDev F{};
  F.shape = shape;
  F.Nm = Nm;
  F.Nseg = (Nm - 1) / 8;
  F.envelope = d->envelope;
  F.length = d->length;
  F.Tperiod = d->Tperiod;
  F.phaseShift = d->phaseShift;
  F.env_a0 = d->env_a0;
  F.env_a2 = d->env_a2;
  F.waveLength = d->waveLength;
  const size_t nd = (size_t)Nm * 10 + (size_t)F.Nseg * FISH_SEGD + 4 + 8;
  double* buf = nullptr;
  CUP2D_HIP_CHECK(c, hipMalloc(&buf, sizeof(double) * nd));
  CUP2D_HIP_CHECK(c, hipMemset(buf, 0, sizeof(double) * nd));
  F.rS = buf;
  F.width = buf + Nm;
  F.mid = buf + 2 * (size_t)Nm;
  F.seg = buf + 10 * (size_t)Nm;
  F.box = F.seg + (size_t)F.Nseg * FISH_SEGD;
  F.st = F.box + 4;
  CUP2D_HIP_CHECK(c, hipMemcpy(F.rS, rS.data(), sizeof(double) * Nm, hipMemcpyHostToDevice));
  CUP2D_HIP_CHECK(c, hipMemcpy(F.width, width.data(), sizeof(double) * Nm, hipMemcpyHostToDevice));
  if (!c->fishd) CUP2D_HIP_CHECK(c, hipMalloc(&c->fishd, sizeof(FishDev) * MAX_FISH));
  const int idx = (int)c->hfish.size();
  CUP2D_HIP_CHECK(c, hipMemcpy(c->fishd + idx, &F, sizeof(F), hipMemcpyHostToDevice));
  c->hfish.push_back(F);
  c->g.nfish = (int)c->hfish.size();
  return CUP2D_OK;
}

void fish_free_all(cup2d_ctx* c) {
  for (FishDev& F : c->hfish)
    if (F.rS) (void)hipFree(F.rS);
  c->hfish.clear();
  if (c->fishd) (void)hipFree(c->fishd);
  c->fishd = nullptr;
}

int enqueue_fish_midline(cup2d_ctx* c, hipStream_t s) {
  if (c->hfish.empty()) return CUP2D_OK;
  hipLaunchKernelGGL(k_fish_midline, dim3((unsigned)c->hfish.size()), dim3(NT), 0, s, c->fishd,
                     c->ss, c->st, c->g.h);
  return launch_check(c, "fish midline");
}

int enqueue_fish_place(cup2d_ctx* c, hipStream_t s) {
  if (c->hfish.empty()) return CUP2D_OK;
  const int G = c->ob.cap > 0 ? c->ob.cap : 1;
  hipLaunchKernelGGL(k_fish_place, dim3(G), dim3(NT), 0, s, c->g, c->sp, c->ss, c->fishd, c->ob);
  return launch_check(c, "fish place");
}

}  // namespace cup2d
