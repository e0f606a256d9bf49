// fish.hip -- swimmers (CarlingFish) on the device: midline kinematics and
// internal-frame moments (Fish::create steps 1-5, Fish.cpp:60-150) and the
// per-block shape construction PutFishOnBlocks (FishLibrary.cpp:403-638).
//
// Arithmetic follows the reference expression by expression (-ffp-contract=off);
// trigonometry uses det_sincos and the midline integrals use a fixed
// 256-lane summation order, so a CPU restatement that does the same produces
// identical bits.  PutFishOnBlocks is a scatter in the reference (each
// midline point writes into the cells around it); here the minimum-distance
// surface winner comes from exact LDS atomic minima with the reference's
// first-wins order, and each cell gathers its internal contributions in the
// reference's loop order, so the sums are the same without float atomics.
#include <cmath>
#include <vector>

#include "device_util.hpp"
#include "fish.hpp"
#include "internal.hpp"

namespace cup2d {
namespace {

constexpr int NT = 256;
constexpr double EPSD = 2.220446049250313e-16;  // std::numeric_limits<double>::epsilon()

// AmplitudeFish::midlineLatPos / midlineLatVel (CarlingFish.cpp:19-27), or
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
__device__ inline void facs(const double* rS, const double* width, int Nm, const Mid& M, int i,
                            double* ds, double* fac1, double* fac2, double* fac3) {
  *ds = (i == 0) ? rS[1] - rS[0] : ((i == Nm - 1) ? rS[Nm - 1] - rS[Nm - 2] : rS[i + 1] - rS[i - 1]);
  const double w = width[i], w3 = w * w * w;
  const double dnorXi = d_ds(rS, Nm, i, M.norX), dnorYi = d_ds(rS, Nm, i, M.norY);
  *fac1 = 2 * w;
  *fac2 = 2 * w3 * (dnorXi * M.norY[i] - dnorYi * M.norX[i]) / 3;
  *fac3 = 2 * w3 / 3;
}

// value of v in lane `lane` (uniform) of this wave
__device__ inline double readlane_d(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, lane);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
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

// Fish::create steps 1-3 (Fish.cpp:78-95), first part, one workgroup per
// swimmer: AmplitudeFish::computeMidline (CarlingFish.cpp:51-88),
// integrateLinearMomentum, changeToCoMFrameLinear, integrateAngularMomentum
// and the angular-velocity correction of changeToCoMFrameAngular
// (FishLibrary.cpp:86-176), i.e. everything that depends only on the time.
// next = 0: at the current time; next = 1: at the end of the current step
// (time + dt, the value the step end forms), so the next step finds it ready.
// The midline arrays live in the swimmer's global work area (wk, L2-resident)
// and only the reduction scratch in LDS: this kernel runs beside the FFT row
// pass, whose workgroups need 2 x 68 KB of a CU's 160 KB LDS, and a larger
// LDS footprint here would take both of that pass's slots on one CU.
// Results go to midA / stA.
__global__ __launch_bounds__(NT) void k_fish_mid_a(const FishDev* __restrict__ fd,
                                                   const StepState* __restrict__ st, int next,
                                                   int parity, int buf) {
  __shared__ double red[5 * NT];
  const FishDev F = fd[blockIdx.x];
  const int Nm = F.Nm, t = threadIdx.x;
  const double* rS = F.rS;
  const double* width = F.width;
  double* w = F.wk;
  const Mid M{w, w + Nm, w + 2 * Nm, w + 3 * Nm, w + 4 * Nm, w + 5 * Nm, w + 6 * Nm, w + 7 * Nm};
  const double time = next ? st->tnext[parity] : st->time;
  const double T = F.Tperiod;
  const double rampFac = time < T ? det_sin(0.5 * M_PI * time / T) : 1.0;
  const double rampFacVel = time < T ? 0.5 * M_PI / T * det_cos(0.5 * M_PI * time / T) : 0.0;
  __syncthreads();
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
    M.rX[i] = dx;
    M.vX[i] = dVx;
    M.norX[i - 1] = -dy / ds;
    M.norY[i - 1] = dx / ds;
    M.vNorX[i - 1] = -dVy / ds;
    M.vNorY[i - 1] = dVx / ds;
  }
  __syncthreads();
  // running sums of the increments dx, dVx (rX[i] += rX[i-1]) in a fixed
  // two-level order, one wave per array: lane l sums its chunk of
  // c = ceil((Nm-1)/64) consecutive increments in index order, the lane totals
  // give the chunk offsets one lane after the other, and each element is its
  // chunk-local sum plus the chunk offset.  (The reference adds strictly in
  // index order; the two orders round differently in the last bit.)
  if (t < 128) {
    const int l = t & 63;
    double* a = t < 64 ? M.rX : M.vX;
    const int c = (Nm - 1 + 63) / 64;
    const int i0 = 1 + l * c, i1 = i0 + c < Nm ? i0 + c : Nm;
    double p = 0;
    for (int i = i0; i < i1; ++i) {
      p = a[i] + p;
      a[i] = p;
    }
    const double total = i0 < i1 ? p : 0.0;
    double off = 0, mine = 0;
    for (int j = 0; j < 64; ++j) {  // offset of lane j: totals of lanes < j
      mine = l == j ? off : mine;
      off = readlane_d(total, j) + off;
    }
    if (l > 0)
      for (int i = i0; i < i1; ++i) a[i] = a[i] + mine;
    if (l == 0) a[0] = 0;
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
      facs(rS, width, Nm, M, i, &ds, &fac1, &fac2, &fac3);
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
      facs(rS, width, Nm, M, i, &ds, &fac1, &fac2, &fac3);
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
  // changeToCoMFrameAngular, first half: v += angvel x r (before the rotation)
  for (int i = t; i < Nm; i += NT) {
    double* mA = F.midA + (size_t)buf * 4 * Nm;
    mA[i] = M.rX[i];
    mA[Nm + i] = M.rY[i];
    mA[2 * Nm + i] = M.vX[i] + angvel * M.rY[i];
    mA[3 * Nm + i] = M.vY[i] - angvel * M.rX[i];
  }
  if (t == 0) {
    double* sA = F.stA + buf * 8;
    sA[0] = area; sA[1] = J;
    sA[2] = C0; sA[3] = C1;
    sA[4] = V0; sA[5] = V1;
    sA[6] = angvel;
  }
}

// Fish::updatePosition's theta_internal update (Fish.cpp:180-185) and
// Fish::create steps 3-5, second part, one workgroup per swimmer, after the
// bodies have moved: rotation by theta_internal (changeToCoMFrameAngular),
// _computeMidlineNormals (FishLibrary.cpp:66-84), and the AreaSegment boxes
// in the computational frame (Fish.cpp:118-150, FishLibrary.cpp:295-342).
__global__ __launch_bounds__(NT) void k_fish_mid_b(const FishDev* __restrict__ fd,
                                                   const cup2d_shape_state* __restrict__ ss,
                                                   const StepState* __restrict__ st, double h,
                                                   int buf) {
  extern __shared__ double lds[];
  const FishDev F = fd[blockIdx.x];
  const int Nm = F.Nm, t = threadIdx.x;
  double* rS = lds;
  double* width = lds + Nm;
  const Mid M{lds + 2 * Nm, lds + 3 * Nm, lds + 4 * Nm, lds + 5 * Nm,
              lds + 6 * Nm, lds + 7 * Nm, lds + 8 * Nm, lds + 9 * Nm};
  double* red = lds + 10 * Nm;
  {
    // every load of the six arrays is issued before the first LDS store (a
    // loop of load-store pairs waits one memory round trip per element)
    constexpr int PER = (FISH_MAX_NM + NT - 1) / NT;
    const double* mA = F.midA + (size_t)buf * 4 * Nm;
    double r[6][PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int i = t + k * NT;
      const bool in = i < Nm;
      r[0][k] = in ? F.rS[i] : 0.0;
      r[1][k] = in ? F.width[i] : 0.0;
      r[2][k] = in ? mA[i] : 0.0;
      r[3][k] = in ? mA[Nm + i] : 0.0;
      r[4][k] = in ? mA[2 * Nm + i] : 0.0;
      r[5][k] = in ? mA[3 * Nm + i] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int i = t + k * NT;
      if (i < Nm) {
        rS[i] = r[0][k];
        width[i] = r[1][k];
        M.rX[i] = r[2][k];
        M.rY[i] = r[3][k];
        M.vX[i] = r[4][k];
        M.vY[i] = r[5][k];
      }
    }
  }
  const double theta = F.st[6] - st->dt * F.st[7];
  const double* sA = F.stA + buf * 8;
  const double angvel = sA[6];
  const double area = sA[0], J = sA[1];
  const double C0 = sA[2], C1 = sA[3], V0 = sA[4], V1 = sA[5];
  __syncthreads();
  // changeToCoMFrameAngular(theta_internal, angvel_internal), second half
  {
    double sn, cs;
    det_sincos(theta, &sn, &cs);
    const double R00 = cs, R01 = -sn, R10 = sn, R11 = cs;
    for (int i = t; i < Nm; i += NT) {
      const double vX = M.vX[i], vY = M.vY[i];
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
  // the step's midline for PutFishOnBlocks (rX rY vX vY norX norY vNorX vNorY)
  for (int i = t; i < 8 * Nm; i += NT) F.mid[i] = M.rX[i];
  // 4-5. AreaSegments in the computational frame of (center, orientation)
  const cup2d_shape_state& S = ss[F.shape];
  double sn, cs;
  det_sincos(S.orientation, &sn, &cs);
  const double R[2][2] = {{cs, -sn}, {sn, cs}};
  double ub[4] = {1e300, -1e300, 1e300, -1e300};
  // bound on any component of a deformation velocity of PutFishOnBlocks,
  // |R (v + o vNor)| <= |vX| + |vY| + w (|vNorX| + |vNorY|), for its fixed-point sums
  double vmax = 0;
  for (int i = t; i < Nm; i += NT)
    vmax = dmax(vmax, fabs(M.vX[i]) + fabs(M.vY[i]) + width[i] * (fabs(M.vNorX[i]) + fabs(M.vNorY[i])));
  for (int i = t; i < F.Nseg; i += NT) {
    const int next_idx = (i + 1) * (Nm - 1) / F.Nseg, idx = i * (Nm - 1) / F.Nseg;
    double bbox[2][2] = {{1e9, -1e9}, {1e9, -1e9}};
    for (int s = idx; s <= next_idx; ++s) {
      const double w = width[s];
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
  // min / max over the workgroup (exact, so the order does not matter): a
  // butterfly per wave, then the waves' values through LDS
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    ub[0] = dmin(ub[0], __shfl_xor(ub[0], o, 64));
    ub[1] = dmax(ub[1], __shfl_xor(ub[1], o, 64));
    ub[2] = dmin(ub[2], __shfl_xor(ub[2], o, 64));
    ub[3] = dmax(ub[3], __shfl_xor(ub[3], o, 64));
    vmax = dmax(vmax, __shfl_xor(vmax, o, 64));
  }
  __syncthreads();
  if (t % 64 == 0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) red[(t / 64) * 5 + q] = ub[q];
    red[(t / 64) * 5 + 4] = vmax;
  }
  __syncthreads();
  if (t == 0) {
    for (int w = 1; w < NT / 64; ++w) {
      red[0] = dmin(red[0], red[w * 5]);
      red[1] = dmax(red[1], red[w * 5 + 1]);
      red[2] = dmin(red[2], red[w * 5 + 2]);
      red[3] = dmax(red[3], red[w * 5 + 3]);
      red[4] = dmax(red[4], red[w * 5 + 4]);
    }
    for (int q = 0; q < 4; ++q) F.box[q] = red[q];
    int e = 0;
    (void)frexp(red[4], &e);  // vmax < 2^e
    F.st[8] = ldexp(1.0, 52 - e);  // velocity scale Su: |w ud Su| < 2^52
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
// slot that belongs to a swimmer: one workgroup per slot.
//   constructSurface (:438-573): a cell takes its signed squared distance,
//     chi weight W and W*udef from the first surface point, in loop order,
//     with the smallest squared distance among those whose 4x4 support covers
//     it.  Found in two parallel passes over (point, cell) pairs: an LDS
//     atomic minimum of the distance (non-negative doubles order like their
//     bit patterns), then an atomic minimum of the loop index among the points
//     that attain it exactly.
//   constructInternl (:575-638): hat-kernel P2M of the internal slices, each
//     cell summing its contributions in loop order (a gather: slices whose
//     extent misses a wave's rows are skipped by the whole wave).
//   signedDistanceSqrt (:419-436): udef /= chi, signed square root.
// Writes the slot's dist, udef and rho (chi is set by the Towers pass).
namespace {
enum { E_NW = 0, E_RX, E_RY, E_NX, E_NY, E_VX, E_VY, E_VNX, E_VNY, ENTD };
union PlaceLds {
  SurfPt pts[NT];
  double ent[NT][ENTD];
};
}  // namespace

__global__ __launch_bounds__(NT) void k_fish_place(Grid g, const ShapeParams* __restrict__ sp,
                                                   const cup2d_shape_state* __restrict__ ss,
                                                   const FishDev* __restrict__ fd,
                                                   Obstacles ob) {
  __shared__ int sl[FISH_MAX_SEG];
  __shared__ int s_tmp[NT];
  __shared__ unsigned long long s_best[BS * BS];
  __shared__ int s_win[BS * BS];
  __shared__ unsigned long long s_qw[BS * BS], s_qu[BS * BS], s_qv[BS * BS];
  __shared__ int s_touch[BS * BS];
  __shared__ PlaceLds U;
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
  const unsigned long long ONE = (unsigned long long)__double_as_longlong(1.0);
  for (int i = t; i < BS * BS; i += NT) {
    s_best[i] = ONE;  // |dist| starts at 1 (dist = -1)
    s_win[i] = 0x7fffffff;
  }
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
  // ---- constructSurface
  const int nchunk = (nsl + SEG_PER_CHUNK - 1) / SEG_PER_CHUNK;
  for (int pass = 0; pass < 2; ++pass) {
    for (int ch = 0; ch < nchunk; ++ch) {
      const int j0 = ch * SEG_PER_CHUNK;
      const int nj = nsl - j0 < SEG_PER_CHUNK ? nsl - j0 : SEG_PER_CHUNK;
      const int np = nj * PTS_PER_SEG;
      if (pass == 0 || nchunk > 1) {  // (one chunk: pass 1 reuses the points)
        __syncthreads();
        if (t < np) {
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
          U.pts[t] = P;
        }
        __syncthreads();
      }
      for (int idx = t; idx < np * 16; idx += NT) {
        const int p = idx >> 4, q = idx & 15;
        if (!U.pts[p].valid) continue;
        const int sx = U.pts[p].iap[0] - 1 + (q & 3), sy = U.pts[p].iap[1] - 1 + (q >> 2);
        if (sx < 0 || sx >= BS || sy < 0 || sy >= BS) continue;
        const double pc[2] = {cell_pos(bx, sx, g.hb, h), cell_pos(by, sy, g.hb, h)};
        const double dist0 = dist_sq(pc, U.pts[p].myP);
        const double distP = dist_sq(pc, U.pts[p].pP);
        const double distM = dist_sq(pc, U.pts[p].pM);
        const double d1 = dmin(dmin(dist0, distP), distM);
        const unsigned long long bits = (unsigned long long)__double_as_longlong(d1);
        if (pass == 0) {
          atomicMin(&s_best[sy * BS + sx], bits);
        } else if (d1 < 1.0 && bits == s_best[sy * BS + sx]) {
          atomicMin(&s_win[sy * BS + sx], j0 * PTS_PER_SEG + p);
        }
      }
    }
    __syncthreads();
  }
  const int cx = t % BS, cy0 = (t / BS) * 4;
  double dist[4], chi[4], udx[4], udy[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    dist[c] = -1; chi[c] = 0; udx[c] = 0; udy[c] = 0;
    const int win = s_win[(cy0 + c) * BS + cx];
    if (win == 0x7fffffff) continue;
    const int j = win / PTS_PER_SEG, r = win % PTS_PER_SEG;
    const int s0 = 8 * sl[j];
    const int first = s0 > 1 ? s0 : 1;
    const int sidx = first + r / 2, signp = (r & 1) ? 1 : -1;
    SurfPt P;
    surf_point(F, M, Fr, sidx, signp, org0, org1, invh, P);
    double p[2] = {cell_pos(bx, cx, g.hb, h), cell_pos(by, cy0 + c, g.hb, h)};
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
  // ---- constructInternl: hat P2M of the slice points at offsets iw*h,
  // every (point, patch cell) pair in parallel, summed in 64-bit fixed point
  // (weights x 2^52, weighted velocities x Su) so the sums do not depend on
  // the order of the additions
  __syncthreads();
  for (int i = t; i < BS * BS; i += NT) {
    s_qw[i] = 0; s_qu[i] = 0; s_qv[i] = 0; s_touch[i] = 0;
  }
  const double Su = F.st[8];
  for (int j0 = 0; j0 < nsl; j0 += SEG_PER_ICHUNK) {
    const int nj = nsl - j0 < SEG_PER_ICHUNK ? nsl - j0 : SEG_PER_ICHUNK;
    const int ne = nj * ENT_PER_SEG;
    __syncthreads();
    int npts = 0;
    if (t < ne) {
      const int j = j0 + t / ENT_PER_SEG, r = t % ENT_PER_SEG;
      const int s0 = 8 * sl[j], s1 = 8 * sl[j] + 8;
      const int first = s0 > 1 ? s0 : 1, last = s1 < Nm - 2 ? s1 : Nm - 2;
      const int sidx = first + r;
      if (sidx <= last) {
        const int Nw = (int)floor(F.width[sidx] / h);
        npts = Nw >= 1 ? 2 * Nw - 1 : 0;
        double* E = U.ent[t];
        E[E_NW] = Nw;
        E[E_RX] = M.rX[sidx]; E[E_RY] = M.rY[sidx];
        E[E_NX] = M.norX[sidx]; E[E_NY] = M.norY[sidx];
        E[E_VX] = M.vX[sidx]; E[E_VY] = M.vY[sidx];
        E[E_VNX] = M.vNorX[sidx]; E[E_VNY] = M.vNorY[sidx];
      }
    }
    s_tmp[t] = npts;
    __syncthreads();
    for (int off = 1; off < NT; off <<= 1) {  // inclusive scan of the point counts
      const int v = t >= off ? s_tmp[t - off] : 0;
      __syncthreads();
      s_tmp[t] += v;
      __syncthreads();
    }
    const int P = s_tmp[NT - 1];
    for (int idx = t; idx < P; idx += NT) {
      int lo = 0, hi = ne - 1;  // first entry whose inclusive count exceeds idx
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (s_tmp[mid] > idx) hi = mid; else lo = mid + 1;
      }
      const double* E = U.ent[lo];
      const int Nw = (int)E[E_NW];
      const int iw = -Nw + 1 + (idx - (lo ? s_tmp[lo - 1] : 0));
      const double offsetW = iw * h;
      double xp[2] = {E[E_RX] + offsetW * E[E_NX], E[E_RY] + offsetW * E[E_NY]};
      Fr.to_comp(xp);
      xp[0] = (xp[0] - org0) * invh;
      xp[1] = (xp[1] - org1) * invh;
      const double ap0 = floor(xp[0]), ap1 = floor(xp[1]);
      const int iap0 = (int)ap0, iap1 = (int)ap1;
      if (iap0 + 2 <= 0 || iap0 >= BS || iap1 + 2 <= 0 || iap1 >= BS) continue;
      double ud[2] = {E[E_VX] + offsetW * E[E_VNX], E[E_VY] + offsetW * E[E_VNY]};
      Fr.vel_to_comp(ud);
      const double wx[2] = {1 - fabs(xp[0] - ap0), 1 - fabs(xp[0] - (ap0 + 1))};
      const double wy[2] = {1 - fabs(xp[1] - ap1), 1 - fabs(xp[1] - (ap1 + 1))};
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int sx = q & 1, sy = q >> 1;
        const int cxq = iap0 + sx, cyq = iap1 + sy;
        if (cxq < 0 || cxq >= BS || cyq < 0 || cyq >= BS) continue;
        const double wxwy = wy[sy] * wx[sx];
        const int cell = cyq * BS + cxq;
        atomicAdd(&s_qw[cell], (unsigned long long)__double2ll_rn(wxwy * 0x1p52));
        atomicAdd(&s_qu[cell], (unsigned long long)__double2ll_rn(wxwy * ud[0] * Su));
        atomicAdd(&s_qv[cell], (unsigned long long)__double2ll_rn(wxwy * ud[1] * Su));
        s_touch[cell] = 1;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int cell = (cy0 + c) * BS + cx;
    if (!s_touch[cell]) continue;
    if (fabs(dist[c] + 1) < EPSD) dist[c] = 1;
    chi[c] += (double)(long long)s_qw[cell] * 0x1p-52;
    udx[c] += (double)(long long)s_qu[cell] / Su;
    udy[c] += (double)(long long)s_qv[cell] / Su;
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
// half-width table (AmplitudeFish::_width, CarlingFish.cpp:41-48, or the
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

static size_t fish_midline_lds(int Nm) { return sizeof(double) * ((size_t)10 * Nm + 5 * NT); }

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
  const size_t nd = (size_t)Nm * 26 + (size_t)F.Nseg * FISH_SEGD + 4 + 16 + 16;
  double* buf = nullptr;
  CUP2D_HIP_CHECK(c, hipMalloc(&buf, sizeof(double) * nd));
  CUP2D_HIP_CHECK(c, hipMemset(buf, 0, sizeof(double) * nd));
  F.rS = buf;
  F.width = buf + Nm;
  F.mid = buf + 2 * (size_t)Nm;
  F.seg = buf + 10 * (size_t)Nm;
  F.box = F.seg + (size_t)F.Nseg * FISH_SEGD;
  F.st = F.box + 4;
  F.stA = F.st + 16;
  F.midA = F.stA + 16;
  F.wk = F.midA + 8 * (size_t)Nm;
  CUP2D_HIP_CHECK(c, hipMemcpy(F.rS, rS.data(), sizeof(double) * Nm, hipMemcpyHostToDevice));
  CUP2D_HIP_CHECK(c, hipMemcpy(F.width, width.data(), sizeof(double) * Nm, hipMemcpyHostToDevice));
  if (!c->fishd) CUP2D_HIP_CHECK(c, hipMalloc(&c->fishd, sizeof(FishDev) * MAX_FISH));
  CUP2D_HIP_CHECK(c, hipFuncSetAttribute((const void*)k_fish_mid_b,
                                         hipFuncAttributeMaxDynamicSharedMemorySize,
                                         (int)fish_midline_lds(FISH_MAX_NM)));
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

// The time-only part lives in buffer host_step & 1: computed here when it is
// not ready, else by enqueue_fish_mid_next of the previous step.
int enqueue_fish_midline(cup2d_ctx* c, hipStream_t s) {
  if (c->hfish.empty()) return CUP2D_OK;
  const int buf = c->host_step & 1;
  // one workgroup per swimmer, LDS sized for the largest number of points
  const int nf = (int)c->hfish.size();
  size_t lds = 0;
  for (const auto& F : c->hfish) lds = fish_midline_lds(F.Nm) > lds ? fish_midline_lds(F.Nm) : lds;
  if (!c->midA_ready)
    hipLaunchKernelGGL(k_fish_mid_a, dim3(nf), dim3(NT), 0, s, c->fishd, c->st, 0, 0, buf);
  hipLaunchKernelGGL(k_fish_mid_b, dim3(nf), dim3(NT), lds, s, c->fishd, c->ss, c->st, c->g.h,
                     buf);
  c->midA_ready = false;
  return launch_check(c, "fish midline");
}

// The next step's time-only part, on the midline stream once this step's
// list builder has stored time + dt in st->tnext[parity] (event ev_begin;
// the next step's list builder writes the other slot), into the other
// buffer than the one this step's k_fish_mid_b reads.  With join_aux the
// side stream waits for it behind this step's obstacle chain, where it is
// idle, and the main stream reaches it through the side stream's next event;
// otherwise the caller orders the main stream after the midline stream.
int enqueue_fish_mid_next(cup2d_ctx* c, int parity, hipEvent_t after, bool join_aux) {
  if (c->hfish.empty()) return CUP2D_OK;
  // `after` follows this step's midline kernels, which share the work arrays
  CUP2D_HIP_CHECK(c, hipStreamWaitEvent(c->mids, after, 0));
  hipLaunchKernelGGL(k_fish_mid_a, dim3((int)c->hfish.size()), dim3(NT), 0, c->mids, c->fishd,
                     c->st, 1, parity, parity ^ 1);
  if (join_aux) {
    CUP2D_HIP_CHECK(c, hipEventRecord(c->ev_mid, c->mids));
    CUP2D_HIP_CHECK(c, hipStreamWaitEvent(c->aux, c->ev_mid, 0));
  }
  c->midA_ready = true;
  return launch_check(c, "fish midline (next step)");
}

int enqueue_fish_place(cup2d_ctx* c, hipStream_t s) {
  if (c->hfish.empty()) return CUP2D_OK;
  const int G = c->ob.cap > 0 ? c->ob.cap : 1;
  hipLaunchKernelGGL(k_fish_place, dim3(G), dim3(NT), 0, s, c->g, c->sp, c->ss, c->fishd, c->ob);
  return launch_check(c, "fish place");
}

}  // namespace cup2d
