// device_util.hpp -- small device helpers shared by the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "internal.hpp"

namespace cup2d {

// std::max / std::min semantics ((a<b)?b:a), so NaN handling and signed
// zeros follow the reference's calls exactly.
__host__ __device__ inline double dmax(double a, double b) { return (a < b) ? b : a; }
__host__ __device__ inline double dmin(double a, double b) { return (b < a) ? b : a; }
// The maximum of the findMaxU reductions: equal to dmax for numbers, but a
// NaN in either argument is kept, so a non-finite velocity reaches maxU
// (where the step raises CUP2D_ERR_NONFINITE, the role of the reference's
// Checker, Helpers.cpp:178-288) instead of being dropped.
__host__ __device__ inline double nmax(double a, double b) {
  return (a < b || b != b) ? b : a;
}
// maxU is finite (not NaN, not +inf)
__host__ __device__ inline bool finite_max(double m) { return m == m && m <= 1.7976931348623157e308; }

// sin and cos on the body path: Cody-Waite reduction by pi/2 (three-part
// constant, the first two parts with 33 significant bits so k*P1 and k*P2 are
// exact for |k| < 2^20) and Taylor polynomials on [-pi/4, pi/4] (sine to
// r^17, cosine to r^18; truncation below 1e-17).  Plain IEEE operations only,
// so host and device (and any CPU restatement written the same way) round
// identically; accuracy is about one ulp.
__host__ __device__ inline void det_sincos(double x, double* sn, double* cs) {
  const double P1 = 0x1.921fb544p+0, P2 = 0x1.0b4611a6p-34, P3 = 0x1.3198a2e037073p-69;
  const double k = floor(x * 0x1.45f306dc9c883p-1 + 0.5);
  const double r = ((x - k * P1) - k * P2) - k * P3;
  const double z = r * r;
  const double sp = r + (r * z) * (-1.0 / 6 + z * (1.0 / 120 + z * (-1.0 / 5040 + z * (1.0 / 362880 +
                    z * (-1.0 / 39916800 + z * (1.0 / 6227020800.0 + z * (-1.0 / 1307674368000.0 +
                    z * (1.0 / 355687428096000.0))))))));
  const double cp = 1 + z * (-0.5 + z * (1.0 / 24 + z * (-1.0 / 720 + z * (1.0 / 40320 +
                    z * (-1.0 / 3628800 + z * (1.0 / 479001600 + z * (-1.0 / 87178291200.0 +
                    z * (1.0 / 20922789888000.0 + z * (-1.0 / 6402373705728000.0)))))))));
  const int q = (int)(k - 4 * floor(k * 0.25));
  if (q == 0) { *sn = sp; *cs = cp; }
  else if (q == 1) { *sn = cp; *cs = -sp; }
  else if (q == 2) { *sn = -sp; *cs = -cp; }
  else { *sn = -cp; *cs = sp; }
}
__host__ __device__ inline double det_sin(double x) { double a, b; det_sincos(x, &a, &b); return a; }
__host__ __device__ inline double det_cos(double x) { double a, b; det_sincos(x, &a, &b); return b; }

// Cubism BlockInfo::pos: origin + h*(i+0.5), origin = block index * hb.
__host__ __device__ inline double cell_pos(int blk_index, int i, double hb, double h) {
  return (double)blk_index * hb + h * (i + 0.5);
}

// PoissonSolver::cub2rhs `_is_touching` (PoissonSolver.cpp:71-79): block
// (bx, by) lies within fadeLen of a domain face.
// band_block = band_col || band_row: the x and y tests separately
__host__ __device__ inline bool band_col(const Grid& g, int bx) {
  const double mn0 = cell_pos(bx, 0, g.hb, g.h), mx0 = cell_pos(bx, BS - 1, g.hb, g.h);
  return g.fadeLenX >= g.extent0 - mx0 || g.fadeLenX >= mn0;  // touchE || touchW
}
__host__ __device__ inline bool band_row(const Grid& g, int by) {
  const double mn1 = cell_pos(by, 0, g.hb, g.h), mx1 = cell_pos(by, BS - 1, g.hb, g.h);
  return g.fadeLenY >= g.extent1 - mx1 || g.fadeLenY >= mn1;  // touchN || touchS
}
__host__ __device__ inline bool band_block(const Grid& g, int bx, int by) {
  return band_row(g, by) || band_col(g, bx);
}

// cub2rhs fadeArg (PoissonSolver.cpp:88-93) at cell centre (px, py).
__host__ __device__ inline double fade_arg(const Grid& g, double px, double py) {
  constexpr double eps = 2.220446049250313e-16;  // std::numeric_limits<double>::epsilon()
  const double invFadeX = 1 / dmax(g.fadeLenX, eps), invFadeY = 1 / dmax(g.fadeLenY, eps);
  const double yt = invFadeY * dmax(0.0, g.fadeLenY - g.extent1 + py);
  const double yb = invFadeY * dmax(0.0, g.fadeLenY - py);
  const double xt = invFadeX * dmax(0.0, g.fadeLenX - g.extent0 + px);
  const double xb = invFadeX * dmax(0.0, g.fadeLenX - px);
  return dmin(dmax(dmax(dmax(yt, yb), xt), xb), 1.0);
}

// cub2rhs correction factor from the totals (sum, sum of positive parts,
// sum of negative parts) of the faded RHS (PoissonSolver.cpp:100-103).
__host__ __device__ inline double fade_corr_factor(double sumRHS, double sumPOS, double sumNEG) {
  constexpr double eps = 2.220446049250313e-16;
  sumNEG = dmax(sumNEG, eps);
  sumPOS = dmax(sumPOS, eps);
  return sumRHS > 0 ? sumRHS / sumPOS : sumRHS / sumNEG;
}

// The faded RHS of one cell of a band block (PoissonSolver.cpp:105-121).
__host__ __device__ inline double fade_cell(double r, double fa, double corr) {
  const double fadeRHS = r * (fa * fa);
  const double fadeCorrection = corr * (corr > 0 ? dmax(fadeRHS, 0.0) : -dmin(fadeRHS, 0.0));
  return r * (1 - fa * fa) + fadeCorrection;
}

template <int NT>
__device__ inline double block_sum(double v, double* lds) {
  // fixed-order tree: wave butterflies, then the per-wave results in order
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  constexpr int NW = NT / 64;
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  __syncthreads();
  if (l == 0) lds[w] = v;
  __syncthreads();
  double s = 0;
  if (NW > 1) {
#pragma unroll
    for (int i = 0; i < NW; ++i) s += lds[i];
  } else {
    s = lds[0];
  }
  return s;
}

template <int NT>
__device__ inline double block_max(double v, double* lds) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = nmax(v, __shfl_xor(v, o, 64));
  constexpr int NW = NT / 64;
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  __syncthreads();
  if (l == 0) lds[w] = v;
  __syncthreads();
  double s = lds[0];
#pragma unroll
  for (int i = 1; i < NW; ++i) s = nmax(s, lds[i]);
  return s;
}

// The same reductions for workgroups of any power-of-two size blockDim.x
// (32..1024): wave butterflies over min(64, blockDim.x) lanes.
__device__ inline double block_sum_dyn(double v, double* lds) {
  const int nt = blockDim.x, W = nt < 64 ? nt : 64;
  for (int o = W / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, W);
  const int NW = (nt + 63) / 64, w = threadIdx.x / 64, l = threadIdx.x % 64;
  __syncthreads();
  if (l == 0) lds[w] = v;
  __syncthreads();
  double s = lds[0];
  for (int i = 1; i < NW; ++i) s += lds[i];
  return s;
}
__device__ inline double block_max_dyn(double v, double* lds) {
  const int nt = blockDim.x, W = nt < 64 ? nt : 64;
  for (int o = W / 2; o > 0; o >>= 1) v = nmax(v, __shfl_xor(v, o, W));
  const int NW = (nt + 63) / 64, w = threadIdx.x / 64, l = threadIdx.x % 64;
  __syncthreads();
  if (l == 0) lds[w] = v;
  __syncthreads();
  double s = lds[0];
  for (int i = 1; i < NW; ++i) s = nmax(s, lds[i]);
  return s;
}

// NQ workgroup sums at once: a[j] <- sum over the workgroup of a[j], with the
// same tree as block_sum for each j (so the results are bit-identical to NQ
// separate block_sum calls) but one pair of barriers.  lds holds NT/64*NQ.
template <int NT, int NQ>
__device__ inline void block_sum_multi(double* a, double* lds) {
#pragma unroll
  for (int j = 0; j < NQ; ++j) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) a[j] += __shfl_xor(a[j], o, 64);
  }
  constexpr int NW = NT / 64;
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  __syncthreads();
  if (l == 0) {
#pragma unroll
    for (int j = 0; j < NQ; ++j) lds[w * NQ + j] = a[j];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < NQ; ++j) {
    double s = 0;
    if (NW > 1) {
#pragma unroll
      for (int i = 0; i < NW; ++i) s += lds[i * NQ + j];
    } else {
      s = lds[j];
    }
    a[j] = s;
  }
}

// Partial column sums of an n x NQ table (row i at a + i*stride) over the
// rows i0, i0+NL, ... in order.  Loads are issued U rows at a time ahead of
// the additions, so up to U*NL rows cost one memory latency.
template <int NL, int NQ, int U>
__device__ inline void strided_table_sum(const double* a, int n, int stride, int i0, double* s) {
#pragma unroll
  for (int j = 0; j < NQ; ++j) s[j] = 0;
  int i = i0;
  for (; i + (U - 1) * NL < n; i += U * NL) {
    double v[U][NQ];
#pragma unroll
    for (int r = 0; r < U; ++r)
#pragma unroll
      for (int j = 0; j < NQ; ++j) v[r][j] = a[(size_t)(i + r * NL) * stride + j];
#pragma unroll
    for (int r = 0; r < U; ++r)
#pragma unroll
      for (int j = 0; j < NQ; ++j) s[j] += v[r][j];
  }
  if (i < n) {  // fewer than U rows left: load them all, then add in order
    double v[U][NQ];
#pragma unroll
    for (int r = 0; r < U; ++r)
#pragma unroll
      for (int j = 0; j < NQ; ++j)
        v[r][j] = i + r * NL < n ? a[(size_t)(i + r * NL) * stride + j] : 0.0;
#pragma unroll
    for (int r = 0; r < U; ++r)
      if (i + r * NL < n)
#pragma unroll
        for (int j = 0; j < NQ; ++j) s[j] += v[r][j];
  }
}

// Column sums of an n x NQ table (row i at a + i*stride): thread t adds rows
// t, t+NT, ... in order, then block_sum_multi.  A table of up to U*NT rows
// costs one memory latency; the order of the additions, and so the result,
// is the same in every workgroup.
template <int NT, int NQ, int U = 4>
__device__ inline void block_sum_table(const double* a, int n, int stride, double* out,
                                       double* lds) {
  double s[NQ];
  strided_table_sum<NT, NQ, U>(a, n, stride, threadIdx.x, s);
  block_sum_multi<NT, NQ>(s, lds);
#pragma unroll
  for (int j = 0; j < NQ; ++j) out[j] = s[j];
}

// The same column sums by one wave (lane l adds rows l, l+64, ..., then a
// butterfly): no barrier, so the waves of a workgroup can reduce different
// tables at once.  Every lane gets the totals; the result is the same in
// every workgroup (it is not bit-identical to block_sum_table's).
template <int NQ, int U = 4>
__device__ inline void wave_sum_table(const double* a, int n, int stride, double* out) {
  double s[NQ];
  strided_table_sum<64, NQ, U>(a, n, stride, threadIdx.x % 64, s);
#pragma unroll
  for (int j = 0; j < NQ; ++j) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s[j] += __shfl_xor(s[j], o, 64);
    out[j] = s[j];
  }
}

// Sum of n values (fixed order independent of the caller): thread t sums
// t, t+NT, ... sequentially, then block_sum.  Every workgroup that calls it
// on the same data gets bit-identical results.
template <int NT>
__device__ inline double block_sum_array(const double* a, int n, int stride, double* lds) {
  double s;
  block_sum_table<NT, 1>(a, n, stride, &s, lds);
  return s;
}

// dst[0..n) = src[0..n) (n <= NT * U) by NT threads, every load issued
// before the first store: one memory round trip instead of one per element.
template <int NT, int U>
__device__ inline void stage_copy(double* dst, const double* __restrict__ src, int n) {
  double v[U];
#pragma unroll
  for (int r = 0; r < U; ++r) {
    const int i = (int)threadIdx.x + r * NT;
    v[r] = i < n ? src[i] : 0.0;
  }
#pragma unroll
  for (int r = 0; r < U; ++r) {
    const int i = (int)threadIdx.x + r * NT;
    if (i < n) dst[i] = v[r];
  }
}

// Max of n non-negative values with U loads in flight per thread (max is
// independent of the order; a NaN is kept).
template <int NT, int U = 8>
__device__ inline double max_array(const double* a, int n) {
  double m = 0;
  int i = threadIdx.x;
  for (; i + (U - 1) * NT < n; i += U * NT) {
    double v[U];
#pragma unroll
    for (int r = 0; r < U; ++r) v[r] = a[i + r * NT];
#pragma unroll
    for (int r = 0; r < U; ++r) m = nmax(m, v[r]);
  }
  for (; i < n; i += NT) m = nmax(m, a[i]);
  return m;
}

// ---------------------------------------------------------- advDiff flux
// The workgroups on the edge of a gx x gy stencil grid, numbered in blockID
// order (row 0, the first and last workgroup of each middle row, the last
// row); -1 for the others.  Only they see boundary cells.
__host__ __device__ inline int edge_count(int gx, int gy) {
  return gy == 1 ? gx : 2 * gx + (gx == 1 ? 1 : 2) * (gy - 2);
}
__device__ inline int edge_index(int bx, int by, int gx, int gy) {
  if (by == 0) return bx;
  if (by == gy - 1) return gx + (gx == 1 ? 1 : 2) * (gy - 2) + bx;
  if (bx == 0) return gx + (gx == 1 ? 1 : 2) * (by - 1);
  if (bx == gx - 1) return gx + 2 * (by - 1) + 1;
  return -1;
}

// corr = IF / (2 NY + 2 NX) (advDiff.cpp:105) from the edge workgroups'
// partials of the boundary outflow, summed in a fixed order (any workgroup
// size; the same size gives the same value).
__device__ inline double flux_corr(const double* __restrict__ fluxpart, int ne, const Grid& g,
                                   double* lds) {
  const int nt = blockDim.x;
  double s = 0;
  int e = threadIdx.x;
  for (; e < ne; e += 4 * nt) {  // four loads in flight, added in order
    const double a0 = fluxpart[e];
    const double a1 = e + nt < ne ? fluxpart[e + nt] : 0.0;
    const double a2 = e + 2 * nt < ne ? fluxpart[e + 2 * nt] : 0.0;
    const double a3 = e + 3 * nt < ne ? fluxpart[e + 3 * nt] : 0.0;
    s += a0;
    if (e + nt < ne) s += a1;
    if (e + 2 * nt < ne) s += a2;
    if (e + 3 * nt < ne) s += a3;
  }
  const double IF = block_sum_dyn(s, lds);
  return IF / (2 * (BS * g.bpdy) + 2 * (BS * g.bpdx));
}

// A velocity component as the advDiff flux correction leaves it on the
// boundary (advDiff.cpp:107-115): + corr on the first cell of its direction,
// - corr on the last (n cells).
// T is the velocity planes' type (double, or float at FP32 precision, where
// the correction is applied in float as well).
template <typename T>
__device__ inline T flux_fixed(T val, int i, int n, T corr) {
  if (i == 0) return val + corr;
  if (i == n - 1) return val - corr;
  return val;
}

// PressureSingle::updatePressureRHS's divergence term at cell (x, y)
// (PressureSingle.cpp:112-128), r = facDiv (u_E - u_W + v_N - v_S) with
// zero-gradient ghosts, on the advDiff output u, v whose boundary flux
// correction corr is still to be applied; the same operations as k_rhs.
template <typename T>
__device__ inline T rhs_cell(const T* __restrict__ u, const T* __restrict__ v, int NX, int NY,
                             int x, int y, T corr, T facDiv) {
  const int xe = x + 1 < NX ? x + 1 : x, xw = x > 0 ? x - 1 : x;
  const int yn = y + 1 < NY ? y + 1 : y, ys = y > 0 ? y - 1 : y;
  const T ue = flux_fixed<T>(u[(size_t)y * NX + xe], xe, NX, corr);
  const T uw = flux_fixed<T>(u[(size_t)y * NX + xw], xw, NX, corr);
  const T vn = flux_fixed<T>(v[(size_t)yn * NX + x], yn, NY, corr);
  const T vs = flux_fixed<T>(v[(size_t)ys * NX + x], ys, NY, corr);
  const T divVx = ue - uw;
  const T divVy = vn - vs;
  return facDiv * (divVx + divVy);
}

// The cub2rhs fade sums of one band block (PoissonSolver.cpp:81-98): of the
// faded RHS r*fa^2, of its positive parts and of its negative parts, by a
// 256-thread workgroup in a fixed order, so the partial is the same whichever
// kernel forms it.  from_plane: the RHS is read from r, else formed from the
// advDiff output u, v with the deferred flux correction (rhs_cell).
template <typename T>
__device__ inline void fade_block_sums(const Grid& g, int b, const T* __restrict__ r,
                                       const T* __restrict__ u, const T* __restrict__ v,
                                       bool from_plane, T corr, T facDiv, double* out,
                                       double* lds) {
  const int bx = b % g.bpdx, by = b / g.bpdx;
  double sR = 0, sP = 0, sN = 0;
#pragma unroll
  for (int it = 0, k = threadIdx.x; it < BS * BS / 256; ++it, k += 256) {
    const int ix = k % BS, iy = k / BS, x = bx * BS + ix, y = by * BS + iy;
    const double px = cell_pos(bx, ix, g.hb, g.h), py = cell_pos(by, iy, g.hb, g.h);
    const double fa = fade_arg(g, px, py);
    const double rc = from_plane ? (double)r[(size_t)y * g.NX + x]
                                 : (double)rhs_cell<T>(u, v, g.NX, g.NY, x, y, corr, facDiv);
    const double f = rc * (fa * fa);
    if (f < 0) sN -= f;
    if (f > 0) sP += f;
    sR += f;
  }
  out[0] = sR; out[1] = sP; out[2] = sN;
  block_sum_multi<256, 3>(out, lds);
}

// A value of the lane to the left (lane - 1) or right (lane + 1) in the
// wave, by a whole-wave DPP shift (a VALU move, no memory); the lane without
// such a neighbour in the wave gets its own value back, and the callers
// replace it.  A DPP move reads a disabled lane as the destination's old
// value, so the shift must run with every lane of the wave active: callers
// shift unconditionally, outside any branch or lazily evaluated operand, and
// the empty volatile asm keeps the compiler from moving the move into one.
__device__ __forceinline__ int dpp_move(int x, int ctrl_shl) {
  int r = ctrl_shl ? __builtin_amdgcn_update_dpp(x, x, 0x130 /* wave_shl:1 */, 0xf, 0xf, false)
                   : __builtin_amdgcn_update_dpp(x, x, 0x138 /* wave_shr:1 */, 0xf, 0xf, false);
  asm volatile("" : "+v"(r));
  return r;
}
__device__ __forceinline__ float lane_from_left(float x) {
  return __int_as_float(dpp_move(__float_as_int(x), 0));
}
__device__ __forceinline__ float lane_from_right(float x) {
  return __int_as_float(dpp_move(__float_as_int(x), 1));
}
__device__ __forceinline__ double lane_from_left(double x) {
  return __hiloint2double(dpp_move(__double2hiint(x), 0), dpp_move(__double2loint(x), 0));
}
__device__ __forceinline__ double lane_from_right(double x) {
  return __hiloint2double(dpp_move(__double2hiint(x), 1), dpp_move(__double2loint(x), 1));
}

// The imposed surface velocity (SurfVel) at offset (p0, p1) from the
// shape's centre: z = e^{i phi} = (p0, p1)/|p|, its powers by complex
// multiplication (no transcendental calls), normal n = (cos, sin), tangent
// t = (-sin, cos).  At the centre itself phi = 0.
__device__ inline void surf_vel(const SurfVel& sv, double p0, double p1, double* ux,
                                double* uy) {
  const double r = sqrt(p0 * p0 + p1 * p1);
  const double zr = r > 0 ? p0 / r : 1.0, zi = r > 0 ? p1 / r : 0.0;
  double wr = 1, wi = 0, un = 0, ut = 0;
  const int M = sv.M;
  for (int m = 0; m < M; ++m) {
    const double tr = wr * zr - wi * zi;
    wi = wr * zi + wi * zr;
    wr = tr;
    un += sv.coef[m] * wr + sv.coef[M + m] * wi;
    ut += sv.coef[2 * M + m] * wr + sv.coef[3 * M + m] * wi;
  }
  *ux = un * zr - ut * zi;
  *uy = un * zi + ut * zr;
}

}  // namespace cup2d
