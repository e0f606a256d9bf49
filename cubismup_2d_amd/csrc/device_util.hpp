// device_util.hpp -- small device helpers shared by the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "internal.hpp"

namespace cup2d {

// std::max / std::min semantics ((a<b)?b:a), so NaN handling and signed
// zeros follow the reference's calls exactly.
__host__ __device__ inline double dmax(double a, double b) { return (a < b) ? b : a; }
__host__ __device__ inline double dmin(double a, double b) { return (b < a) ? b : a; }

// Cubism BlockInfo::pos: origin + h*(i+0.5), origin = block index * hb.
__device__ inline double cell_pos(int blk_index, int i, double hb, double h) {
  return (double)blk_index * hb + h * (i + 0.5);
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
  for (int o = 32; o > 0; o >>= 1) v = dmax(v, __shfl_xor(v, o, 64));
  constexpr int NW = NT / 64;
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  __syncthreads();
  if (l == 0) lds[w] = v;
  __syncthreads();
  double s = lds[0];
#pragma unroll
  for (int i = 1; i < NW; ++i) s = dmax(s, lds[i]);
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
  for (int o = W / 2; o > 0; o >>= 1) v = dmax(v, __shfl_xor(v, o, W));
  const int NW = (nt + 63) / 64, w = threadIdx.x / 64, l = threadIdx.x % 64;
  __syncthreads();
  if (l == 0) lds[w] = v;
  __syncthreads();
  double s = lds[0];
  for (int i = 1; i < NW; ++i) s = dmax(s, lds[i]);
  return s;
}

// Sum of n values (fixed order independent of the caller): thread t sums
// t, t+NT, ... sequentially, then block_sum.  Every workgroup that calls it
// on the same data gets bit-identical results.
template <int NT>
__device__ inline double block_sum_array(const double* a, int n, int stride, double* lds) {
  double s = 0;
  for (int i = threadIdx.x; i < n; i += NT) s += a[(size_t)i * stride];
  return block_sum<NT>(s, lds);
}

}  // namespace cup2d
