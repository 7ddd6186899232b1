// Wave64 and workgroup reductions (fixed order, so results are deterministic).
#pragma once

#include <hip/hip_runtime.h>

namespace emc {

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
template <class T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    T u = __shfl_xor(v, o, 64);
    v = u > v ? u : v;
  }
  return v;
}

// Block-wide reductions for blockDim.x a multiple of 64, <= 1024.  Every
// thread receives the result.
template <class T, class Op>
__device__ __forceinline__ T block_reduce(T v, Op op, T (*wave)(T)) {
  __shared__ T part[16];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  v = wave(v);
  __syncthreads();  // part[] may still be read by a previous call
  if (lane == 0) part[w] = v;
  __syncthreads();
  T r = part[0];
  for (int k = 1; k < nw; k++) r = op(r, part[k]);
  return r;
}

struct OpAdd {
  template <class T>
  __device__ T operator()(T a, T b) const { return a + b; }
};
struct OpMax {
  template <class T>
  __device__ T operator()(T a, T b) const { return a > b ? a : b; }
};

__device__ __forceinline__ double block_sum_f64(double v) { return block_reduce<double>(v, OpAdd(), wave_sum<double>); }
__device__ __forceinline__ float block_sum_f32(float v) { return block_reduce<float>(v, OpAdd(), wave_sum<float>); }
__device__ __forceinline__ float block_max_f32(float v) { return block_reduce<float>(v, OpMax(), wave_max<float>); }
__device__ __forceinline__ unsigned int block_max_u32(unsigned int v) {
  return block_reduce<unsigned int>(v, OpMax(), wave_max<unsigned int>);
}

// Per-pattern normalisation from nmerge (max, sum of exp(x - max)) partials
// [nmerge][N]: the overall maximum m (at least -1e10, the reference's floor)
// and 1 / sum exp(x - m), merged in partial order.  Every kernel that
// normalises values of one pattern held in different places calls this on
// the same partials and so uses bit-identical m and 1 / sum.
__device__ __forceinline__ void merge_stats(const float* __restrict__ smax, const double* __restrict__ ssum, int nmerge,
                                            int N, int i, float& m, double& inv) {
  m = -1.0e10f;
  for (int k = 0; k < nmerge; k++) m = fmaxf(m, smax[(size_t)k * N + i]);
  double s = 0.0;
  for (int k = 0; k < nmerge; k++) s += ssum[(size_t)k * N + i] * exp((double)smax[(size_t)k * N + i] - (double)m);
  inv = 1.0 / s;
}

}  // namespace emc
