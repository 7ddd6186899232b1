// Per-iteration summaries of the responsibilities for the viewer-era
// outputs (best rotation of each pattern, responsibility of each rotation
// averaged over the patterns).  No counterpart kernel in the reference; the
// outputs are those viewer/rotations_module.py and viewer/slice_module.py
// read.
#include <hip/hip_runtime.h>

#include "engine.hpp"
#include "reduce.hpp"

namespace emc {

// best[i] = argmax_r P[r][i] (lowest r on ties), one lane per pattern
// (coalesced rows).
__global__ __launch_bounds__(256) void k_best_rotation(const float* __restrict__ P, int R, int N,
                                                       int* __restrict__ best) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= N) return;
  float m = -1.0f;
  int b = 0;
  for (int r = 0; r < R; r++) {
    const float v = P[(size_t)r * N + i];
    if (v > m) {
      m = v;
      b = r;
    }
  }
  best[i] = b;
}

// avg[r] = sum_i P[r][i] / N, one workgroup per rotation, FP64 sum.
__global__ __launch_bounds__(256) void k_average_respons(const float* __restrict__ P, int R, int N,
                                                         float* __restrict__ avg) {
  const int r = blockIdx.x;
  double s = 0.0;
  for (int i = threadIdx.x; i < N; i += 256) s += P[(size_t)r * N + i];
  s = block_sum_f64(s);
  if (threadIdx.x == 0) avg[r] = (float)(s / N);
}

}  // namespace emc

using namespace emc;

extern "C" int emc_hip_respons_summary(const float* d_respons, int N_slices, int N_images, int* best,
                                       float* average) {
  if (!ready()) return 1;
  Engine& e = engine();
  if (!d_respons || N_slices <= 0 || N_images <= 0) {
    set_error("emc_hip_respons_summary: invalid argument");
    return 1;
  }
  DBuf b, a;
  if ((best && !b.ensure(sizeof(int) * N_images)) || (average && !a.ensure(sizeof(float) * N_slices))) return 1;
  if (best) {
    hipLaunchKernelGGL(k_best_rotation, dim3((N_images + 255) / 256), dim3(256), 0, e.stream, d_respons, N_slices,
                       N_images, b.as<int>());
    EMC_LAUNCHED();
    EMC_CHECK(hipMemcpyAsync(best, b.p, sizeof(int) * N_images, hipMemcpyDeviceToHost, e.stream));
  }
  if (average) {
    hipLaunchKernelGGL(k_average_respons, dim3(N_slices), dim3(256), 0, e.stream, d_respons, N_slices, N_images,
                       a.as<float>());
    EMC_LAUNCHED();
    EMC_CHECK(hipMemcpyAsync(average, a.p, sizeof(float) * N_slices, hipMemcpyDeviceToHost, e.stream));
  }
  EMC_CHECK(hipStreamSynchronize(e.stream));
  b.release();
  a.release();
  return failed() ? 1 : 0;
}
