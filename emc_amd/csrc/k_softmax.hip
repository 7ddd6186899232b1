// Per-pattern normalisation of the log-likelihoods into responsibilities, the
// sum P log P diagnostic and the scaling update.
//
// Responsibilities follow emc_cuda.cu:445-476 (intended semantics, i.e. the
// true per-pattern maximum): m_i = max(-1e10, max_r L_ri),
// P_ri = exp(L_ri - m_i) if L_ri - m_i > -1e10 else 0, P_ri /= sum_r P_ri.
// The exponentials are FP32 as in the reference; their per-pattern sum is
// accumulated in FP64 (log-sum-exp with an online maximum).
//
// Layout is the ABI's [R][N]: one lane per pattern, so every load of a wave
// is 64 consecutive patterns of one rotation (256 contiguous bytes).  The
// rotation axis is split into chunks for parallelism; chunk partials
// (max, sum) are merged in a fixed order.
#include <hip/hip_runtime.h>

#include "engine.hpp"
#include "reduce.hpp"

namespace emc {

int softmax_chunks(int N, int R) {
  const int blocks = (N + 255) / 256;
  int c = 1;
  while (blocks * c < 2048 && c * 2 <= R && c < 64) c *= 2;
  return c;
}

constexpr int kSmBatch = 8;

__global__ __launch_bounds__(256) void k_softmax_partial(const float* __restrict__ L, int R, int N, int rows,
                                                         float* __restrict__ smax, double* __restrict__ ssum) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int c = blockIdx.y;
  if (i >= N) return;
  const int r0 = c * rows, r1 = min(R, r0 + rows);
  float m = -1.0e10f;
  double s = 0.0;
  // the rows' values are loaded kSmBatch at a time before they are used (the
  // arithmetic and its order are those of one row at a time)
  for (int rb = r0; rb < r1; rb += kSmBatch) {
    float vb[kSmBatch];
#pragma unroll
    for (int u = 0; u < kSmBatch; u++) vb[u] = rb + u < r1 ? L[(size_t)(rb + u) * N + i] : 0.0f;
#pragma unroll
    for (int u = 0; u < kSmBatch; u++) {
      if (rb + u >= r1) break;
      const float v = vb[u];
      if (m < v) {
        s *= exp((double)m - (double)v);
        m = v;
      }
      const float d = v - m;
      if (d > -1.0e10f) s += (double)expf(d);
    }
  }
  smax[(size_t)c * N + i] = m;
  ssum[(size_t)c * N + i] = s;
}

// sacc layout: [4][nmerge][N] = sum P log P, sum P, sum P A_r, sum P L_ri.
// The normalisation merges nmerge >= chunks (max, sum) partials: the chunks
// of L, and beyond them partials of values held elsewhere (the candidate
// cells of the rotation proposal), merged by merge_stats in one order.
__global__ __launch_bounds__(256) void k_softmax_finish(float* __restrict__ L, int R, int N, int rows, int nmerge,
                                                        const float* __restrict__ smax,
                                                        const double* __restrict__ ssum,
                                                        double* __restrict__ sacc,
                                                        const double* __restrict__ A, int fuse) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int c = blockIdx.y;
  if (i >= N) return;
  float m;
  double inv;
  merge_stats(smax, ssum, nmerge, N, i, m, inv);
  const int r0 = c * rows, r1 = min(R, r0 + rows);
  double plogp = 0.0, psum = 0.0, pa = 0.0, pl = 0.0;
  for (int rb = r0; rb < r1; rb += kSmBatch) {
    float vb[kSmBatch];
#pragma unroll
    for (int u = 0; u < kSmBatch; u++) vb[u] = rb + u < r1 ? L[(size_t)(rb + u) * N + i] : 0.0f;
#pragma unroll
    for (int u = 0; u < kSmBatch; u++) {
      if (rb + u >= r1) break;
      const int r = rb + u;
      const float v = vb[u];
      const float d = v - m;
      const float e = d > -1.0e10f ? expf(d) : 0.0f;
      const float p = (float)((double)e * inv);
      L[(size_t)r * N + i] = p;
      if (p > 0.0f) plogp += (double)(p * logf(p));
      psum += p;
      if (fuse) {
        pa += (double)p * A[r];
        pl += (double)p * (double)v;
      }
    }
  }
  const size_t cn = (size_t)nmerge * N;
  sacc[0 * cn + (size_t)c * N + i] = plogp;
  sacc[1 * cn + (size_t)c * N + i] = psum;
  sacc[2 * cn + (size_t)c * N + i] = pa;
  sacc[3 * cn + (size_t)c * N + i] = pl;
}

// Per pattern: merge chunk partials; with the likelihood's A_r, B_i and phi
// known, sum_r P_ri C_ir follows from L without recomputing C:
//   C_ir = (phi_i / 2) (A_r + B_i / phi_i^2 + 2 n_m sigma^2 L_ri).
// red[blockIdx.x] receives the block's sum of P log P.  Patterns with
// skip[i] set are left out (no phiden written, no P log P counted).
__global__ __launch_bounds__(256) void k_softmax_reduce(int N, int chunks, const double* __restrict__ sacc, int fuse,
                                                        const double* __restrict__ B, const float* __restrict__ phi,
                                                        double sigma2x, double* __restrict__ phiden,
                                                        double* __restrict__ red, int poisson,
                                                        const unsigned char* __restrict__ skip) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  double plogp = 0.0;
  if (i < N && !(skip && skip[i])) {
    const size_t cn = (size_t)chunks * N;
    double ps = 0.0, pa = 0.0, pl = 0.0;
    for (int k = 0; k < chunks; k++) {
      const size_t o = (size_t)k * N + i;
      plogp += sacc[o];
      ps += sacc[cn + o];
      pa += sacc[2 * cn + o];
      pl += sacc[3 * cn + o];
    }
    if (fuse) {
      const double f = (double)phi[i];
      // Poisson: the denominator is sum_r P_ri A_r itself
      phiden[i] = poisson ? pa : 0.5 * f * (pa + B[i] / (f * f) * ps + sigma2x * pl);
    }
  }
  plogp = block_sum_f64(plogp);
  if (threadIdx.x == 0) red[blockIdx.x] = plogp;
}

// Sum of n doubles in a fixed order, result in out[0].
__global__ __launch_bounds__(1024) void k_sum_f64(const double* __restrict__ in, int n, double* __restrict__ out) {
  double s = 0.0;
  for (int k = threadIdx.x; k < n; k += 1024) s += in[k];
  s = block_sum_f64(s);
  if (threadIdx.x == 0) out[0] = s;
}

void launch_softmax_partial(const float* L, int R, int N, int chunks, float* smax, double* ssum, hipStream_t st) {
  const int rows = (R + chunks - 1) / chunks;
  const dim3 grid((N + 255) / 256, chunks);
  hipLaunchKernelGGL(k_softmax_partial, grid, dim3(256), 0, st, L, R, N, rows, smax, ssum);
  EMC_LAUNCHED();
}

void launch_softmax_finish(float* L, int R, int N, int chunks, int nmerge, const float* smax, const double* ssum,
                           double* sacc, const double* A, bool fuse_phi, hipStream_t st) {
  const int rows = (R + chunks - 1) / chunks;
  const dim3 grid((N + 255) / 256, chunks);
  hipLaunchKernelGGL(k_softmax_finish, grid, dim3(256), 0, st, L, R, N, rows, nmerge, smax, ssum, sacc, A,
                     (int)fuse_phi);
  EMC_LAUNCHED();
}

void launch_softmax(float* L, int R, int N, int chunks, float* smax, double* ssum, double* sacc,
                    const double* A, bool fuse_phi, hipStream_t st) {
  launch_softmax_partial(L, R, N, chunks, smax, ssum, st);
  launch_softmax_finish(L, R, N, chunks, chunks, smax, ssum, sacc, A, fuse_phi, st);
}

// red must hold ceil(N/256) + 1 doubles; red[ceil(N/256)] = sum P log P.
void launch_softmax_reduce(int N, int chunks, const double* sacc, bool fuse_phi, const double* B,
                           const float* phi_used, double sigma2x, double* phiden, double* red,
                           hipStream_t st, bool poisson, const unsigned char* skip) {
  const int nb = (N + 255) / 256;
  hipLaunchKernelGGL(k_softmax_reduce, dim3(nb), dim3(256), 0, st, N, chunks, sacc, (int)fuse_phi, B, phi_used,
                     sigma2x, phiden, red, (int)poisson, skip);
  EMC_LAUNCHED();
  hipLaunchKernelGGL(k_sum_f64, dim3(1), dim3(256), 0, st, red, nb, red + nb);
  EMC_LAUNCHED();
}

// sum over n of P log P (terms with P <= 0 count 0), emc_cuda.cu:484-505.
__global__ __launch_bounds__(256) void k_plogp_partial(const float* __restrict__ P, size_t n, double* __restrict__ red) {
  double s = 0.0;
  for (size_t k = (size_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (size_t)gridDim.x * 256) {
    const float p = P[k];
    if (p > 0.0f) s += (double)(p * logf(p));
  }
  s = block_sum_f64(s);
  if (threadIdx.x == 0) red[blockIdx.x] = s;
}

void launch_sum_f64(const double* v, int n, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_sum_f64, dim3(1), dim3(1024), 0, st, v, n, out);
  EMC_LAUNCHED();
}

// red must hold 1025 doubles; red[1024] = result.
void launch_total_respons(const float* P, size_t n, double* red, hipStream_t st) {
  const int nb = 1024;
  hipLaunchKernelGGL(k_plogp_partial, dim3(nb), dim3(256), 0, st, P, n, red);
  EMC_LAUNCHED();
  hipLaunchKernelGGL(k_sum_f64, dim3(1), dim3(256), 0, st, red, nb, red + nb);
  EMC_LAUNCHED();
}

// scaling_i = sum_p m K_ip^2 / sum_r P_ri C_ir  (emc_cuda.cu:195-228).
__global__ __launch_bounds__(256) void k_scaling_finish(int N, const double* __restrict__ B,
                                                        const double* __restrict__ D, float* __restrict__ scaling,
                                                        float* __restrict__ invphi) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= N) return;
  const float s = (float)(B[i] / D[i]);
  scaling[i] = s;
  invphi[i] = 1.0f / s;
}

__global__ void k_reciprocal(const float* __restrict__ s, int N, float* __restrict__ inv) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < N) inv[i] = 1.0f / s[i];
}

void launch_reciprocal(const float* s, int N, float* inv, hipStream_t st) {
  if (N <= 0) return;
  hipLaunchKernelGGL(k_reciprocal, dim3((N + 255) / 256), dim3(256), 0, st, s, N, inv);
  EMC_LAUNCHED();
}

void launch_scaling_finish(int N, const double* B, const double* phiden, float* scaling, float* invphi,
                           hipStream_t st) {
  hipLaunchKernelGGL(k_scaling_finish, dim3((N + 255) / 256), dim3(256), 0, st, N, B, phiden, scaling, invphi);
  EMC_LAUNCHED();
}

}  // namespace emc
