// FP32 matrix-core contractions (plain form: the reference's method in FP32).
//
// E-step (likelihood, emc_cuda.cu:113-150 restated as a GEMM):
//   C[i][r] = sum_p m_p K_ip S_rp  over the packed masked pixels p, then
//   L[r][i] = -(A_r - 2 C/phi_i + B_i/phi_i^2) / (2 n_m sigma^2)
// with A_r = sum m S^2 and B_i = sum m K^2 (double), which is the expansion of
// the reference's -sum m (S - K/phi)^2 / (2 n_m sigma^2).
//
// M-step (emc_atomic.cu:76-113 restated as a GEMM):
//   U[r][p] = sum_i (P_ri / phi_i) K_ip,  S'_rp = U / T_r when T_r > 1e-10.
//
// Both are "NT" products of two row-major operands whose rows are contiguous
// along the reduction index.  Workgroup tile 128 x 128, 4 wave64s in a 2 x 2
// arrangement, each wave 64 x 64 = 4 x 4 blocks of v_mfma_f32_16x16x4_f32.
// Operand tiles are staged through LDS, double buffered, with the next tile's
// global loads issued before the current tile's MFMAs.
//
// MFMA 16x16 operand/result layout (wave64): lane l supplies row (l % 16) of
// A and column (l % 16) of B for the reduction slots of group g = l / 16;
// result element (4g + j, l % 16) lands in accumulator register j of lane l.
// A lane holds a float4 of consecutive k; the four MFMAs of a 16-wide k step
// consume element s of it, so A and B see the same k in every slot.
#include <hip/hip_runtime.h>

#include "engine.hpp"
#include "gemm_common.hpp"

namespace emc {

typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int BM = kTileM, BN = kTileN;
constexpr int BK32 = 16;
constexpr int LD32 = BK32 + 4;  // padded LDS row pitch (floats)

// ---------------------------------------------------------------------------
// FP32 E-step
// ---------------------------------------------------------------------------
template <bool PC>
__global__ __launch_bounds__(256) void k_estep_f32(const float* __restrict__ K, const float* __restrict__ S,
                                                   int n_mp, int tiles_m, int tiles_n, LikeEpilogue ep) {
  __shared__ __attribute__((aligned(16))) float sA[2][BM][LD32];
  __shared__ __attribute__((aligned(16))) float sB[2][BN][LD32];
  int tm, tn;
  tile_of_block(tiles_m, tiles_n, tm, tn);
  const int m0 = tm * BM, n0 = tn * BN;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6, wm = w >> 1, wn = w & 1;
  const int lr = t >> 2, lc = (t & 3) * 4;  // global->LDS: rows lr, lr+64; 4 floats at lc
  const float* gA = K + (size_t)(m0 + lr) * n_mp + lc;
  const float* gB = S + (size_t)(n0 + lr) * n_mp + lc;
  const size_t step64 = (size_t)64 * n_mp;

  floatx4 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 4; b++) acc[a][b] = floatx4{0, 0, 0, 0};

  floatx4 ra0, ra1, rb0, rb1;
  ra0 = *reinterpret_cast<const floatx4*>(gA);
  ra1 = *reinterpret_cast<const floatx4*>(gA + step64);
  rb0 = *reinterpret_cast<const floatx4*>(gB);
  rb1 = *reinterpret_cast<const floatx4*>(gB + step64);
  *reinterpret_cast<floatx4*>(&sA[0][lr][lc]) = ra0;
  *reinterpret_cast<floatx4*>(&sA[0][lr + 64][lc]) = ra1;
  *reinterpret_cast<floatx4*>(&sB[0][lr][lc]) = rb0;
  *reinterpret_cast<floatx4*>(&sB[0][lr + 64][lc]) = rb1;
  __syncthreads();

  const int nk = n_mp / BK32;
  const int fr = lane & 15, fk = (lane >> 4) * 4;
  for (int kt = 0; kt < nk; kt++) {
    const int buf = kt & 1;
    const bool more = kt + 1 < nk;
    if (more) {
      const size_t off = (size_t)(kt + 1) * BK32;
      ra0 = *reinterpret_cast<const floatx4*>(gA + off);
      ra1 = *reinterpret_cast<const floatx4*>(gA + off + step64);
      rb0 = *reinterpret_cast<const floatx4*>(gB + off);
      rb1 = *reinterpret_cast<const floatx4*>(gB + off + step64);
    }
    floatx4 fa[4], fb[4];
#pragma unroll
    for (int a = 0; a < 4; a++) fa[a] = *reinterpret_cast<const floatx4*>(&sA[buf][wm * 64 + a * 16 + fr][fk]);
#pragma unroll
    for (int b = 0; b < 4; b++) fb[b] = *reinterpret_cast<const floatx4*>(&sB[buf][wn * 64 + b * 16 + fr][fk]);
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
      for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[a][s], fb[b][s], acc[a][b], 0, 0, 0);
    if (more) {
      const int nb = buf ^ 1;
      *reinterpret_cast<floatx4*>(&sA[nb][lr][lc]) = ra0;
      *reinterpret_cast<floatx4*>(&sA[nb][lr + 64][lc]) = ra1;
      *reinterpret_cast<floatx4*>(&sB[nb][lr][lc]) = rb0;
      *reinterpret_cast<floatx4*>(&sB[nb][lr + 64][lc]) = rb1;
    }
    __syncthreads();
  }

  const int g = lane >> 4;
  if (!PC) {
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
      for (int b = 0; b < 4; b++)
        like_store(ep, m0 + wm * 64 + a * 16 + 4 * g, n0 + wn * 64 + b * 16 + fr, acc[a][b]);
  } else {
#pragma unroll
    for (int a = 0; a < 4; a++) {
      const int i0 = m0 + wm * 64 + a * 16 + 4 * g;
      double v[4] = {0, 0, 0, 0};
#pragma unroll
      for (int b = 0; b < 4; b++) {
        const int r = n0 + wn * 64 + b * 16 + fr;
        if (r >= ep.R) continue;
#pragma unroll
        for (int j = 0; j < 4; j++)
          if (i0 + j < ep.N) v[j] += (double)ep.P[(size_t)r * ep.N + i0 + j] * (double)acc[a][b][j];
      }
      like_pc_flush(ep, i0, v);
    }
  }
}

void launch_estep(int precision, bool ksplit, const void* K, const void* Klo, const void* S,
                  const void* Slo, int N_pad, int R_pad, int n_mp, const LikeEpilogue& ep,
                  hipStream_t st) {
  if (precision == PREC_FP16) {
    launch_estep16(ksplit, K, Klo, S, Slo, N_pad, R_pad, n_mp, ep, st);
    return;
  }
  const int tiles_m = N_pad / BM, tiles_n = R_pad / BN;
  const dim3 grid(tiles_m * tiles_n), block(256);
  if (ep.P)
    hipLaunchKernelGGL(k_estep_f32<true>, grid, block, 0, st, (const float*)K, (const float*)S, n_mp, tiles_m, tiles_n, ep);
  else
    hipLaunchKernelGGL(k_estep_f32<false>, grid, block, 0, st, (const float*)K, (const float*)S, n_mp, tiles_m, tiles_n, ep);
  EMC_LAUNCHED();
}

// ---------------------------------------------------------------------------
// FP32 M-step
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_mstep_f32(const float* __restrict__ Kt, int tiles_m, int tiles_n, MstepArgs a) {
  __shared__ __attribute__((aligned(16))) float sA[2][BM][LD32];
  __shared__ __attribute__((aligned(16))) float sB[2][BN][LD32];
  int tm, tn;
  tile_of_block(tiles_m, tiles_n, tm, tn);
  const int m0 = tm * BM, n0 = tn * BN;  // rotations, pixels
  const int t = threadIdx.x, lane = t & 63, w = t >> 6, wm = w >> 1, wn = w & 1;
  const int lr = t >> 2, lc = (t & 3) * 4;
  const float* gB = Kt + (size_t)(n0 + lr) * a.N_pad + lc;
  const size_t step64 = (size_t)64 * a.N_pad;

  floatx4 acc[4][4];
#pragma unroll
  for (int x = 0; x < 4; x++)
#pragma unroll
    for (int y = 0; y < 4; y++) acc[x][y] = floatx4{0, 0, 0, 0};

  floatx4 ra0 = load_p4(a, m0 + lr, lc, 1.0f), ra1 = load_p4(a, m0 + lr + 64, lc, 1.0f);
  floatx4 rb0 = *reinterpret_cast<const floatx4*>(gB), rb1 = *reinterpret_cast<const floatx4*>(gB + step64);
  *reinterpret_cast<floatx4*>(&sA[0][lr][lc]) = ra0;
  *reinterpret_cast<floatx4*>(&sA[0][lr + 64][lc]) = ra1;
  *reinterpret_cast<floatx4*>(&sB[0][lr][lc]) = rb0;
  *reinterpret_cast<floatx4*>(&sB[0][lr + 64][lc]) = rb1;
  __syncthreads();

  const int nk = a.N_pad / BK32;
  const int fr = lane & 15, fk = (lane >> 4) * 4;
  for (int kt = 0; kt < nk; kt++) {
    const int buf = kt & 1;
    const bool more = kt + 1 < nk;
    if (more) {
      const int i = (kt + 1) * BK32 + lc;
      ra0 = load_p4(a, m0 + lr, i, 1.0f);
      ra1 = load_p4(a, m0 + lr + 64, i, 1.0f);
      const size_t off = (size_t)(kt + 1) * BK32;
      rb0 = *reinterpret_cast<const floatx4*>(gB + off);
      rb1 = *reinterpret_cast<const floatx4*>(gB + off + step64);
    }
    floatx4 fa[4], fb[4];
#pragma unroll
    for (int x = 0; x < 4; x++) fa[x] = *reinterpret_cast<const floatx4*>(&sA[buf][wm * 64 + x * 16 + fr][fk]);
#pragma unroll
    for (int y = 0; y < 4; y++) fb[y] = *reinterpret_cast<const floatx4*>(&sB[buf][wn * 64 + y * 16 + fr][fk]);
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
      for (int x = 0; x < 4; x++)
#pragma unroll
        for (int y = 0; y < 4; y++) acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[x][s], fb[y][s], acc[x][y], 0, 0, 0);
    if (more) {
      const int nb = buf ^ 1;
      *reinterpret_cast<floatx4*>(&sA[nb][lr][lc]) = ra0;
      *reinterpret_cast<floatx4*>(&sA[nb][lr + 64][lc]) = ra1;
      *reinterpret_cast<floatx4*>(&sB[nb][lr][lc]) = rb0;
      *reinterpret_cast<floatx4*>(&sB[nb][lr + 64][lc]) = rb1;
    }
    __syncthreads();
  }
  const int g = lane >> 4;
#pragma unroll
  for (int x = 0; x < 4; x++)
#pragma unroll
    for (int y = 0; y < 4; y++) mstep_store(a, m0 + wm * 64 + x * 16 + 4 * g, n0 + wn * 64 + y * 16 + fr, acc[x][y], 1.0f);
}

void launch_mstep32(const float* Kt, const MstepArgs& a, hipStream_t st) {
  const int tiles_m = a.R_pad / BM, tiles_n = a.n_mp / BN;  // n_mp is a multiple of BN
  const dim3 grid(tiles_m * tiles_n), block(256);
  hipLaunchKernelGGL(k_mstep_f32, grid, block, 0, st, Kt, tiles_m, tiles_n, a);
  EMC_LAUNCHED();
}

}  // namespace emc
