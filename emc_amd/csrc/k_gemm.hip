// Matrix-core contractions of the EMC iteration.
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
// arrangement, each wave 64 x 64 = 4 x 4 blocks of the 16x16 MFMA output.
// Operand tiles are staged through LDS, double buffered, with the next tile's
// global loads issued before the current tile's MFMAs.
//
// FP32 form: v_mfma_f32_16x16x4_f32.  FP16 form: v_mfma_f32_16x16x32_f16 with
// the inexact operand split into hi + lo halves (the exact integer counts stay
// a single FP16 operand), FP32 accumulation.
//
// MFMA 16x16 operand/result layout (wave64): lane l supplies row (l % 16) of
// A and column (l % 16) of B for the reduction slots of group g = l / 16;
// result element (4g + j, l % 16) lands in accumulator register j of lane l.
#include <hip/hip_runtime.h>

#include "engine.hpp"

namespace emc {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef _Float16 halfx8 __attribute__((ext_vector_type(8)));
typedef _Float16 halfx4 __attribute__((ext_vector_type(4)));

constexpr int BM = kTileM, BN = kTileN;

// Workgroup -> output tile.  Consecutive workgroup ids are dispatched round
// robin over the 8 XCDs, each with its own L2; remap so that every XCD works
// on a contiguous range of tiles, and walk that range in groups of 8 tile
// rows so neighbouring workgroups share operand panels in L2.
__device__ __forceinline__ void tile_of_block(int tiles_m, int tiles_n, int& tm, int& tn) {
  const int nt = tiles_m * tiles_n;
  const int bid = blockIdx.x;
  const int xcd = bid & 7, slot = bid >> 3;
  const int q = nt >> 3, rem = nt & 7;
  const int logical = xcd * q + min(xcd, rem) + slot;
  constexpr int G = 8;
  const int per_group = G * tiles_n;
  const int group = logical / per_group;
  const int first_m = group * G;
  const int gsize = min(tiles_m - first_m, G);
  const int in_group = logical - group * per_group;
  tm = first_m + in_group % gsize;
  tn = in_group / gsize;
}

// ---------------------------------------------------------------------------
// Epilogues
// ---------------------------------------------------------------------------

// E-step: 4 consecutive patterns (rows) x 1 rotation (column) per lane block.
__device__ __forceinline__ void like_store(const LikeEpilogue& ep, int i0, int r, floatx4 acc, bool pc) {
  if (r >= ep.R) return;
  const double cs = (double)(ep.cscale ? ep.cscale[r] : 1.0f) * (double)ep.kscale_inv;
  if (!pc) {
    const double a = ep.A[r];
    float out[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int i = i0 + j;
      out[j] = 0.0f;
      if (i < ep.N) {
        const double c = (double)acc[j] * cs;
        const double phi = (double)ep.phi[i];
        out[j] = (float)(-(a - 2.0 * c / phi + ep.B[i] / (phi * phi)) * ep.inv_2nms2);
      }
    }
    float* dst = ep.L + (size_t)r * ep.N + i0;
    if (ep.vecL && i0 + 3 < ep.N) {
      *reinterpret_cast<floatx4*>(dst) = floatx4{out[0], out[1], out[2], out[3]};
    } else {
#pragma unroll
      for (int j = 0; j < 4; j++)
        if (i0 + j < ep.N) dst[j] = out[j];
    }
  }
}

// PC mode accumulates sum_r P_ri C_ir into D[i]: reduce the lane's values over
// its columns (rotations) across the 16 lanes of a group, one atomic per row.
__device__ __forceinline__ void like_pc_flush(const LikeEpilogue& ep, int i0, double v[4]) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    double s = v[j];
#pragma unroll
    for (int o = 8; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((lane & 15) == 0 && i0 + j < ep.N) atomicAdd(&ep.D[i0 + j], s);
  }
}

// ---------------------------------------------------------------------------
// FP32 E-step
// ---------------------------------------------------------------------------
constexpr int BK32 = 16;
constexpr int LD32 = BK32 + 4;  // padded LDS row pitch (floats)

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
        like_store(ep, m0 + wm * 64 + a * 16 + 4 * g, n0 + wn * 64 + b * 16 + fr, acc[a][b], false);
  } else {
#pragma unroll
    for (int a = 0; a < 4; a++) {
      const int i0 = m0 + wm * 64 + a * 16 + 4 * g;
      double v[4] = {0, 0, 0, 0};
#pragma unroll
      for (int b = 0; b < 4; b++) {
        const int r = n0 + wn * 64 + b * 16 + fr;
        if (r >= ep.R) continue;
        const double cs = (double)(ep.cscale ? ep.cscale[r] : 1.0f) * (double)ep.kscale_inv;
#pragma unroll
        for (int j = 0; j < 4; j++)
          if (i0 + j < ep.N) v[j] += (double)ep.P[(size_t)r * ep.N + i0 + j] * (double)acc[a][b][j] * cs;
      }
      like_pc_flush(ep, i0, v);
    }
  }
}

// ---------------------------------------------------------------------------
// FP16 hi/lo E-step
// ---------------------------------------------------------------------------
constexpr int BK16 = 32;
constexpr int LD16 = BK16 + 8;  // padded LDS row pitch (halves)

template <bool KSPLIT, bool PC>
__global__ __launch_bounds__(256) void k_estep_f16(const _Float16* __restrict__ K, const _Float16* __restrict__ Klo,
                                                   const _Float16* __restrict__ S, const _Float16* __restrict__ Slo,
                                                   int n_mp, int tiles_m, int tiles_n, LikeEpilogue ep) {
  constexpr int NA = KSPLIT ? 2 : 1;
  __shared__ __attribute__((aligned(16))) _Float16 sA[2][NA][BM][LD16];
  __shared__ __attribute__((aligned(16))) _Float16 sB[2][2][BN][LD16];
  int tm, tn;
  tile_of_block(tiles_m, tiles_n, tm, tn);
  const int m0 = tm * BM, n0 = tn * BN;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6, wm = w >> 1, wn = w & 1;
  const int lr = t >> 2, lc = (t & 3) * 8;  // rows lr, lr+64; 8 halves at lc
  const size_t rowA = (size_t)(m0 + lr) * n_mp + lc, rowB = (size_t)(n0 + lr) * n_mp + lc;
  const size_t step64 = (size_t)64 * n_mp;
  const _Float16* gA[2] = {K + rowA, KSPLIT ? Klo + rowA : nullptr};
  const _Float16* gB[2] = {S + rowB, Slo + rowB};

  floatx4 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 4; b++) acc[a][b] = floatx4{0, 0, 0, 0};

  uint4 ra[NA][2], rb[2][2];
  auto gload = [&](size_t off) {
#pragma unroll
    for (int h = 0; h < NA; h++) {
      ra[h][0] = *reinterpret_cast<const uint4*>(gA[h] + off);
      ra[h][1] = *reinterpret_cast<const uint4*>(gA[h] + off + step64);
    }
#pragma unroll
    for (int h = 0; h < 2; h++) {
      rb[h][0] = *reinterpret_cast<const uint4*>(gB[h] + off);
      rb[h][1] = *reinterpret_cast<const uint4*>(gB[h] + off + step64);
    }
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int h = 0; h < NA; h++) {
      *reinterpret_cast<uint4*>(&sA[buf][h][lr][lc]) = ra[h][0];
      *reinterpret_cast<uint4*>(&sA[buf][h][lr + 64][lc]) = ra[h][1];
    }
#pragma unroll
    for (int h = 0; h < 2; h++) {
      *reinterpret_cast<uint4*>(&sB[buf][h][lr][lc]) = rb[h][0];
      *reinterpret_cast<uint4*>(&sB[buf][h][lr + 64][lc]) = rb[h][1];
    }
  };
  gload(0);
  lstore(0);
  __syncthreads();

  const int nk = n_mp / BK16;
  const int fr = lane & 15, fk = (lane >> 4) * 8;
  for (int kt = 0; kt < nk; kt++) {
    const int buf = kt & 1;
    const bool more = kt + 1 < nk;
    if (more) gload((size_t)(kt + 1) * BK16);
    halfx8 fa[NA][4], fb[2][4];
#pragma unroll
    for (int h = 0; h < NA; h++)
#pragma unroll
      for (int a = 0; a < 4; a++) fa[h][a] = *reinterpret_cast<const halfx8*>(&sA[buf][h][wm * 64 + a * 16 + fr][fk]);
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
      for (int b = 0; b < 4; b++) fb[h][b] = *reinterpret_cast<const halfx8*>(&sB[buf][h][wn * 64 + b * 16 + fr][fk]);
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
      for (int b = 0; b < 4; b++) {
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(fa[0][a], fb[0][b], acc[a][b], 0, 0, 0);
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(fa[0][a], fb[1][b], acc[a][b], 0, 0, 0);
        if (KSPLIT) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(fa[NA - 1][a], fb[0][b], acc[a][b], 0, 0, 0);
      }
    if (more) lstore(buf ^ 1);
    __syncthreads();
  }

  const int g = lane >> 4;
  if (!PC) {
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
      for (int b = 0; b < 4; b++)
        like_store(ep, m0 + wm * 64 + a * 16 + 4 * g, n0 + wn * 64 + b * 16 + fr, acc[a][b], false);
  } else {
#pragma unroll
    for (int a = 0; a < 4; a++) {
      const int i0 = m0 + wm * 64 + a * 16 + 4 * g;
      double v[4] = {0, 0, 0, 0};
#pragma unroll
      for (int b = 0; b < 4; b++) {
        const int r = n0 + wn * 64 + b * 16 + fr;
        if (r >= ep.R) continue;
        const double cs = (double)ep.cscale[r] * (double)ep.kscale_inv;
#pragma unroll
        for (int j = 0; j < 4; j++)
          if (i0 + j < ep.N) v[j] += (double)ep.P[(size_t)r * ep.N + i0 + j] * (double)acc[a][b][j] * cs;
      }
      like_pc_flush(ep, i0, v);
    }
  }
}

void launch_estep(int precision, bool ksplit, const void* K, const void* Klo, const void* S,
                  const void* Slo, int N_pad, int R_pad, int n_mp, const LikeEpilogue& ep,
                  hipStream_t st) {
  const int tiles_m = N_pad / BM, tiles_n = R_pad / BN;
  const dim3 grid(tiles_m * tiles_n), block(256);
  const bool pc = ep.P != nullptr;
  if (precision == PREC_FP32) {
    if (pc)
      hipLaunchKernelGGL(k_estep_f32<true>, grid, block, 0, st, (const float*)K, (const float*)S, n_mp, tiles_m, tiles_n, ep);
    else
      hipLaunchKernelGGL(k_estep_f32<false>, grid, block, 0, st, (const float*)K, (const float*)S, n_mp, tiles_m, tiles_n, ep);
  } else {
    const _Float16 *k = (const _Float16*)K, *kl = (const _Float16*)Klo, *s = (const _Float16*)S,
                   *sl = (const _Float16*)Slo;
    if (ksplit) {
      if (pc) hipLaunchKernelGGL((k_estep_f16<true, true>), grid, block, 0, st, k, kl, s, sl, n_mp, tiles_m, tiles_n, ep);
      else hipLaunchKernelGGL((k_estep_f16<true, false>), grid, block, 0, st, k, kl, s, sl, n_mp, tiles_m, tiles_n, ep);
    } else {
      if (pc) hipLaunchKernelGGL((k_estep_f16<false, true>), grid, block, 0, st, k, kl, s, sl, n_mp, tiles_m, tiles_n, ep);
      else hipLaunchKernelGGL((k_estep_f16<false, false>), grid, block, 0, st, k, kl, s, sl, n_mp, tiles_m, tiles_n, ep);
    }
  }
  EMC_LAUNCHED();
}

// ---------------------------------------------------------------------------
// M-step
// ---------------------------------------------------------------------------

// Row r of the responsibilities, 4 patterns from i: P_ri * invphi_i * rs,
// zero outside the matrix.
__device__ __forceinline__ floatx4 load_p4(const MstepArgs& a, int r, int i, float rs) {
  floatx4 v = {0, 0, 0, 0};
  if (r >= a.R) return v;
  const float* row = a.P + (size_t)r * a.N;
  if (a.vecP) {
    if (i < a.N) {
      v = *reinterpret_cast<const floatx4*>(row + i);
      const floatx4 ip = *reinterpret_cast<const floatx4*>(a.invphi + i);
      v = v * ip * rs;
    }
  } else {
#pragma unroll
    for (int c = 0; c < 4; c++)
      if (i + c < a.N) v[c] = row[i + c] * a.invphi[i + c] * rs;
  }
  return v;
}

__device__ __forceinline__ void mstep_store(const MstepArgs& a, int r0, int k, floatx4 acc, float extra) {
  if (k >= a.n_m) return;
  const int p = a.pix[k];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int r = r0 + j;
    if (r >= a.R) continue;
    float u = acc[j] * extra;
    if (a.rowscale) u = u / a.rowscale[r];
    const float T = a.Tf[r];
    const float s = T > 1e-10f ? u / T : u;
    a.slices[(size_t)r * a.N2d + p] = s;
    a.Sc[(size_t)r * a.n_mp + k] = s;
  }
}

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

template <bool KSPLIT>
__global__ __launch_bounds__(256) void k_mstep_f16(const _Float16* __restrict__ Kt, const _Float16* __restrict__ Ktlo,
                                                   int tiles_m, int tiles_n, MstepArgs a) {
  constexpr int NB = KSPLIT ? 2 : 1;
  __shared__ __attribute__((aligned(16))) _Float16 sA[2][2][BM][LD16];
  __shared__ __attribute__((aligned(16))) _Float16 sB[2][NB][BN][LD16];
  int tm, tn;
  tile_of_block(tiles_m, tiles_n, tm, tn);
  const int m0 = tm * BM, n0 = tn * BN;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6, wm = w >> 1, wn = w & 1;
  // A (P rows) staging: 128 rows x 32 patterns of fp32 = 4 float4 per thread
  const int ar = t >> 3, ac = (t & 7) * 4;  // rows ar + 32 j
  // B (Kt rows) staging: rows lr, lr + 64, 8 halves at lc
  const int lr = t >> 2, lc = (t & 3) * 8;
  const size_t rowB = (size_t)(n0 + lr) * a.N_pad + lc;
  const size_t step64 = (size_t)64 * a.N_pad;
  const _Float16* gB[2] = {Kt + rowB, KSPLIT ? Ktlo + rowB : nullptr};
  float rs[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int r = m0 + ar + 32 * j;
    rs[j] = r < a.R ? a.rowscale[r] : 0.0f;
  }

  floatx4 acc[4][4];
#pragma unroll
  for (int x = 0; x < 4; x++)
#pragma unroll
    for (int y = 0; y < 4; y++) acc[x][y] = floatx4{0, 0, 0, 0};

  floatx4 ra[4];
  uint4 rb[NB][2];
  auto gload = [&](int kbase) {
#pragma unroll
    for (int j = 0; j < 4; j++) ra[j] = load_p4(a, m0 + ar + 32 * j, kbase + ac, rs[j]);
#pragma unroll
    for (int h = 0; h < NB; h++) {
      rb[h][0] = *reinterpret_cast<const uint4*>(gB[h] + kbase);
      rb[h][1] = *reinterpret_cast<const uint4*>(gB[h] + kbase + step64);
    }
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int j = 0; j < 4; j++) {
      halfx4 hi, lo;
#pragma unroll
      for (int c = 0; c < 4; c++) {
        hi[c] = (_Float16)ra[j][c];
        lo[c] = (_Float16)(ra[j][c] - (float)hi[c]);
      }
      *reinterpret_cast<halfx4*>(&sA[buf][0][ar + 32 * j][ac]) = hi;
      *reinterpret_cast<halfx4*>(&sA[buf][1][ar + 32 * j][ac]) = lo;
    }
#pragma unroll
    for (int h = 0; h < NB; h++) {
      *reinterpret_cast<uint4*>(&sB[buf][h][lr][lc]) = rb[h][0];
      *reinterpret_cast<uint4*>(&sB[buf][h][lr + 64][lc]) = rb[h][1];
    }
  };
  gload(0);
  lstore(0);
  __syncthreads();

  const int nk = a.N_pad / BK16;
  const int fr = lane & 15, fk = (lane >> 4) * 8;
  for (int kt = 0; kt < nk; kt++) {
    const int buf = kt & 1;
    const bool more = kt + 1 < nk;
    if (more) gload((kt + 1) * BK16);
    halfx8 fa[2][4], fb[NB][4];
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
      for (int x = 0; x < 4; x++) fa[h][x] = *reinterpret_cast<const halfx8*>(&sA[buf][h][wm * 64 + x * 16 + fr][fk]);
#pragma unroll
    for (int h = 0; h < NB; h++)
#pragma unroll
      for (int y = 0; y < 4; y++) fb[h][y] = *reinterpret_cast<const halfx8*>(&sB[buf][h][wn * 64 + y * 16 + fr][fk]);
#pragma unroll
    for (int x = 0; x < 4; x++)
#pragma unroll
      for (int y = 0; y < 4; y++) {
        acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_f16(fa[0][x], fb[0][y], acc[x][y], 0, 0, 0);
        acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_f16(fa[1][x], fb[0][y], acc[x][y], 0, 0, 0);
        if (KSPLIT) acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_f16(fa[0][x], fb[NB - 1][y], acc[x][y], 0, 0, 0);
      }
    if (more) lstore(buf ^ 1);
    __syncthreads();
  }
  const int g = lane >> 4;
#pragma unroll
  for (int x = 0; x < 4; x++)
#pragma unroll
    for (int y = 0; y < 4; y++)
      mstep_store(a, m0 + wm * 64 + x * 16 + 4 * g, n0 + wn * 64 + y * 16 + fr, acc[x][y], a.kscale_inv);
}

void launch_mstep(int precision, bool ksplit, const void* Kt, const void* Ktlo, const MstepArgs& a,
                  hipStream_t st) {
  const int tiles_m = a.R_pad / BM, tiles_n = a.n_mp / BN;  // n_mp is a multiple of BN
  const dim3 grid(tiles_m * tiles_n), block(256);
  if (precision == PREC_FP32) {
    hipLaunchKernelGGL(k_mstep_f32, grid, block, 0, st, (const float*)Kt, tiles_m, tiles_n, a);
  } else if (ksplit) {
    hipLaunchKernelGGL(k_mstep_f16<true>, grid, block, 0, st, (const _Float16*)Kt, (const _Float16*)Ktlo, tiles_m, tiles_n, a);
  } else {
    hipLaunchKernelGGL(k_mstep_f16<false>, grid, block, 0, st, (const _Float16*)Kt, (const _Float16*)Ktlo, tiles_m, tiles_n, a);
  }
  EMC_LAUNCHED();
}

}  // namespace emc
