// FP16 matrix-core contractions (precision form) of the E-step and M-step.
//
// C[m][n] = sum_k A[m][k] B[n][k] with both operands row-major along k.  The
// operand that is not exact in FP16 is carried as a hi + lo pair
// (x = hi + lo to ~22 bits); the products hi*hi + hi*lo (+ lo*hi when both
// sides are split) are accumulated in FP32 by v_mfma_f32_16x16x32_f16.
//
// Workgroup: 4 wave64s in a 2 x 2 grid; wave tile (16 TM) x (16 TN); the
// split operand sits on the 128-wide side so each stage stages 256 + 2 x 128
// (or 2 x 128 + 256) rows of 32 halves: 64 MFMAs per wave per stage against
// 32 KB of global reads.  LDS double buffered (80 KB, two workgroups per
// CU), next stage's global loads in flight during the current MFMAs.
#include <hip/hip_runtime.h>

#include "engine.hpp"
#include "gemm_common.hpp"

namespace emc {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef _Float16 halfx8 __attribute__((ext_vector_type(8)));
typedef _Float16 halfx4 __attribute__((ext_vector_type(4)));

constexpr int BK = 32;      // halves per stage along k
constexpr int LDK = BK + 8;  // padded LDS row pitch (80 bytes: conflict-free b128 reads)

// Plain FP16 rows: global [rows][ld] (+ optional lo copy) -> LDS.
template <int ROWS, int NS>
struct RowsF16 {
  static constexpr int PER = ROWS / 64;  // 16-byte chunks per thread per split
  const _Float16* g[NS];
  uint4 r[NS][PER];
  int lr, lc;
  __device__ void init(const _Float16* hi, const _Float16* lo, size_t row0, size_t ld) {
    lr = threadIdx.x >> 2;
    lc = (threadIdx.x & 3) * 8;
    g[0] = hi + (row0 + lr) * ld + lc;
    if (NS > 1) g[NS - 1] = lo + (row0 + lr) * ld + lc;
    ldm = ld;
  }
  size_t ldm;
  __device__ __forceinline__ void load(int k0) {
#pragma unroll
    for (int s = 0; s < NS; s++)
#pragma unroll
      for (int j = 0; j < PER; j++) r[s][j] = *reinterpret_cast<const uint4*>(g[s] + (size_t)j * 64 * ldm + k0);
  }
  __device__ __forceinline__ void store(_Float16 (*lds)[ROWS][LDK]) {
#pragma unroll
    for (int s = 0; s < NS; s++)
#pragma unroll
      for (int j = 0; j < PER; j++) *reinterpret_cast<uint4*>(&lds[s][lr + 64 * j][lc]) = r[s][j];
  }
};

// M-step A operand: responsibilities P[r][i] (fp32), scaled by invphi[i] and
// a per-row power of two, split into hi/lo halves on the way to LDS.
template <int ROWS>
struct RowsP {
  static constexpr int PER = ROWS / 32;  // float4 per thread
  const MstepArgs* a;
  floatx4 r[PER];
  float rs[PER];
  int ar, ac, m0;
  __device__ void init(const MstepArgs* args, int row0) {
    a = args;
    m0 = row0;
    ar = threadIdx.x >> 3;
    ac = (threadIdx.x & 7) * 4;
#pragma unroll
    for (int j = 0; j < PER; j++) {
      const int row = m0 + ar + 32 * j;
      rs[j] = row < a->R ? a->rowscale[row] : 0.0f;
    }
  }
  __device__ __forceinline__ void load(int k0) {
#pragma unroll
    for (int j = 0; j < PER; j++) r[j] = load_p4(*a, m0 + ar + 32 * j, k0 + ac, rs[j]);
  }
  __device__ __forceinline__ void store(_Float16 (*lds)[ROWS][LDK]) {
#pragma unroll
    for (int j = 0; j < PER; j++) {
      halfx4 hi, lo;
#pragma unroll
      for (int c = 0; c < 4; c++) {
        hi[c] = (_Float16)r[j][c];
        lo[c] = (_Float16)(r[j][c] - (float)hi[c]);
      }
      *reinterpret_cast<halfx4*>(&lds[0][ar + 32 * j][ac]) = hi;
      *reinterpret_cast<halfx4*>(&lds[1][ar + 32 * j][ac]) = lo;
    }
  }
};

// Main loop.  NA / NB: splits of A / B (1 or 2).  Returns the accumulators.
template <int TM, int TN, int NA, int NB, class LA, class LB>
__device__ __forceinline__ void mma_loop(LA& la, LB& lb, int nk, floatx4 (&acc)[TM][TN],
                                         _Float16 (*sA)[NA][32 * TM][LDK], _Float16 (*sB)[NB][32 * TN][LDK]) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, wm = w >> 1, wn = w & 1;
  const int fr = lane & 15, fk = (lane >> 4) * 8;
#pragma unroll
  for (int x = 0; x < TM; x++)
#pragma unroll
    for (int y = 0; y < TN; y++) acc[x][y] = floatx4{0, 0, 0, 0};
  la.load(0);
  lb.load(0);
  la.store(sA[0]);
  lb.store(sB[0]);
  __syncthreads();
  for (int kt = 0; kt < nk; kt++) {
    const int buf = kt & 1;
    const bool more = kt + 1 < nk;
    if (more) {
      la.load((kt + 1) * BK);
      lb.load((kt + 1) * BK);
    }
    halfx8 fb[NB][TN];
#pragma unroll
    for (int s = 0; s < NB; s++)
#pragma unroll
      for (int y = 0; y < TN; y++)
        fb[s][y] = *reinterpret_cast<const halfx8*>(&sB[buf][s][wn * 16 * TN + y * 16 + fr][fk]);
#pragma unroll
    for (int x = 0; x < TM; x++) {
      halfx8 fa[NA];
#pragma unroll
      for (int s = 0; s < NA; s++)
        fa[s] = *reinterpret_cast<const halfx8*>(&sA[buf][s][wm * 16 * TM + x * 16 + fr][fk]);
#pragma unroll
      for (int y = 0; y < TN; y++) {
        acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_f16(fa[0], fb[0][y], acc[x][y], 0, 0, 0);
        if (NB > 1) acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_f16(fa[0], fb[NB - 1][y], acc[x][y], 0, 0, 0);
        if (NA > 1) acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_f16(fa[NA - 1], fb[0][y], acc[x][y], 0, 0, 0);
      }
    }
    if (more) {
      la.store(sA[buf ^ 1]);
      lb.store(sB[buf ^ 1]);
    }
    __syncthreads();
  }
}

// E-step: A = packed counts (rows = patterns, 256 per tile), B = packed
// slices hi/lo (rows = rotations, 128 per tile).
template <bool KSPLIT, bool PC>
__global__ __launch_bounds__(256, 2) void k_estep16(const _Float16* __restrict__ K, const _Float16* __restrict__ Klo,
                                                 const _Float16* __restrict__ S, const _Float16* __restrict__ Slo,
                                                 int n_mp, int tiles_m, int tiles_n, LikeEpilogue ep) {
  constexpr int TM = 8, TN = 4, NA = KSPLIT ? 2 : 1, NB = 2;
  constexpr int BMM = 32 * TM, BNN = 32 * TN;
  __shared__ __attribute__((aligned(16))) _Float16 sA[2][NA][BMM][LDK];
  __shared__ __attribute__((aligned(16))) _Float16 sB[2][NB][BNN][LDK];
  int tm, tn;
  tile_of_block(tiles_m, tiles_n, tm, tn);
  const int m0 = tm * BMM, n0 = tn * BNN;
  RowsF16<BMM, NA> la;
  RowsF16<BNN, NB> lb;
  la.init(K, Klo, (size_t)m0, (size_t)n_mp);
  lb.init(S, Slo, (size_t)n0, (size_t)n_mp);
  floatx4 acc[TM][TN];
  mma_loop<TM, TN, NA, NB>(la, lb, n_mp / BK, acc, sA, sB);

  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, wm = w >> 1, wn = w & 1;
  const int g = lane >> 4, fr = lane & 15;
  if (!PC) {
#pragma unroll
    for (int x = 0; x < TM; x++)
#pragma unroll
      for (int y = 0; y < TN; y++)
        like_store(ep, m0 + wm * 16 * TM + x * 16 + 4 * g, n0 + wn * 16 * TN + y * 16 + fr, acc[x][y]);
  } else {
#pragma unroll
    for (int x = 0; x < TM; x++) {
      const int i0 = m0 + wm * 16 * TM + x * 16 + 4 * g;
      double v[4] = {0, 0, 0, 0};
#pragma unroll
      for (int y = 0; y < TN; y++) {
        const int r = n0 + wn * 16 * TN + y * 16 + fr;
        if (r >= ep.R) continue;
        const double cs = (double)ep.cscale[r] * (double)ep.kscale_inv;
#pragma unroll
        for (int j = 0; j < 4; j++)
          if (i0 + j < ep.N) v[j] += (double)ep.P[(size_t)r * ep.N + i0 + j] * (double)acc[x][y][j] * cs;
      }
      like_pc_flush(ep, i0, v);
    }
  }
}

// M-step: A = P/phi hi/lo (rows = rotations, 128 per tile), B = transposed
// packed counts (rows = pixels, 256 per tile).
template <bool KSPLIT>
__global__ __launch_bounds__(256, 2) void k_mstep16(const _Float16* __restrict__ Kt, const _Float16* __restrict__ Ktlo,
                                                 int tiles_m, int tiles_n, MstepArgs a) {
  constexpr int TM = 4, TN = 8, NA = 2, NB = KSPLIT ? 2 : 1;
  constexpr int BMM = 32 * TM, BNN = 32 * TN;
  __shared__ __attribute__((aligned(16))) _Float16 sA[2][NA][BMM][LDK];
  __shared__ __attribute__((aligned(16))) _Float16 sB[2][NB][BNN][LDK];
  int tm, tn;
  tile_of_block(tiles_m, tiles_n, tm, tn);
  const int m0 = tm * BMM, n0 = tn * BNN;
  RowsP<BMM> la;
  RowsF16<BNN, NB> lb;
  la.init(&a, m0);
  lb.init(Kt, Ktlo, (size_t)n0, (size_t)a.N_pad);
  floatx4 acc[TM][TN];
  mma_loop<TM, TN, NA, NB>(la, lb, a.N_pad / BK, acc, sA, sB);

  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, wm = w >> 1, wn = w & 1;
  const int g = lane >> 4, fr = lane & 15;
#pragma unroll
  for (int x = 0; x < TM; x++)
#pragma unroll
    for (int y = 0; y < TN; y++)
      mstep_store(a, m0 + wm * 16 * TM + x * 16 + 4 * g, n0 + wn * 16 * TN + y * 16 + fr, acc[x][y], a.kscale_inv);
}

void launch_estep16(bool ksplit, const void* K, const void* Klo, const void* S, const void* Slo, int N_pad,
                    int R_pad, int n_mp, const LikeEpilogue& ep, hipStream_t st) {
  const int tiles_m = N_pad / kTileM16, tiles_n = R_pad / kTileN;
  const dim3 grid(tiles_m * tiles_n), block(256);
  const bool pc = ep.P != nullptr;
  const _Float16 *k = (const _Float16*)K, *kl = (const _Float16*)Klo, *s = (const _Float16*)S,
                 *sl = (const _Float16*)Slo;
  if (ksplit) {
    if (pc) hipLaunchKernelGGL((k_estep16<true, true>), grid, block, 0, st, k, kl, s, sl, n_mp, tiles_m, tiles_n, ep);
    else hipLaunchKernelGGL((k_estep16<true, false>), grid, block, 0, st, k, kl, s, sl, n_mp, tiles_m, tiles_n, ep);
  } else {
    if (pc) hipLaunchKernelGGL((k_estep16<false, true>), grid, block, 0, st, k, kl, s, sl, n_mp, tiles_m, tiles_n, ep);
    else hipLaunchKernelGGL((k_estep16<false, false>), grid, block, 0, st, k, kl, s, sl, n_mp, tiles_m, tiles_n, ep);
  }
  EMC_LAUNCHED();
}

void launch_mstep16(bool ksplit, const void* Kt, const void* Ktlo, const MstepArgs& a, hipStream_t st) {
  const int tiles_m = a.R_pad / kTileN, tiles_n = a.n_mp / kPadK;
  const dim3 grid(tiles_m * tiles_n), block(256);
  if (ksplit)
    hipLaunchKernelGGL(k_mstep16<true>, grid, block, 0, st, (const _Float16*)Kt, (const _Float16*)Ktlo, tiles_m, tiles_n, a);
  else
    hipLaunchKernelGGL(k_mstep16<false>, grid, block, 0, st, (const _Float16*)Kt, (const _Float16*)Ktlo, tiles_m, tiles_n, a);
  EMC_LAUNCHED();
}

}  // namespace emc
