// FP16 matrix-core contractions (precision form) of the E-step and M-step.
//
// C[m][n] = sum_k A[m][k] B[n][k] with both operands FP16 row-major along k.
// The operand that is not exact in FP16 is carried as a hi + lo pair
// (x = hi + lo to ~22 bits); hi*hi + hi*lo (+ lo*hi when both sides are
// split) accumulate in FP32 through v_mfma_f32_16x16x32_f16.
//
// Workgroup: 256 x 256 output tile, 16 wave64s; wave tile 128 x 32 (E-step,
// 2 x 8 waves) or 64 x 64 (M-step, 4 x 4 waves).  Operands reach LDS by
// direct global->LDS loads (global_load_lds_dwordx4): each wave-level load
// writes 1 KB = 16 rows of 32 halves; no staging registers.  Three LDS stages
// (48 KB each) keep two stages of loads in flight while the third is
// consumed; one barrier per stage.  LDS rows are 64 B with the four 16-byte
// chunks of a row XOR-swizzled by the row index, so the 16-byte fragment
// reads of 16 consecutive rows hit distinct banks.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "engine.hpp"
#include "gemm16.hpp"
#include "gemm_common.hpp"

namespace emc {

// ---------------------------------------------------------------------------
// E-step: A = packed counts (256 patterns), B = packed slices hi/lo (256
// rotations).  KSPLIT: counts not exact in FP16, A carried hi/lo too (then
// two stages to stay within LDS).
// ---------------------------------------------------------------------------
// Tiles: the first n_main blocks take one whole tile each (XCD-aware order);
// the last tiles_m * tiles_n - n_main tiles (the tail, fewer than the device
// has compute units) are each split into `split` ranges along k, one block
// per range, whose FP32 sums go to `part`; k_estep16_tail adds them in range
// order and applies the epilogue.  One workgroup fills a compute unit (LDS),
// so without the split a tail of a few tiles (or a launch of fewer tiles
// than units) would take a whole tile time on a few units while the rest of
// the device idles.
// Tile shapes: 256 x 256 (16 waves of 128 x 32, one workgroup per compute
// unit, three 48 KB stages: the stage's 48 global->LDS loads spread over 16
// waves, 3 each, so each wave's loads land sooner -- the main loop waits on
// that latency, not on bandwidth; measured 9% faster than 8 waves of
// 128 x 64, which stays as option gemm_tile "256w8"); "128": 128 x 128 (4 waves
// of 64 x 64, 72 KB of LDS, two independent workgroups per compute unit, so
// one computes while the other waits at its stage barrier); "256x128":
// 8 waves of 128 x 32, four 32 KB stages (three stages of loads in flight).
template <int TM, int TN, int WGM, int WGN, int MINB, int ST, bool KSPLIT, bool PC>
__global__ __launch_bounds__(64 * WGM * WGN, MINB) void k_estep16(Operands op, int n_mp, int tiles_m, int tiles_n,
                                                                  int n_main, int split, float* __restrict__ part,
                                                                  LikeEpilogue ep) {
  constexpr int NA = KSPLIT ? 2 : 1, NB = 2, STAGES = KSPLIT ? 2 : ST, NT = 64 * WGM * WGN;
  constexpr int BM = 16 * WGM * TM, BN = 16 * WGN * TN;
  __shared__ __attribute__((aligned(1024))) _Float16 lds[STAGES * (NA * BM + NB * BN) * 32];
  int tm, tn, logical, range = -1, k0 = 0, k1 = n_mp / BK;
  if ((int)blockIdx.x < n_main) {
    logical = logical_of_block(n_main);
  } else {
    const int u = (int)blockIdx.x - n_main, nk = n_mp / BK;
    logical = n_main + u / split;
    range = u % split;
    k0 = range * nk / split;
    k1 = (range + 1) * nk / split;
  }
  tile_of_logical(logical, tiles_m, tiles_n, tm, tn);
  const int m0 = tm * BM, n0 = tn * BN;
  floatx4 acc[TM][TN];
  dl_mma_loop<TM, TN, WGM, WGN, NA, NB, STAGES>(op, m0, n0, k1 - k0, acc, lds, k0 * BK);
  if (range >= 0) {
    // [tail tile][range][x][y][thread] float4, coalesced over threads
    float* out = part + ((size_t)(logical - n_main) * split + range) * (TM * TN * NT * 4);
#pragma unroll
    for (int x = 0; x < TM; x++)
#pragma unroll
      for (int y = 0; y < TN; y++)
        reinterpret_cast<floatx4*>(out)[(x * TN + y) * NT + threadIdx.x] = acc[x][y];
    return;
  }

  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, wm = w / WGN, wn = w % WGN;
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

// The split tail tiles: blocks (tile, x), ranges added in order.
template <int TM, int TN, int WGM, int WGN>
__global__ __launch_bounds__(64 * WGM * WGN) void k_estep16_tail(const float* __restrict__ part, int tiles_m,
                                                                 int tiles_n, int n_main, int split, LikeEpilogue ep) {
  constexpr int NT = 64 * WGM * WGN;
  const int logical = n_main + (int)blockIdx.x;
  int tm, tn;
  tile_of_logical(logical, tiles_m, tiles_n, tm, tn);
  const int m0 = tm * 16 * WGM * TM, n0 = tn * 16 * WGN * TN;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, wm = w / WGN, wn = w % WGN;
  const int g = lane >> 4, fr = lane & 15;
  const floatx4* in = reinterpret_cast<const floatx4*>(part) + (size_t)blockIdx.x * split * (TM * TN * NT);
  const int x = blockIdx.y;
#pragma unroll
  for (int y = 0; y < TN; y++) {
    floatx4 acc = in[(x * TN + y) * NT + threadIdx.x];
    for (int r = 1; r < split; r++) acc += in[((size_t)r * TM * TN + x * TN + y) * NT + threadIdx.x];
    like_store(ep, m0 + wm * 16 * TM + x * 16 + 4 * g, n0 + wn * 16 * TN + y * 16 + fr, acc);
  }
}

// ---------------------------------------------------------------------------
// M-step: A = (P/phi * rowscale) hi/lo (256 rotations), B = transposed packed
// counts (256 pixels), reduction over patterns.
// ---------------------------------------------------------------------------
template <bool KSPLIT>
__global__ __launch_bounds__(1024, 1) void k_mstep16(Operands op, int tiles_m, int tiles_n, MstepArgs a) {
  constexpr int TM = 4, TN = 4, WGM = 4, WGN = 4, NA = 2, NB = KSPLIT ? 2 : 1, STAGES = KSPLIT ? 2 : 3;
  constexpr int BM = 16 * WGM * TM, BN = 16 * WGN * TN;
  __shared__ __attribute__((aligned(1024))) _Float16 lds[STAGES * (NA * BM + NB * BN) * 32];
  int tm, tn;
  tile_of_block(tiles_m, tiles_n, tm, tn);
  const int m0 = tm * BM, n0 = tn * BN;
  floatx4 acc[TM][TN];
  dl_mma_loop<TM, TN, WGM, WGN, NA, NB, STAGES>(op, m0, n0, a.N_pad / BK, acc, lds);

  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, wm = w / WGN, wn = w % WGN;
  const int g = lane >> 4, fr = lane & 15;
#pragma unroll
  for (int x = 0; x < TM; x++)
#pragma unroll
    for (int y = 0; y < TN; y++)
      mstep_store(a, m0 + wm * 16 * TM + x * 16 + 4 * g, n0 + wn * 16 * TN + y * 16 + fr, acc[x][y], a.kscale_inv);
}

// P'[r][i] = P_ri / phi_i * rowscale_r split into hi/lo halves, zero padded to
// [R_pad][N_pad] (the M-step's A operand).
__global__ __launch_bounds__(256) void k_convert_p(const float* __restrict__ P, int R, int N, int N_pad,
                                                   const float* __restrict__ invphi, const float* __restrict__ rowscale,
                                                   int vecP, _Float16* __restrict__ hi, _Float16* __restrict__ lo) {
  const int r = blockIdx.x;
  const float rs = r < R ? rowscale[r] : 0.0f;
  const float* row = P + (size_t)r * N;
  for (int i = threadIdx.x * 4; i < N_pad; i += 256 * 4) {
    floatx4 v = {0, 0, 0, 0};
    if (r < R) {
      if (vecP && i + 3 < N) {
        v = *reinterpret_cast<const floatx4*>(row + i) * *reinterpret_cast<const floatx4*>(invphi + i) * rs;
      } else {
#pragma unroll
        for (int c = 0; c < 4; c++)
          if (i + c < N) v[c] = row[i + c] * invphi[i + c] * rs;
      }
    }
    halfx4 h, l;
#pragma unroll
    for (int c = 0; c < 4; c++) {
      h[c] = (_Float16)v[c];
      l[c] = (_Float16)(v[c] - (float)h[c]);
    }
    *reinterpret_cast<halfx4*>(hi + (size_t)r * N_pad + i) = h;
    *reinterpret_cast<halfx4*>(lo + (size_t)r * N_pad + i) = l;
  }
}

void launch_convert_p(const MstepArgs& a, _Float16* hi, _Float16* lo, hipStream_t st) {
  hipLaunchKernelGGL(k_convert_p, dim3(a.R_pad), dim3(256), 0, st, a.P, a.R, a.N, a.N_pad, a.invphi, a.rowscale,
                     a.vecP, hi, lo);
  EMC_LAUNCHED();
}

template <int TM, int TN, int WGM, int WGN, int MINB, int ST>
static void estep16_shape(bool ksplit, const Operands& op, int N_pad, int R_pad, int n_mp, const LikeEpilogue& ep,
                          hipStream_t st) {
  constexpr int BM = 16 * WGM * TM, BN = 16 * WGN * TN, NT = 64 * WGM * WGN;
  const int tiles_m = N_pad / BM, tiles_n = R_pad / BN, nt = tiles_m * tiles_n;
  const bool pc = ep.P != nullptr;
  // MINB workgroups per compute unit (LDS): the tiles of the last, partial
  // round (all of them when there are fewer tiles than workgroup slots) are
  // split along k into the number of ranges that finishes them soonest,
  // ceil(tail * split / slots) rounds of 1 / split of a tile each
  Engine& e = engine();
  const int slots = e.n_cu * MINB;
  const int rounds_full = nt / slots, tail = nt - rounds_full * slots, nk = n_mp / BK;
  int split = 1, n_main = nt;
  if (e.tail_split && !pc && tail > 0) {
    double best = 1.0;
    for (int sp = 2; sp <= std::min(16, nk); sp++) {
      const double t = (double)((tail * sp + slots - 1) / slots) / sp;
      if (t < best - 1e-9) {
        best = t;
        split = sp;
      }
    }
    if (split > 1 && e.d_tail_part.ensure(sizeof(float) * (size_t)tail * split * BM * BN))
      n_main = nt - tail;
    else
      split = 1;
  }
  float* part = e.d_tail_part.as<float>();
  const dim3 grid(n_main + (nt - n_main) * split), block(NT);
  if (ksplit) {
    if (pc)
      hipLaunchKernelGGL((k_estep16<TM, TN, WGM, WGN, MINB, ST, true, true>), grid, block, 0, st, op, n_mp, tiles_m,
                         tiles_n, n_main, split, part, ep);
    else
      hipLaunchKernelGGL((k_estep16<TM, TN, WGM, WGN, MINB, ST, true, false>), grid, block, 0, st, op, n_mp, tiles_m,
                         tiles_n, n_main, split, part, ep);
  } else {
    if (pc)
      hipLaunchKernelGGL((k_estep16<TM, TN, WGM, WGN, MINB, ST, false, true>), grid, block, 0, st, op, n_mp, tiles_m,
                         tiles_n, n_main, split, part, ep);
    else
      hipLaunchKernelGGL((k_estep16<TM, TN, WGM, WGN, MINB, ST, false, false>), grid, block, 0, st, op, n_mp, tiles_m,
                         tiles_n, n_main, split, part, ep);
  }
  EMC_LAUNCHED();
  if (n_main < nt) {
    hipLaunchKernelGGL((k_estep16_tail<TM, TN, WGM, WGN>), dim3(nt - n_main, TM), dim3(NT), 0, st, part, tiles_m,
                       tiles_n, n_main, split, ep);
    EMC_LAUNCHED();
  }
}

void launch_estep16(bool ksplit, const void* K, const void* Klo, const void* S, const void* Slo, int N_pad,
                    int R_pad, int n_mp, const LikeEpilogue& ep, hipStream_t st) {
  Operands op{};
  op.a[0] = (const _Float16*)K;
  op.a[1] = (const _Float16*)Klo;
  op.b[0] = (const _Float16*)S;
  op.b[1] = (const _Float16*)Slo;
  op.lda = op.ldb = (size_t)n_mp;
  const int shape = engine().gemm_shape;
  if (shape == 1)
    estep16_shape<4, 4, 2, 2, 2, 3>(ksplit, op, N_pad, R_pad, n_mp, ep, st);
  else if (shape == 2 && R_pad % 128 == 0)
    estep16_shape<8, 2, 2, 4, 1, 4>(ksplit, op, N_pad, R_pad, n_mp, ep, st);
  else if (shape == 3)
    estep16_shape<8, 4, 2, 4, 1, 3>(ksplit, op, N_pad, R_pad, n_mp, ep, st);
  else
    estep16_shape<8, 2, 2, 8, 1, 3>(ksplit, op, N_pad, R_pad, n_mp, ep, st);
}

void launch_mstep16(bool ksplit, const void* Phi, const void* Plo, const void* Kt, const void* Ktlo,
                    const MstepArgs& a, hipStream_t st) {
  const int tiles_m = a.R_pad / kTile16, tiles_n = a.n_mp / kTile16;
  const dim3 grid(tiles_m * tiles_n), block(1024);
  Operands op{};
  op.a[0] = (const _Float16*)Phi;
  op.a[1] = (const _Float16*)Plo;
  op.b[0] = (const _Float16*)Kt;
  op.b[1] = (const _Float16*)Ktlo;
  op.lda = op.ldb = (size_t)a.N_pad;
  if (ksplit)
    hipLaunchKernelGGL(k_mstep16<true>, grid, block, 0, st, op, tiles_m, tiles_n, a);
  else
    hipLaunchKernelGGL(k_mstep16<false>, grid, block, 0, st, op, tiles_m, tiles_n, a);
  EMC_LAUNCHED();
}

}  // namespace emc
