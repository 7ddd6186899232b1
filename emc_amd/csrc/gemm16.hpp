// FP16 matrix-core main loop shared by the dense and candidate-restricted
// contractions.
//
// C[m][n] = sum_k A[m][k] B[n][k] with both operands FP16 row-major along k.
// Operands reach LDS by direct global->LDS loads (global_load_lds_dwordx4):
// each wave-level load writes 1 KB = 16 rows of 32 halves; no staging
// registers.  STAGES LDS buffers keep STAGES-1 stages of loads in flight
// while one is consumed; one barrier per stage.  LDS rows are 64 B with the
// four 16-byte chunks of a row XOR-swizzled by the row index, so the 16-byte
// fragment reads of 16 consecutive rows hit distinct banks.
#pragma once

#include <hip/hip_runtime.h>

namespace emc {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef _Float16 halfx8 __attribute__((ext_vector_type(8)));
typedef _Float16 halfx4 __attribute__((ext_vector_type(4)));
#define EMC_LDS __attribute__((address_space(3)))

constexpr int BK = 32;  // halves per stage along k (one MFMA k-step)

__device__ __forceinline__ int chunk_swz(int row) { return ((row >> 1) ^ (row >> 2)) & 3; }

template <int N>
__device__ __forceinline__ void wait_vmcnt() {
  asm volatile("s_waitcnt vmcnt(%0)" ::"n"(N) : "memory");
}

// Operand description for the main loop: NA splits of A (rows m0..m0+BM),
// NB splits of B (rows n0..n0+BN), row pitch ld (halves).  With a row list
// (arow / brow), tile row m reads global row arow[m] for m < arow_n and the
// zero row otherwise (m0 / n0 are then ignored).
struct Operands {
  const _Float16* a[2];
  const _Float16* b[2];
  size_t lda, ldb;
  const int* arow = nullptr;
  int arow_n = 0;
  const int* brow = nullptr;
  int brow_n = 0;
  const _Float16* zero = nullptr;
};

// Main loop.  WGM x WGN waves, each TM x TN blocks of 16x16; STAGES LDS
// buffers; nk steps of BK along k from k = kofs.  On return acc holds this
// wave's block of C.
template <int TM, int TN, int WGM, int WGN, int NA, int NB, int STAGES>
__device__ __forceinline__ void dl_mma_loop(const Operands& op, int m0, int n0, int nk, floatx4 (&acc)[TM][TN],
                                            _Float16* lds, int kofs = 0) {
  constexpr int BM = 16 * WGM * TM, BN = 16 * WGN * TN, NW = WGM * WGN;
  constexpr int ROWS = NA * BM + NB * BN;  // rows of 32 halves per stage
  constexpr int STAGE_HALVES = ROWS * 32;
  constexpr int PERW = ROWS / 16 / NW;     // 1 KB loads per wave per stage
  static_assert((ROWS / 16) % NW == 0, "stage rows must split evenly over waves");
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, wm = w / WGN, wn = w % WGN;

  // global source of each of this wave's loads: row r of the stage, chunk
  // (lane & 3) ^ swizzle(r) so that LDS slot lane holds the swizzled chunk
  const _Float16* src[PERW];
#pragma unroll
  for (int j = 0; j < PERW; j++) {
    const int row = (w * PERW + j) * 16 + (lane >> 2);
    const int c = (lane & 3) ^ chunk_swz(row);
    const _Float16* p;
    if (row < NA * BM) {
      const int s = row / BM, m = row - s * BM;
      if (op.arow) {
        p = m < op.arow_n ? op.a[s] + (size_t)op.arow[m] * op.lda : op.zero;
      } else {
        p = op.a[s] + (size_t)(m0 + m) * op.lda;
      }
    } else {
      const int rb = row - NA * BM;
      const int s = rb / BN, n = rb - s * BN;
      if (op.brow) {
        p = n < op.brow_n ? op.b[s] + (size_t)op.brow[n] * op.ldb : op.zero;
      } else {
        p = op.b[s] + (size_t)(n0 + n) * op.ldb;
      }
    }
    src[j] = p + c * 8 + kofs;
  }
  auto issue = [&](int stage, int k0) {
#pragma unroll
    for (int j = 0; j < PERW; j++)
      __builtin_amdgcn_global_load_lds((const void*)(src[j] + k0),
                                       (EMC_LDS void*)(lds + stage * STAGE_HALVES + (w * PERW + j) * 16 * 32), 16, 0, 0);
  };
#pragma unroll
  for (int x = 0; x < TM; x++)
#pragma unroll
    for (int y = 0; y < TN; y++) acc[x][y] = floatx4{0, 0, 0, 0};
#pragma unroll
  for (int s = 0; s < STAGES - 1; s++)
    if (s < nk) issue(s, s * BK);
  const int fr = lane & 15, g = lane >> 4;
  for (int kt = 0; kt < nk; kt++) {
    // this wave's loads of stage kt are complete (later stages may still be
    // in flight); the barrier makes every wave's part visible and also
    // guarantees that all waves finished reading the buffer refilled below
    if (kt + STAGES - 2 < nk) wait_vmcnt<(STAGES - 2) * PERW>(); else wait_vmcnt<0>();
    __builtin_amdgcn_s_barrier();
    if (kt + STAGES - 1 < nk) issue((kt + STAGES - 1) % STAGES, (kt + STAGES - 1) * BK);
    const _Float16* L = lds + (kt % STAGES) * STAGE_HALVES;
    halfx8 fb[NB][TN];
#pragma unroll
    for (int s = 0; s < NB; s++)
#pragma unroll
      for (int y = 0; y < TN; y++) {
        const int row = NA * BM + s * BN + wn * 16 * TN + y * 16 + fr;
        fb[s][y] = *reinterpret_cast<const halfx8*>(&L[row * 32 + ((g ^ chunk_swz(row)) * 8)]);
      }
#pragma unroll
    for (int x = 0; x < TM; x++) {
      halfx8 fa[NA];
#pragma unroll
      for (int s = 0; s < NA; s++) {
        const int row = s * BM + wm * 16 * TM + x * 16 + fr;
        fa[s] = *reinterpret_cast<const halfx8*>(&L[row * 32 + ((g ^ chunk_swz(row)) * 8)]);
      }
#pragma unroll
      for (int y = 0; y < TN; y++) {
        acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_f16(fa[0], fb[0][y], acc[x][y], 0, 0, 0);
        if (NB > 1) acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_f16(fa[0], fb[NB - 1][y], acc[x][y], 0, 0, 0);
        if (NA > 1) acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_f16(fa[NA - 1], fb[0][y], acc[x][y], 0, 0, 0);
      }
    }
  }
  wait_vmcnt<0>();
}

}  // namespace emc
