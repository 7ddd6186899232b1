// Pattern groups of the candidate E-step.  Patterns whose candidate lists
// start with the same cell (with re-centring: the nearest free cell of the
// same best cell) share most of their candidate cells.  For such a group,
// the core cells (those among the candidates of at least 1 / core_div of its
// patterns, so at most core_div * topk of them; the kMaxCore lowest-numbered
// are kept and the number left out is counted in `trunc`, a proposal
// statistic) are scored as one dense block: every
// pattern of the group against every rotation of the core, gathered count
// rows and gathered slice rows through the FP16 main loop, each count row
// read once per 256 core rotations instead of once per candidate cell.  The
// block's values for (pattern, rotation) pairs that are candidates go to Lc;
// the others are dropped.  Candidate cells outside their group's core are
// scored by the per-cell kernel as before.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include "engine.hpp"
#include "gemm16.hpp"
#include "proposal_kernels.hpp"
#include "reduce.hpp"

namespace emc {

__global__ __launch_bounds__(256) void k_group_keys(const int* __restrict__ cand, int N, int topk, int C,
                                                    int* __restrict__ key, int* __restrict__ count) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= N) return;
  const int c = cand[(size_t)i * topk];
  key[i] = c;
  atomicAdd(&count[c], 1);
}

// hist[key][cell] += 1 for every candidate entry
__global__ __launch_bounds__(256) void k_group_hist(const int* __restrict__ cand, const int* __restrict__ key, int N,
                                                    int topk, int C, int* __restrict__ hist) {
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t >= (long)N * topk) return;
  const int i = (int)(t / topk);
  atomicAdd(&hist[(size_t)key[i] * C + cand[t]], 1);
}

// One workgroup per group key: its core cells in increasing order (held by
// at least 1 / core_div of the group's patterns; groups smaller than min_size get
// none), their count and fine-rotation total; qualifying cells beyond the
// first kMaxCore are left to the per-cell kernel and counted in *trunc.
__global__ __launch_bounds__(256) void k_group_core(const int* __restrict__ hist, const int* __restrict__ count,
                                                    int C, int min_size, int core_div,
                                                    const int* __restrict__ cell_off,
                                                    int* __restrict__ core, int* __restrict__ ncore,
                                                    int* __restrict__ nrot, int* __restrict__ trunc) {
  const int g = blockIdx.x;
  const int n = count[g];
  __shared__ int s_n, s_rot;
  if (threadIdx.x == 0) {
    s_n = 0;
    s_rot = 0;
  }
  __syncthreads();
  if (n >= min_size && n > 0) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __shared__ int s_wc[4];
    for (int c0 = 0; c0 < C; c0 += 256) {
      const int c = c0 + threadIdx.x;
      const bool in = c < C && core_div * hist[(size_t)g * C + c] >= n;
      const unsigned long long m = __ballot(in);
      if (lane == 0) s_wc[w] = __popcll(m);
      __syncthreads();
      int before = s_n;
      for (int k = 0; k < w; k++) before += s_wc[k];
      const int pos = before + __popcll(m & ((1ull << lane) - 1ull));
      if (in && pos < kMaxCore) {
        core[(size_t)g * kMaxCore + pos] = c;
        atomicAdd(&s_rot, cell_off[c + 1] - cell_off[c]);
      }
      __syncthreads();
      if (threadIdx.x == 0) s_n += s_wc[0] + s_wc[1] + s_wc[2] + s_wc[3];
      __syncthreads();
    }
  }
  if (threadIdx.x == 0) {
    ncore[g] = min(s_n, kMaxCore);
    nrot[g] = s_rot;
    if (s_n > kMaxCore) atomicAdd(trunc, s_n - kMaxCore);
  }
}

// One workgroup per group key with a core: its core rotations at cofs[g]:
// rotation, core index of its cell, slot in the cell.
__global__ __launch_bounds__(64) void k_group_rots(const int* __restrict__ core, const int* __restrict__ ncore,
                                                   const int* __restrict__ cofs, const int* __restrict__ cell_off,
                                                   const int* __restrict__ cell_rot, int* __restrict__ crot,
                                                   unsigned char* __restrict__ ccell, unsigned char* __restrict__ cslot) {
  const int g = blockIdx.x;
  const int q = threadIdx.x;
  if (q >= ncore[g]) return;
  int at = cofs[g];
  for (int k = 0; k < q; k++) {
    const int c = core[(size_t)g * kMaxCore + k];
    at += cell_off[c + 1] - cell_off[c];
  }
  const int c = core[(size_t)g * kMaxCore + q];
  for (int s = 0; s < cell_off[c + 1] - cell_off[c]; s++) {
    crot[at + s] = cell_rot[cell_off[c] + s];
    ccell[at + s] = (unsigned char)q;
    cslot[at + s] = (unsigned char)s;
  }
}

// One thread per pattern: rank[i][q] = rank of core cell q of its group in
// its candidate list (-1: not a candidate); rkey = the candidate's cell for
// the per-cell kernel, or C (no entry) when its group's block covers it.
__global__ __launch_bounds__(256) void k_group_rank(const int* __restrict__ cand, const int* __restrict__ key, int N,
                                                    int topk, int C, const int* __restrict__ core,
                                                    const int* __restrict__ ncore, signed char* __restrict__ rank,
                                                    int* __restrict__ rkey) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= N) return;
  signed char* ri = rank + (size_t)i * kMaxCore;
  for (int q = 0; q < kMaxCore; q++) ri[q] = -1;
  const int g = key[i], nq = ncore[g];
  const int* cg = core + (size_t)g * kMaxCore;
  for (int j = 0; j < topk; j++) {
    const int c = cand[(size_t)i * topk + j];
    int q = 0;
    while (q < nq && cg[q] != c) q++;
    if (q < nq) {
      ri[q] = (signed char)j;
      rkey[(size_t)i * topk + j] = C;
    } else {
      rkey[(size_t)i * topk + j] = c;
    }
  }
}

bool build_group_keys(const int* cand, int N, int topk, int C, int* key, int* count, int* gpat, DBuf& key_sorted,
                      DBuf& temp, hipStream_t st) {
  launch_fill(reinterpret_cast<float*>(count), (size_t)C, 0.0f, st);
  hipLaunchKernelGGL(k_group_keys, dim3((N + 255) / 256), dim3(256), 0, st, cand, N, topk, C, key, count);
  EMC_LAUNCHED();
  int bits = 1;
  while ((1 << bits) < C) bits++;
  if (!key_sorted.ensure(sizeof(int) * N)) return false;
  rocprim::counting_iterator<int> ids(0);
  size_t tbytes = 0;
  EMC_CHECK(rocprim::radix_sort_pairs(nullptr, tbytes, key, key_sorted.as<int>(), ids, gpat, (size_t)N, 0, bits, st));
  if (!temp.ensure(std::max(tbytes, (size_t)1))) return false;
  EMC_CHECK(rocprim::radix_sort_pairs(temp.p, tbytes, key, key_sorted.as<int>(), ids, gpat, (size_t)N, 0, bits, st));
  return !failed();
}

void launch_group_core(const int* cand, const int* key, const int* count, int N, int topk, int C, int min_size,
                       int core_div, const int* cell_off, int* hist, int* core, int* ncore, int* nrot, int* trunc,
                       hipStream_t st) {
  launch_fill(reinterpret_cast<float*>(hist), (size_t)C * C, 0.0f, st);
  launch_fill(reinterpret_cast<float*>(trunc), 1, 0.0f, st);
  const long n = (long)N * topk;
  hipLaunchKernelGGL(k_group_hist, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, cand, key, N, topk, C, hist);
  EMC_LAUNCHED();
  hipLaunchKernelGGL(k_group_core, dim3(C), dim3(256), 0, st, hist, count, C, min_size, core_div, cell_off, core, ncore,
                     nrot, trunc);
  EMC_LAUNCHED();
}

void launch_group_rots(const int* core, const int* ncore, const int* cofs, const int* cell_off, const int* cell_rot,
                       int C, int* crot, unsigned char* ccell, unsigned char* cslot, hipStream_t st) {
  hipLaunchKernelGGL(k_group_rots, dim3(C), dim3(64), 0, st, core, ncore, cofs, cell_off, cell_rot, crot, ccell, cslot);
  EMC_LAUNCHED();
}

void launch_group_rank(const int* cand, const int* key, int N, int topk, int C, const int* core, const int* ncore,
                       signed char* rank, int* rkey, hipStream_t st) {
  hipLaunchKernelGGL(k_group_rank, dim3((N + 255) / 256), dim3(256), 0, st, cand, key, N, topk, C, core, ncore, rank,
                     rkey);
  EMC_LAUNCHED();
}

// ---------------------------------------------------------------------------
// Dense block of a group: 256 patterns x 256 core rotations per workgroup
// (16 waves of 128 x 32, one workgroup per compute unit as in the dense
// GEMM): every
// gathered count row is read once per 256 core rotations and every slice row
// once per 256 patterns of the group.
// ---------------------------------------------------------------------------
template <bool KSPLIT>
__global__ __launch_bounds__(1024, 1) void k_group_estep(GroupEstepArgs a) {
  constexpr int TM = 8, TN = 2, WGM = 2, WGN = 8, NA = KSPLIT ? 2 : 1, NB = 2, STAGES = KSPLIT ? 2 : 3;
  constexpr int BM = 16 * WGM * TM, BN = 16 * WGN * TN;
  static_assert(BM == kGroupTile && BN == kGroupTile, "tile must match the work items");
  __shared__ __attribute__((aligned(1024))) _Float16 lds[STAGES * (NA * BM + NB * BN) * 32];
  const int4 w = a.work[blockIdx.x];  // first pattern, pattern count, first core rotation, rotation count
  Operands op{};
  op.a[0] = a.K;
  op.a[1] = a.Klo;
  op.b[0] = a.S;
  op.b[1] = a.Slo;
  op.lda = op.ldb = (size_t)a.n_mp;
  op.arow = a.gpat + w.x;
  op.arow_n = w.y;
  op.brow = a.crot + w.z;
  op.brow_n = w.w;
  op.zero = a.zero;
  floatx4 acc[TM][TN];
  dl_mma_loop<TM, TN, WGM, WGN, NA, NB, STAGES>(op, 0, 0, a.n_mp / BK, acc, lds);

  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, wm = wv / WGN, wn = wv % WGN;
  const int g = lane >> 4, fr = lane & 15;
#pragma unroll
  for (int y = 0; y < TN; y++) {
    const int b = wn * 16 * TN + y * 16 + fr;
    if (b >= w.w) continue;
    const int cb = w.z + b;
    const int r = a.crot[cb], q = a.ccell[cb], s = a.cslot[cb];
    const double cs = (double)a.sscale[r] * (double)a.kscale_inv;
    const double Ar = a.A[r];
#pragma unroll
    for (int x = 0; x < TM; x++)
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int m = wm * 16 * TM + x * 16 + 4 * g + j;
        if (m >= w.y) continue;
        const int i = a.gpat[w.x + m];
        const int rk = a.rank[(size_t)i * kMaxCore + q];
        if (rk < 0) continue;
        const double phi = (double)a.phi[i];
        const double cc = (double)acc[x][y][j] * cs;
        const float L = a.poisson ? (float)(cc + a.B[i] * log(phi) - phi * Ar)
                                  : (float)(-(Ar - 2.0 * cc / phi + a.B[i] / (phi * phi)) * a.inv_2nms2);
        a.Lc[((size_t)i * a.topk + rk) * kSlot + s] = L;
      }
  }
}

void launch_group_estep(bool ksplit, const GroupEstepArgs& a, hipStream_t st) {
  if (a.n_work <= 0) return;
  if (ksplit)
    hipLaunchKernelGGL(k_group_estep<true>, dim3(a.n_work), dim3(1024), 0, st, a);
  else
    hipLaunchKernelGGL(k_group_estep<false>, dim3(a.n_work), dim3(1024), 0, st, a);
  EMC_LAUNCHED();
}

}  // namespace emc
