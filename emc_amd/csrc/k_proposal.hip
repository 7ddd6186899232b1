// Kernels of the candidate-restricted iteration (learned rotation proposal).
//
// Every pattern i has topk candidate coarse cells cand[i][j]; each cell c
// holds the fine rotations cell_rot[cell_off[c] .. cell_off[c+1]) (at most
// kSlot).  Work is organised by cell: the entries (pattern, j) whose
// candidate j is c form the cell's list ent_off[c] .. ent_off[c+1], patterns
// ascending.  Per-candidate values live in Lc[i][j][s] (s = slot = position
// of the rotation in its cell), first log-likelihoods, then responsibilities.
//
// E-step: one workgroup = one cell x 128 entries; the gathered count rows
// and the cell's slice rows (hi/lo) stream through the FP16 main loop by
// direct global->LDS loads.  M-step: one workgroup = one cell x 256 pixels;
// the responsibility panel (hi/lo, prepared per cell) and the gathered count
// rows are staged in LDS, two stages.  Each rotation belongs to exactly one
// cell, so every M-step output is owned by one workgroup: no atomics.
#include <hip/hip_runtime.h>

#include "engine.hpp"
#include "gemm16.hpp"
#include "gemm_common.hpp"
#include "proposal_kernels.hpp"
#include "reduce.hpp"

namespace emc {

// ---------------------------------------------------------------------------
// E-step over the candidates
// ---------------------------------------------------------------------------
template <bool KSPLIT>
__global__ __launch_bounds__(256) void k_cand_estep(CandEstepArgs a) {
  constexpr int TM = 2, TN = 2, WGM = 4, WGN = 1, NA = KSPLIT ? 2 : 1, NB = 2, STAGES = 3;
  constexpr int BM = 16 * WGM * TM, BN = 16 * WGN * TN;
  static_assert(BM == kCandRows && BN == kSlot, "tile must match the work items");
  __shared__ __attribute__((aligned(1024))) _Float16 lds[STAGES * (NA * BM + NB * BN) * 32];
  // work items in cell order, round robin over the XCDs: the whole chip
  // works on one window of neighbouring cells, whose patterns' count rows
  // then stay in the memory-side cache
  const int2 w = a.work[blockIdx.x];
  const int c = w.x, e0 = w.y;
  const int e1 = min(e0 + BM, a.ent_off[c + 1]);
  const int r0 = a.cell_off[c], nr = a.cell_off[c + 1] - r0;
  Operands op{};
  op.a[0] = a.K;
  op.a[1] = a.Klo;
  op.b[0] = a.S;
  op.b[1] = a.Slo;
  op.lda = op.ldb = (size_t)a.n_mp;
  op.arow = a.ent_pat + e0;
  op.arow_n = e1 - e0;
  op.brow = a.cell_rot + r0;
  op.brow_n = nr;
  op.zero = a.zero;
  floatx4 acc[TM][TN];
  dl_mma_loop<TM, TN, WGM, WGN, NA, NB, STAGES>(op, 0, 0, a.n_mp / BK, acc, lds);

  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int g = lane >> 4, fr = lane & 15;
#pragma unroll
  for (int y = 0; y < TN; y++) {
    const int s = y * 16 + fr;
    if (s >= nr) continue;
    const int r = a.cell_rot[r0 + s];
    const double cs = (double)a.sscale[r] * (double)a.kscale_inv;
    const double Ar = a.A[r];
#pragma unroll
    for (int x = 0; x < TM; x++)
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int e = e0 + wv * 16 * TM + x * 16 + 4 * g + j;
        if (e >= e1) continue;
        const int i = a.ent_pat[e];
        const double phi = (double)a.phi[i];
        const double cc = (double)acc[x][y][j] * cs;
        const float L = a.poisson ? (float)(cc + a.B[i] * log(phi) - phi * Ar)
                                  : (float)(-(Ar - 2.0 * cc / phi + a.B[i] / (phi * phi)) * a.inv_2nms2);
        a.Lc[((size_t)i * a.topk + a.ent_j[e]) * kSlot + s] = L;
      }
  }
}

void launch_cand_estep(bool ksplit, const CandEstepArgs& a, hipStream_t st) {
  if (a.n_work <= 0) return;
  if (ksplit)
    hipLaunchKernelGGL(k_cand_estep<true>, dim3(a.n_work), dim3(256), 0, st, a);
  else
    hipLaunchKernelGGL(k_cand_estep<false>, dim3(a.n_work), dim3(256), 0, st, a);
  EMC_LAUNCHED();
}

// ---------------------------------------------------------------------------
// Softmax over each pattern's candidates (emc_cuda.cu:445-476 semantics:
// m = max(-1e10, max L), P = exp(L - m) if L - m > -1e10 else 0, P /= sum),
// one wave per pattern.  Also: sum P log P of the pattern, the scaling
// denominator sum_r P_ri C_ir from L (C_ir = (phi/2)(A_r + B_i/phi^2 +
// 2 n_m sigma^2 L_ri)), and the responsibilities scattered into the dense
// [R][N] matrix (zeroed beforehand).  Patterns scored on the whole table
// this iteration (insub) get zero candidate responsibilities.
// ---------------------------------------------------------------------------
// One wave per pattern; with kSlot == 32, a lane's k-th entry t = lane + 64k
// is slot lane % 32 of candidate cell 2k + lane / 32.  The lane's entries are
// loaded once into registers (NPL of them, NPL >= ceil(topk / 2)), each load
// kind issued for all entries before the next kind so that no entry waits on
// a dependent load pair of its own; sums run over the entries in t order as
// a loop over t would.
static_assert(kSlot == 32, "k_cand_softmax maps two candidate cells to a wave's 64 lanes");
template <int NPL>
__global__ __launch_bounds__(256) void k_cand_softmax(CandSoftmaxArgs a) {
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= a.N) return;
  float* row = a.Lc + (size_t)i * a.topk * kSlot;
  const int n = a.topk * kSlot;
  const int* cand = a.cand + (size_t)i * a.topk;
  const size_t si = (size_t)a.slot * a.N + i;
  if (a.insub[i]) {
    if (a.mode == 1) {
      if (lane == 0) {
        a.smax[si] = -1.0e10f;
        a.ssum[si] = 0.0;
      }
      return;
    }
    for (int t = lane; t < n; t += 64) row[t] = 0.0f;
    if (lane == 0) {
      if (a.mode == 2)
        for (int q = 0; q < 4; q++) a.sacc[(size_t)q * a.nmerge * a.N + si] = 0.0;
      else
        a.plogp[i] = 0.0;
    }
    return;
  }
  const int s = lane & 31;
  int o[NPL], lim[NPL];
  float v[NPL];
#pragma unroll
  for (int k = 0; k < NPL; k++) o[k] = lane + 64 * k < n ? cand[2 * k + (lane >> 5)] : -1;
#pragma unroll
  for (int k = 0; k < NPL; k++) {
    lim[k] = 0;
    if (o[k] >= 0) {
      const int c = o[k];
      o[k] = a.cell_off[c];
      lim[k] = a.cell_off[c + 1];
    } else {
      o[k] = 0;
    }
  }
#pragma unroll
  for (int k = 0; k < NPL; k++) v[k] = lane + 64 * k < n ? row[lane + 64 * k] : 0.0f;
  unsigned long long on = 0;
#pragma unroll
  for (int k = 0; k < NPL; k++)
    if (lane + 64 * k < n && s < lim[k] - o[k]) on |= 1ull << k;
  float m;
  double inv;
  if (a.mode == 2) {
    merge_stats(a.smax, a.ssum, a.nmerge, a.N, i, m, inv);
  } else {
    m = -1.0e10f;
#pragma unroll
    for (int k = 0; k < NPL; k++)
      if (on >> k & 1) m = fmaxf(m, v[k]);
    m = wave_max(m);
    double sum = 0.0;
#pragma unroll
    for (int k = 0; k < NPL; k++)
      if (on >> k & 1) {
        const float d = v[k] - m;
        if (d > -1.0e10f) sum += (double)expf(d);
      }
    sum = wave_sum(sum);
    if (a.mode == 1) {
      if (lane == 0) {
        a.smax[si] = m;
        a.ssum[si] = sum;
      }
      return;
    }
    inv = 1.0 / sum;
  }
  // Rotation of each entry (entries off the cell read rotation 0 and are
  // not used), then its A.
  int r[NPL];
#pragma unroll
  for (int k = 0; k < NPL; k++) r[k] = a.cell_rot[(on >> k & 1) ? o[k] + s : 0];
  double Ar[NPL];
#pragma unroll
  for (int k = 0; k < NPL; k++) Ar[k] = a.A[r[k]];
  double plogp = 0.0, ps = 0.0, pa = 0.0, pl = 0.0;
#pragma unroll
  for (int k = 0; k < NPL; k++) {
    const int t = lane + 64 * k;
    if (t >= n) continue;
    float p = 0.0f;
    if (on >> k & 1) {
      const float d = v[k] - m;
      const float ex = d > -1.0e10f ? expf(d) : 0.0f;
      p = (float)((double)ex * inv);
      if (p > 0.0f) plogp += (double)(p * logf(p));
      ps += p;
      pa += (double)p * Ar[k];
      pl += (double)p * (double)v[k];
      if (a.respons) a.respons[(size_t)r[k] * a.N + i] = p;
    }
    row[t] = p;
  }
  plogp = wave_sum(plogp);
  ps = wave_sum(ps);
  pa = wave_sum(pa);
  pl = wave_sum(pl);
  if (lane == 0) {
    if (a.mode == 2) {
      const size_t cn = (size_t)a.nmerge * a.N;
      a.sacc[si] = plogp;
      a.sacc[cn + si] = ps;
      a.sacc[2 * cn + si] = pa;
      a.sacc[3 * cn + si] = pl;
    } else {
      const double f = (double)a.phi[i];
      a.phiden[i] = a.poisson ? pa : 0.5 * f * (pa + a.B[i] / (f * f) * ps + a.sigma2x * pl);
      a.plogp[i] = plogp;
    }
  }
}

// The same for topk above 32 (where the register form would run at one wave
// per SIMD): each entry's cell bounds are read per pass.
__global__ __launch_bounds__(256) void k_cand_softmax_wide(CandSoftmaxArgs a) {
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= a.N) return;
  float* row = a.Lc + (size_t)i * a.topk * kSlot;
  const int n = a.topk * kSlot;
  const int* cand = a.cand + (size_t)i * a.topk;
  const size_t si = (size_t)a.slot * a.N + i;
  if (a.insub[i]) {
    if (a.mode == 1) {
      if (lane == 0) {
        a.smax[si] = -1.0e10f;
        a.ssum[si] = 0.0;
      }
      return;
    }
    for (int t = lane; t < n; t += 64) row[t] = 0.0f;
    if (lane == 0) {
      if (a.mode == 2)
        for (int q = 0; q < 4; q++) a.sacc[(size_t)q * a.nmerge * a.N + si] = 0.0;
      else
        a.plogp[i] = 0.0;
    }
    return;
  }
  float m;
  double inv;
  if (a.mode == 2) {
    merge_stats(a.smax, a.ssum, a.nmerge, a.N, i, m, inv);
  } else {
    m = -1.0e10f;
    for (int t = lane; t < n; t += 64) {
      const int j = t / kSlot, s = t - j * kSlot;
      const int c = cand[j];
      if (s < a.cell_off[c + 1] - a.cell_off[c]) m = fmaxf(m, row[t]);
    }
    m = wave_max(m);
    double sum = 0.0;
    for (int t = lane; t < n; t += 64) {
      const int j = t / kSlot, s = t - j * kSlot;
      const int c = cand[j];
      if (s < a.cell_off[c + 1] - a.cell_off[c]) {
        const float d = row[t] - m;
        if (d > -1.0e10f) sum += (double)expf(d);
      }
    }
    sum = wave_sum(sum);
    if (a.mode == 1) {
      if (lane == 0) {
        a.smax[si] = m;
        a.ssum[si] = sum;
      }
      return;
    }
    inv = 1.0 / sum;
  }
  double plogp = 0.0, ps = 0.0, pa = 0.0, pl = 0.0;
  for (int t = lane; t < n; t += 64) {
    const int j = t / kSlot, s = t - j * kSlot;
    const int c = cand[j];
    const int o = a.cell_off[c];
    float p = 0.0f;
    if (s < a.cell_off[c + 1] - o) {
      const float v = row[t];
      const float d = v - m;
      const float ex = d > -1.0e10f ? expf(d) : 0.0f;
      p = (float)((double)ex * inv);
      const int r = a.cell_rot[o + s];
      if (p > 0.0f) plogp += (double)(p * logf(p));
      ps += p;
      pa += (double)p * a.A[r];
      pl += (double)p * (double)v;
      if (a.respons) a.respons[(size_t)r * a.N + i] = p;
    }
    row[t] = p;
  }
  plogp = wave_sum(plogp);
  ps = wave_sum(ps);
  pa = wave_sum(pa);
  pl = wave_sum(pl);
  if (lane == 0) {
    if (a.mode == 2) {
      const size_t cn = (size_t)a.nmerge * a.N;
      a.sacc[si] = plogp;
      a.sacc[cn + si] = ps;
      a.sacc[2 * cn + si] = pa;
      a.sacc[3 * cn + si] = pl;
    } else {
      const double f = (double)a.phi[i];
      a.phiden[i] = a.poisson ? pa : 0.5 * f * (pa + a.B[i] / (f * f) * ps + a.sigma2x * pl);
      a.plogp[i] = plogp;
    }
  }
}

void launch_cand_softmax(const CandSoftmaxArgs& a, hipStream_t st) {
  const int npl = (a.topk + 1) / 2;
  if (npl <= 8)
    hipLaunchKernelGGL(k_cand_softmax<8>, dim3((a.N + 3) / 4), dim3(256), 0, st, a);
  else if (npl <= 16)
    hipLaunchKernelGGL(k_cand_softmax<16>, dim3((a.N + 3) / 4), dim3(256), 0, st, a);
  else
    hipLaunchKernelGGL(k_cand_softmax_wide, dim3((a.N + 3) / 4), dim3(256), 0, st, a);
  EMC_LAUNCHED();
}

// ---------------------------------------------------------------------------
// Whole-table subset (audited patterns and patterns in full search)
// ---------------------------------------------------------------------------
// Rows sub[k] of the packed counts (hi, and lo if given) -> [n_sub_pad][n_mp],
// zero rows from n_sub on.
__global__ __launch_bounds__(256) void k_gather_rows(const _Float16* __restrict__ K, const _Float16* __restrict__ Klo,
                                                     const int* __restrict__ sub, int n_sub, int n_mp,
                                                     _Float16* __restrict__ Ks, _Float16* __restrict__ Kslo) {
  const int k = blockIdx.x;
  typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
  const long src = k < n_sub ? (long)sub[k] * n_mp : -1;
  for (int q = threadIdx.x; q < n_mp / 8; q += 256) {
    u32x4 v = {0, 0, 0, 0}, vl = {0, 0, 0, 0};
    if (src >= 0) {
      v = *reinterpret_cast<const u32x4*>(K + src + q * 8);
      if (Klo) vl = *reinterpret_cast<const u32x4*>(Klo + src + q * 8);
    }
    *reinterpret_cast<u32x4*>(Ks + (long)k * n_mp + q * 8) = v;
    if (Klo) *reinterpret_cast<u32x4*>(Kslo + (long)k * n_mp + q * 8) = vl;
  }
}

// Transposed gather: Kt[p][k] = K[sub[k]][p], [n_mp][n_sub_pad], zero for
// k >= n_sub.  64 x 64 tiles through LDS.
__global__ __launch_bounds__(256) void k_gather_rows_t(const _Float16* __restrict__ K, const int* __restrict__ sub,
                                                       int n_sub, int n_sub_pad, int n_mp, _Float16* __restrict__ Kt) {
  __shared__ _Float16 tile[64][66];
  const int k0 = blockIdx.x * 64, p0 = blockIdx.y * 64;
  for (int idx = threadIdx.x; idx < 64 * 64; idx += 256) {
    const int kk = idx >> 6, pp = idx & 63;
    const int k = k0 + kk;
    tile[kk][pp] = k < n_sub ? K[(long)sub[k] * n_mp + p0 + pp] : (_Float16)0.0f;
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < 64 * 64; idx += 256) {
    const int pp = idx >> 6, kk = idx & 63;
    Kt[(long)(p0 + pp) * n_sub_pad + k0 + kk] = tile[kk][pp];
  }
}

__global__ void k_gather_f64(const double* __restrict__ v, const int* __restrict__ sub, int n, double* __restrict__ out) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k < n) out[k] = v[sub[k]];
}
__global__ void k_gather_f32(const float* __restrict__ v, const int* __restrict__ sub, int n, float* __restrict__ out) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k < n) out[k] = v[sub[k]];
}

void launch_gather_rows(const _Float16* K, const _Float16* Klo, const int* sub, int n_sub, int n_sub_pad, int n_mp,
                        _Float16* Ks, _Float16* Kslo, hipStream_t st) {
  hipLaunchKernelGGL(k_gather_rows, dim3(n_sub_pad), dim3(256), 0, st, K, Klo, sub, n_sub, n_mp, Ks, Kslo);
  EMC_LAUNCHED();
}
void launch_gather_rows_t(const _Float16* K, const int* sub, int n_sub, int n_sub_pad, int n_mp, _Float16* Kt,
                          hipStream_t st) {
  hipLaunchKernelGGL(k_gather_rows_t, dim3(n_sub_pad / 64, n_mp / 64), dim3(256), 0, st, K, sub, n_sub, n_sub_pad,
                     n_mp, Kt);
  EMC_LAUNCHED();
}
void launch_gather_f64(const double* v, const int* sub, int n, double* out, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_gather_f64, dim3((n + 255) / 256), dim3(256), 0, st, v, sub, n, out);
  EMC_LAUNCHED();
}
void launch_gather_f32(const float* v, const int* sub, int n, float* out, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_gather_f32, dim3((n + 255) / 256), dim3(256), 0, st, v, sub, n, out);
  EMC_LAUNCHED();
}

// Per subset pattern k (one wave each): its scaling denominator to the
// pattern's slot, and the whole-table posterior mass its candidates hold.
__global__ __launch_bounds__(256) void k_sub_finish(const float* __restrict__ Psub, int n_sub,
                                                    const int* __restrict__ sub, const double* __restrict__ phiden_sub,
                                                    double* __restrict__ phiden, const int* __restrict__ cand, int topk,
                                                    const int* __restrict__ cell_off, const int* __restrict__ cell_rot,
                                                    const int* __restrict__ sh_rot, int R_sh,
                                                    float* __restrict__ mass) {
  const int lane = threadIdx.x & 63;
  const int k = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (k >= n_sub) return;
  const int i = sub[k];
  double s = 0.0;
  for (int t = lane; t < topk * kSlot; t += 64) {
    const int j = t / kSlot, q = t - j * kSlot;
    const int c = cand[(size_t)i * topk + j];
    const int o = cell_off[c];
    if (q < cell_off[c + 1] - o) s += Psub[(size_t)cell_rot[o + q] * n_sub + k];
  }
  // the shared rotations (disjoint from the candidate cells)
  for (int t = lane; t < R_sh; t += 64) s += Psub[(size_t)sh_rot[t] * n_sub + k];
  s = wave_sum(s);
  if (lane == 0) {
    phiden[i] = phiden_sub[k];
    mass[k] = (float)s;
  }
}

void launch_sub_finish(const float* Psub, int n_sub, const int* sub, const double* phiden_sub, double* phiden,
                       const int* cand, int topk, const int* cell_off, const int* cell_rot, const int* sh_rot,
                       int R_sh, float* mass, hipStream_t st) {
  if (n_sub <= 0) return;
  hipLaunchKernelGGL(k_sub_finish, dim3((n_sub + 3) / 4), dim3(256), 0, st, Psub, n_sub, sub, phiden_sub, phiden,
                     cand, topk, cell_off, cell_rot, sh_rot, R_sh, mass);
  EMC_LAUNCHED();
}

// Shared rotations: respons[sh_rot[r]][i] = Psh[r][i] for the patterns not
// scored on the whole table.  Grid (patterns / 1,024, shared rotations);
// a thread takes four consecutive patterns, with 16-byte loads and stores
// when the rows allow them (vec: N % 4 == 0 and both buffers 16-byte
// aligned).
__global__ __launch_bounds__(256) void k_shared_scatter(const float* __restrict__ Psh, int N,
                                                        const int* __restrict__ sh_rot,
                                                        const unsigned char* __restrict__ insub,
                                                        float* __restrict__ respons, int vec) {
  typedef float f4 __attribute__((ext_vector_type(4)));
  const int r = blockIdx.y;
  float* out = respons + (size_t)sh_rot[r] * N;
  const float* in = Psh + (size_t)r * N;
  const int i0 = (blockIdx.x * 256 + threadIdx.x) * 4;
  if (i0 >= N) return;
  if (vec) {
    const f4 v = *reinterpret_cast<const f4*>(in + i0);
    const uchar4 sk = *reinterpret_cast<const uchar4*>(insub + i0);
    if (!(sk.x | sk.y | sk.z | sk.w)) {
      *reinterpret_cast<f4*>(out + i0) = v;
    } else {
      if (!sk.x) out[i0] = v.x;
      if (!sk.y) out[i0 + 1] = v.y;
      if (!sk.z) out[i0 + 2] = v.z;
      if (!sk.w) out[i0 + 3] = v.w;
    }
    return;
  }
  for (int i = i0; i < min(N, i0 + 4); i++)
    if (!insub[i]) out[i] = in[i];
}

void launch_shared_scatter(const float* Psh, int R_sh, int N, const int* sh_rot, const unsigned char* insub,
                           float* respons, hipStream_t st) {
  if (R_sh <= 0) return;
  const int vec = N % 4 == 0 && (uintptr_t)Psh % 16 == 0 && (uintptr_t)respons % 16 == 0 &&
                  (uintptr_t)insub % 4 == 0;
  hipLaunchKernelGGL(k_shared_scatter, dim3((N + 1023) / 1024, R_sh), dim3(256), 0, st, Psh, N, sh_rot, insub,
                     respons, vec);
  EMC_LAUNCHED();
}

// The dense responsibilities before the candidates are scattered: zero
// everywhere except the columns of the patterns scored on the whole table,
// which take their subset posterior (colmap[i] = subset index or -1).  One
// pass over [R][N], 16-byte stores when N % 4 == 0.
__global__ __launch_bounds__(256) void k_fill_subset(float* __restrict__ respons, int N,
                                                     const int* __restrict__ colmap, const float* __restrict__ Psub,
                                                     int n_sub) {
  typedef float f4 __attribute__((ext_vector_type(4)));
  const int r = blockIdx.x;
  const float* ps = Psub + (size_t)r * n_sub;
  f4* out = reinterpret_cast<f4*>(respons + (size_t)r * N);
  const int4* cm = reinterpret_cast<const int4*>(colmap);
  for (int q = threadIdx.x; q < N / 4; q += 256) {
    const int4 c = cm[q];
    f4 v = {0.0f, 0.0f, 0.0f, 0.0f};
    if (c.x >= 0) v.x = ps[c.x];
    if (c.y >= 0) v.y = ps[c.y];
    if (c.z >= 0) v.z = ps[c.z];
    if (c.w >= 0) v.w = ps[c.w];
    __builtin_nontemporal_store(v, out + q);  // streamed: not re-read before the candidates' scatter
  }
}

void launch_fill_subset(float* respons, int R, int N, const int* colmap, const float* Psub, int n_sub,
                        hipStream_t st) {
  hipLaunchKernelGGL(k_fill_subset, dim3(R), dim3(256), 0, st, respons, N, colmap, Psub, n_sub);
  EMC_LAUNCHED();
}

// respons[r][sub[k]] = Psub[r][k] for every rotation r (one block per r).
__global__ __launch_bounds__(256) void k_sub_scatter(const float* __restrict__ Psub, int n_sub,
                                                     const int* __restrict__ sub, int N, int nonzero_only,
                                                     float* __restrict__ respons) {
  const int r = blockIdx.x;
  for (int k = threadIdx.x; k < n_sub; k += 256) {
    const float v = Psub[(size_t)r * n_sub + k];
    if (!nonzero_only || v != 0.0f) respons[(size_t)r * N + sub[k]] = v;
  }
}

void launch_sub_scatter(const float* Psub, int R, int n_sub, const int* sub, int N, float* respons, bool nonzero_only,
                        hipStream_t st) {
  if (n_sub <= 0) return;
  hipLaunchKernelGGL(k_sub_scatter, dim3(R), dim3(256), 0, st, Psub, n_sub, sub, N, (int)nonzero_only, respons);
  EMC_LAUNCHED();
}

// ---------------------------------------------------------------------------
// M-step over the candidates
// ---------------------------------------------------------------------------
// M-step preparation, in three load-balanced passes (cells differ widely in
// how many patterns chose them):
//  1. per 32-entry chunk: partial sums over its entries of P, P phi and the
//     largest |P * opscale| for each of the 32 slots;
//  2. per cell: those partials added in chunk order (fixed, so T_r is
//     reproducible) + the subset's sums -> T_r, Tphi_r, and the power-of-two
//     row scale placing the row's largest operand near 2^14;
//  3. per chunk: the FP16 operand panel [32 slots][32 entries] in hi and lo.
__device__ __forceinline__ int chunk_first_entry(const CandPrepArgs& a, long q, int& cell) {
  cell = a.chunk_cell[q];
  return a.ent_off[cell] + (int)(q - a.chunk_off[cell]) * 32;
}

__global__ __launch_bounds__(256) void k_prep_partial(CandPrepArgs a) {
  __shared__ double tpart[8][kSlot], ppart[8][kSlot];
  __shared__ float mpart[8][kSlot];
  const long q = blockIdx.x;
  int cell;
  const int e0 = chunk_first_entry(a, q, cell);
  const int ee = min(e0 + 32, a.ent_off[cell + 1]);
  const int tid = threadIdx.x, s = tid & (kSlot - 1), grp = tid >> 5;
  double t = 0.0, tp = 0.0;
  float mx = 0.0f;
  for (int e = e0 + grp; e < ee; e += 8) {
    const int i = a.ent_pat[e];
    const float p = a.Pc[((size_t)i * a.topk + a.ent_j[e]) * kSlot + s];
    t += p;
    if (a.phi) tp += (double)p * (double)a.phi[i];
    mx = fmaxf(mx, fabsf(p * a.invphi[i]));
  }
  tpart[grp][s] = t;
  ppart[grp][s] = tp;
  mpart[grp][s] = mx;
  __syncthreads();
  if (tid < kSlot) {
    double st = 0.0, sp = 0.0;
    float m = 0.0f;
    for (int g = 0; g < 8; g++) {
      st += tpart[g][tid];
      sp += ppart[g][tid];
      m = fmaxf(m, mpart[g][tid]);
    }
    a.part_T[q * kSlot + tid] = st;
    a.part_Tphi[q * kSlot + tid] = sp;
    a.part_max[q * kSlot + tid] = m;
  }
}

__global__ __launch_bounds__(64) void k_prep_reduce(CandPrepArgs a) {
  const int c = blockIdx.x, s = threadIdx.x;
  if (s >= kSlot) return;
  const int r0 = a.cell_off[c], nr = a.cell_off[c + 1] - r0;
  double t = 0.0, tp = 0.0;
  float mx = 0.0f;
  for (long q = a.chunk_off[c]; q < a.chunk_off[c + 1]; q++) {
    t += a.part_T[q * kSlot + s];
    tp += a.part_Tphi[q * kSlot + s];
    mx = fmaxf(mx, a.part_max[q * kSlot + s]);
  }
  float sc = 1.0f;
  if (mx > 0.0f && mx < 3.0e38f) {
    int ex;
    frexpf(mx, &ex);
    sc = ldexpf(1.0f, max(-120, min(120, 14 - ex)));
  }
  if (s < nr) {
    const int r = a.cell_rot[r0 + s];
    const double T = (a.Tsub ? a.Tsub[r] : 0.0) + t;
    a.T[r] = T;
    a.Tf[r] = (float)T;
    a.rowscale[r] = sc;
    if (a.phi) a.Tphi[r] = (float)((a.Tphisub ? (double)a.Tphisub[r] : 0.0) + tp);
  }
}

__global__ __launch_bounds__(256) void k_prep_panels(CandPrepArgs a) {
  __shared__ __attribute__((aligned(16))) _Float16 panel[2][kSlot * 32];
  const long q = blockIdx.x;
  int cell;
  const int e0 = chunk_first_entry(a, q, cell);
  const int ee = a.ent_off[cell + 1];
  const int r0 = a.cell_off[cell], nr = a.cell_off[cell + 1] - r0;
  const int tid = threadIdx.x;
  // thread (entry el = tid & 31, slots 4 * (tid >> 5) .. +3)
  const int el = tid & 31, s4 = (tid >> 5) * 4;
  const int e = e0 + el;
  floatx4 v = {0, 0, 0, 0};
  if (e < ee) {
    const int i = a.ent_pat[e];
    v = *reinterpret_cast<const floatx4*>(a.Pc + ((size_t)i * a.topk + a.ent_j[e]) * kSlot + s4);
    v = v * a.invphi[i];
  }
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const int s = s4 + u;
    const float sc = s < nr ? a.rowscale[a.cell_rot[r0 + s]] : 1.0f;
    const float x = v[u] * sc;
    const _Float16 h = (_Float16)x;
    panel[0][s * 32 + el] = h;
    panel[1][s * 32 + el] = (_Float16)(x - (float)h);
  }
  __syncthreads();
  typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
  const int h = tid >> 7, o = (tid & 127) * 8;
  *reinterpret_cast<u32x4*>(a.panels + (q * 2 + h) * (kSlot * 32) + o) =
      *reinterpret_cast<const u32x4*>(&panel[h][o]);
}

void launch_cand_prep(const CandPrepArgs& a, long n_chunks, hipStream_t st) {
  if (n_chunks > 0) {
    hipLaunchKernelGGL(k_prep_partial, dim3(n_chunks), dim3(256), 0, st, a);
    EMC_LAUNCHED();
  }
  hipLaunchKernelGGL(k_prep_reduce, dim3(a.C), dim3(64), 0, st, a);
  EMC_LAUNCHED();
  if (n_chunks > 0) {
    hipLaunchKernelGGL(k_prep_panels, dim3(n_chunks), dim3(256), 0, st, a);
    EMC_LAUNCHED();
  }
}

// XOR swizzle of the 16-byte chunks of a count row in LDS: the four groups
// of 8 entries read by one MFMA B fragment land in distinct banks.
__device__ __forceinline__ int kswz(int row) { return ((row >> 3) & 3) * 2; }

// S' = U / divisor under the likelihood's rule.
__device__ __forceinline__ float finish_slice(float u, float T, int poisson) {
  return poisson ? (T > 1e-30f ? u / T : 0.0f) : (T > 1e-10f ? u / T : u);
}

// U[r][k] = sum over a segment of the cell's entries e of
// (P/phi)_e,r K[pat(e)][k].  A cell whose entries fit one segment adds the
// subset's U already in Sc and writes S' = U / T to Sc and the full slices;
// the segments of a larger cell write partial sums, added in segment order by
// k_cand_finalize.  Workgroup: one segment x 256 pixels, 4 waves of 64
// pixels; chunk of 32 entries = one MFMA k-step.
template <bool KSPLIT>
__global__ __launch_bounds__(256) void k_cand_mstep(CandMstepArgs a) {
  constexpr int EC = 32, PT = 256, NBK = KSPLIT ? 2 : 1;
  __shared__ __attribute__((aligned(1024))) _Float16 kb[2][NBK][EC * PT];
  __shared__ __attribute__((aligned(1024))) _Float16 ab[2][2][kSlot * EC];
  int seg, pt;
  {
    int tm, tn;
    tile_of_block(a.n_seg * a.n_ptiles, 1, tm, tn);
    seg = tm % a.n_seg;
    pt = tm / a.n_seg;
  }
  const int cell = a.seg_cell[seg];
  const int p0 = pt * PT;
  const long chunk0 = a.seg_chunk0[seg];
  const int nch = a.seg_nch[seg];
  const int part = a.seg_part[seg];
  const int eb = a.ent_off[cell] + (int)(chunk0 - a.chunk_off[cell]) * EC;
  const int ee = min(eb + nch * EC, a.ent_off[cell + 1]);
  const int r0 = a.cell_off[cell], nr = a.cell_off[cell + 1] - r0;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int g = lane >> 4, fr = lane & 15;

  floatx4 acc[2][4];
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 4; y++) acc[x][y] = floatx4{0, 0, 0, 0};

  // pattern of each of this lane's 4 rows for a chunk
  auto rows_of = [&](int q, int (&pat)[4]) {
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int row = (wv * 4 + j) * 2 + (lane >> 5);
      const int e = eb + q * EC + row;
      pat[j] = e < ee ? a.ent_pat[e] : -1;
    }
  };
  // loads of one chunk: 16 wave loads per count split (2 rows of 512 B
  // each), 4 for the two 2 KB panels; a wave issues 4 * NBK + 1
  auto dma_chunk = [&](int q, int st, const int (&pat)[4]) {
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int ld = wv * 4 + j;
      const int row = ld * 2 + (lane >> 5);
      const int ch = (lane & 31) ^ kswz(row);
#pragma unroll
      for (int sp = 0; sp < NBK; sp++) {
        const _Float16* base = sp ? a.Klo : a.K;
        const _Float16* src = pat[j] >= 0 ? base + (size_t)pat[j] * a.n_mp + p0 + ch * 8 : a.zero + ch * 8;
        __builtin_amdgcn_global_load_lds((const void*)src, (EMC_LDS void*)(&kb[st][sp][ld * 512]), 16, 0, 0);
      }
    }
    // panels: hi (2 KB) and lo (2 KB) = 4 wave loads; wave wv takes load wv
    {
      const int h = wv >> 1, part = wv & 1;
      const _Float16* src = a.panels + ((chunk0 + q) * 2 + h) * (kSlot * EC) + part * 512 + lane * 8;
      __builtin_amdgcn_global_load_lds((const void*)src, (EMC_LDS void*)(&ab[st][h][part * 512]), 16, 0, 0);
    }
  };

  int pat_next[4];
  if (nch > 0) {
    int pat0[4];
    rows_of(0, pat0);
    dma_chunk(0, 0, pat0);
    if (nch > 1) rows_of(1, pat_next);
  }
  for (int q = 0; q < nch; q++) {
    const int st = q & 1;
    wait_vmcnt<0>();
    __builtin_amdgcn_s_barrier();
    if (q + 1 < nch) {
      dma_chunk(q + 1, st ^ 1, pat_next);
      if (q + 2 < nch) rows_of(q + 2, pat_next);
    }
    // B fragments: K[e = 8g + jj][pixel wv*64 + 16y + fr]
    halfx8 bf[NBK][4];
#pragma unroll
    for (int y = 0; y < 4; y++) {
      const int px = wv * 64 + y * 16 + fr;
      const int chk = px >> 3, within = px & 7;
#pragma unroll
      for (int sp = 0; sp < NBK; sp++)
#pragma unroll
        for (int jj = 0; jj < 8; jj++) {
          const int row = g * 8 + jj;
          bf[sp][y][jj] = kb[st][sp][row * PT + ((chk ^ kswz(row)) * 8) + within];
        }
    }
#pragma unroll
    for (int x = 0; x < 2; x++) {
      const halfx8 ah = *reinterpret_cast<const halfx8*>(&ab[st][0][(x * 16 + fr) * EC + g * 8]);
      const halfx8 al = *reinterpret_cast<const halfx8*>(&ab[st][1][(x * 16 + fr) * EC + g * 8]);
#pragma unroll
      for (int y = 0; y < 4; y++) {
        acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bf[0][y], acc[x][y], 0, 0, 0);
        acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bf[0][y], acc[x][y], 0, 0, 0);
        if (KSPLIT) acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bf[NBK - 1][y], acc[x][y], 0, 0, 0);
      }
    }
  }
  wait_vmcnt<0>();

  // epilogue: rows = slots 16x + 4g + j, column = pixel
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int s = x * 16 + 4 * g + j;
      if (s >= nr) continue;
      const int r = a.cell_rot[r0 + s];
      const float inv_scale = a.kscale_inv / a.rowscale[r];
      const float T = a.Tf[r];
#pragma unroll
      for (int y = 0; y < 4; y++) {
        const int k = p0 + wv * 64 + y * 16 + fr;
        if (k >= a.n_m) continue;
        float u = acc[x][y][j] * inv_scale;
        if (part >= 0) {
          a.part[((size_t)part * kSlot + s) * a.n_mp + k] = u;
          continue;
        }
        float* dst = a.Sc + (size_t)r * a.n_mp + k;
        if (a.has_sub) u += *dst;
        const float sv = finish_slice(u, T, a.poisson);
        *dst = sv;
        a.slices[(size_t)r * a.N2d + a.pix[k]] = sv;
      }
    }
}

// Cells split into several segments: U = (subset U) + the segments' partial
// sums in segment order, then S' as above.  One block per (cell, 256 pixels).
__global__ __launch_bounds__(256) void k_cand_finalize(CandMstepArgs a) {
  const int mc = blockIdx.x % a.n_mcell, pt = blockIdx.x / a.n_mcell;
  const int cell = a.mcell[mc];
  const int k = pt * 256 + threadIdx.x;
  if (k >= a.n_m) return;
  const int r0 = a.cell_off[cell], nr = a.cell_off[cell + 1] - r0;
  const int q0 = a.mcell_part[mc], nq = a.mcell_part[mc + 1] - q0;
  for (int s = 0; s < nr; s++) {
    const int r = a.cell_rot[r0 + s];
    float* dst = a.Sc + (size_t)r * a.n_mp + k;
    float u = a.has_sub ? *dst : 0.0f;
    for (int q = 0; q < nq; q++) u += a.part[((size_t)(q0 + q) * kSlot + s) * a.n_mp + k];
    const float sv = finish_slice(u, a.Tf[r], a.poisson);
    *dst = sv;
    a.slices[(size_t)r * a.N2d + a.pix[k]] = sv;
  }
}

void launch_cand_mstep(bool ksplit, const CandMstepArgs& a, hipStream_t st) {
  const dim3 grid(a.n_seg * a.n_ptiles);
  if (ksplit)
    hipLaunchKernelGGL(k_cand_mstep<true>, grid, dim3(256), 0, st, a);
  else
    hipLaunchKernelGGL(k_cand_mstep<false>, grid, dim3(256), 0, st, a);
  EMC_LAUNCHED();
  if (a.n_mcell > 0) {
    hipLaunchKernelGGL(k_cand_finalize, dim3(a.n_mcell * a.n_ptiles), dim3(256), 0, st, a);
    EMC_LAUNCHED();
  }
}

}  // namespace emc

namespace emc {

// ---------------------------------------------------------------------------
// Posterior cells for re-centring the candidates
// ---------------------------------------------------------------------------
// Wave per pattern scored on its candidates: mass of each candidate cell
// (sum over its slots), the kTop largest (ties to the lower rank).
__global__ __launch_bounds__(256) void k_cand_topcells(const float* __restrict__ Pc, int N, int topk,
                                                       const unsigned char* __restrict__ insub,
                                                       const int* __restrict__ cand, int ktop,
                                                       int* __restrict__ top, float* __restrict__ topm) {
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= N || insub[i]) return;
  float m = -1.0f;
  if (lane < topk) {
    const float* row = Pc + ((size_t)i * topk + lane) * kSlot;
    m = 0.0f;
    for (int s = 0; s < kSlot; s++) m += row[s];
  }
  for (int t = 0; t < ktop; t++) {
    float bv = m;
    int bj = lane < topk ? lane : 0x7fffffff;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const float ov = __shfl_xor(bv, o, 64);
      const int oj = __shfl_xor(bj, o, 64);
      if (ov > bv || (ov == bv && oj < bj)) {
        bv = ov;
        bj = oj;
      }
    }
    if (lane == 0) {
      top[(size_t)i * ktop + t] = bj < topk && bv >= 0.0f ? cand[(size_t)i * topk + bj] : -1;
      topm[(size_t)i * ktop + t] = bj < topk ? fmaxf(bv, 0.0f) : 0.0f;
    }
    if (lane == bj) m = -1.0f;
  }
}

// Patterns scored on the whole table: cell masses from the columns of Psub
// (lanes over consecutive subset columns, so every row read is coalesced),
// in two passes: the 4 largest of every group of kTopGroup cells, then the
// ktop largest of those.  Order: larger mass first, equal masses to the
// lower cell, so the result does not depend on the order of insertion.

__device__ __forceinline__ bool top_before(float m1, int c1, float m2, int c2) {
  return m1 > m2 || (m1 == m2 && c1 >= 0 && (c2 < 0 || c1 < c2));
}

__device__ __forceinline__ void top4_insert(float m, int c, float (&lm)[4], int (&lc)[4]) {
  if (!top_before(m, c, lm[3], lc[3])) return;
  int at = 3;
  while (at > 0 && top_before(m, c, lm[at - 1], lc[at - 1])) {
    lm[at] = lm[at - 1];
    lc[at] = lc[at - 1];
    at--;
  }
  lm[at] = m;
  lc[at] = c;
}

__global__ __launch_bounds__(64) void k_sub_topcells_part(const float* __restrict__ Psub, int n_sub, int C,
                                                          const int* __restrict__ cell_off,
                                                          const int* __restrict__ cell_rot,
                                                          int* __restrict__ pc, float* __restrict__ pm) {
  const int k = blockIdx.x * 64 + threadIdx.x;
  if (k >= n_sub) return;
  const int g = blockIdx.y, ng = gridDim.y;
  float lm[4] = {-1.0f, -1.0f, -1.0f, -1.0f};
  int lc[4] = {-1, -1, -1, -1};
  const int c1 = min(C, (g + 1) * kTopGroup);
  for (int c = g * kTopGroup; c < c1; c++) {
    // the cell's rotations 8 at a time: their indices, then their
    // responsibilities, loaded together (a load per rotation waited for in
    // turn made this a chain of dependent loads), added in rotation order
    const int j1 = cell_off[c + 1];
    float m = 0.0f;
    int j = cell_off[c];
    for (; j + 8 <= j1; j += 8) {
      int r[8];
      float v[8];
#pragma unroll
      for (int u = 0; u < 8; u++) r[u] = cell_rot[j + u];
#pragma unroll
      for (int u = 0; u < 8; u++) v[u] = Psub[(size_t)r[u] * n_sub + k];
#pragma unroll
      for (int u = 0; u < 8; u++) m += v[u];
    }
    for (; j < j1; j++) m += Psub[(size_t)cell_rot[j] * n_sub + k];
    top4_insert(m, c, lm, lc);
  }
  const size_t o = ((size_t)k * ng + g) * 4;
  for (int t = 0; t < 4; t++) {
    pc[o + t] = lc[t];
    pm[o + t] = lm[t];
  }
}

// One wave per pattern: lanes merge the partial lists of groups lane,
// lane + 64, ..., then the lanes' lists are merged pairwise down the wave.
__global__ __launch_bounds__(256) void k_sub_topcells_merge(const int* __restrict__ pc, const float* __restrict__ pm,
                                                            int n_sub, int ng, const int* __restrict__ sub, int ktop,
                                                            int* __restrict__ top, float* __restrict__ topm) {
  const int lane = threadIdx.x & 63;
  const int k = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (k >= n_sub) return;
  float lm[4] = {-1.0f, -1.0f, -1.0f, -1.0f};
  int lc[4] = {-1, -1, -1, -1};
  for (int g = lane; g < ng; g += 64)
    for (int t = 0; t < 4; t++) {
      const size_t o = ((size_t)k * ng + g) * 4 + t;
      if (pc[o] >= 0) top4_insert(pm[o], pc[o], lm, lc);
    }
  for (int off = 32; off > 0; off >>= 1) {
    float om[4];
    int oc[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
      om[t] = __shfl_down(lm[t], off);
      oc[t] = __shfl_down(lc[t], off);
    }
    if (lane < off)
#pragma unroll
      for (int t = 0; t < 4; t++)
        if (oc[t] >= 0) top4_insert(om[t], oc[t], lm, lc);
  }
  if (lane == 0) {
    const int i = sub[k];
    for (int t = 0; t < ktop; t++) {
      top[(size_t)i * ktop + t] = lc[t];
      topm[(size_t)i * ktop + t] = fmaxf(lm[t], 0.0f);
    }
  }
}

// Merge of each pattern's ktop candidate cells (top / topm, in and out) with
// its ktop cells among the shared ones (local numbers -> sh_cells), for the
// patterns not scored on the whole table.
__global__ __launch_bounds__(256) void k_merge_top(int N, int ktop, const unsigned char* __restrict__ insub,
                                                   int* __restrict__ top, float* __restrict__ topm,
                                                   const int* __restrict__ top_s, const float* __restrict__ topm_s,
                                                   const int* __restrict__ sh_cells) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= N || insub[i]) return;
  float lm[4] = {-1.0f, -1.0f, -1.0f, -1.0f};
  int lc[4] = {-1, -1, -1, -1};
  for (int t = 0; t < ktop; t++) {
    const size_t o = (size_t)i * ktop + t;
    if (top[o] >= 0) top4_insert(topm[o], top[o], lm, lc);
    if (top_s[o] >= 0) top4_insert(topm_s[o], sh_cells[top_s[o]], lm, lc);
  }
  for (int t = 0; t < ktop; t++) {
    top[(size_t)i * ktop + t] = lc[t];
    topm[(size_t)i * ktop + t] = fmaxf(lm[t], 0.0f);
  }
}

void launch_merge_top(int N, int ktop, const unsigned char* insub, int* top, float* topm, const int* top_s,
                      const float* topm_s, const int* sh_cells, hipStream_t st) {
  hipLaunchKernelGGL(k_merge_top, dim3((N + 255) / 256), dim3(256), 0, st, N, ktop, insub, top, topm, top_s, topm_s,
                     sh_cells);
  EMC_LAUNCHED();
}

void launch_topcells(const float* Pc, int N, int topk, const unsigned char* insub, const int* cand,
                     const float* Psub, int n_sub, const int* sub, int C, const int* cell_off, const int* cell_rot,
                     int ktop, int* top, float* topm, int* part_c, float* part_m, hipStream_t st) {
  if (Pc) {
    hipLaunchKernelGGL(k_cand_topcells, dim3((N + 3) / 4), dim3(256), 0, st, Pc, N, topk, insub, cand, ktop, top,
                       topm);
    EMC_LAUNCHED();
  }
  if (n_sub > 0) {
    const int ng = (C + kTopGroup - 1) / kTopGroup;
    hipLaunchKernelGGL(k_sub_topcells_part, dim3((n_sub + 63) / 64, ng), dim3(64), 0, st, Psub, n_sub, C, cell_off,
                       cell_rot, part_c, part_m);
    EMC_LAUNCHED();
    hipLaunchKernelGGL(k_sub_topcells_merge, dim3((n_sub + 3) / 4), dim3(256), 0, st, part_c, part_m, n_sub, ng,
                       sub, ktop, top, topm);
    EMC_LAUNCHED();
  }
}

}  // namespace emc
