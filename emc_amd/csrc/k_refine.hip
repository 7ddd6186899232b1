// Refinement of the significant responsibilities.
//
// The matrix-core E-step accumulates C_ir = sum_p m K_ip f(S_rp) in FP32
// over the masked pixels (f(S) = log max(S, floor) for the Poisson
// likelihood, S for the Gaussian one) and stores L in FP32.  At the
// benchmark depth (15,576 pixels, |L| ~ 4e4) the accumulation and storage
// rounding reach ~5e-3 in L, which moves a responsibility shared by two or
// three rotations by up to ~1e-3 and every responsibility by ~5e-3 of
// itself.  That relative error still matters for small entries: a voxel
// reached only by rotations of small T_r takes its value M / W from them
// alone.  So after the softmax every entry with P >= tau (1e-9 by default;
// a few per pattern under the Poisson likelihood, tens under the Gaussian
// one; a pattern with more than 1,024 such entries refines those >= 1e-6,
// or >= 1e-3, instead) is recomputed: C_ir over the pattern's nonzero
// counts in FP64 (S read from the FP32 slices, f in FP64), L_ir in FP64,
// then per pattern
//   P'_ri = (1 - f_i) exp(L_ri - M_i) / sum_refined exp(L - M_i),
// with M_i the largest refined L and f_i = 1 - sum_refined P_ri the mass of
// the entries left as they were (< tau each).  The refined values are
// written to the responsibility matrix and to the buffer the entry came from
// (candidate slots, shared rows or the whole-table subset of the proposal),
// and the per-pattern sums the scaling update reads (sum_r P_ri A_r for
// Poisson, sum_r P_ri C_ir for Gaussian) and sum P log P are corrected by
// the difference.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include <rocprim/device/device_scan.hpp>

#include "engine.hpp"
#include "reduce.hpp"

namespace emc {

// one entry: source address (nullptr: the matrix only), rotation, pattern,
// responsibility before, L in FP64, C in FP64
struct RefineEntry {
  float* src;
  int r, i;
  float p;
  int pad;
  double L, C;
};

// ---- count lists: every pattern's nonzero masked counts (pixel, count) ----
__global__ __launch_bounds__(256) void k_cnt_fill(const float* __restrict__ images, const int* __restrict__ pix,
                                                  int N, int N2d, int n_mp, const int* __restrict__ off,
                                                  int* __restrict__ px, float* __restrict__ cnt) {
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= N) return;
  int at = off[i];
  for (int k0 = 0; k0 < n_mp; k0 += 64) {
    const int p = pix[k0 + lane];
    const float v = p >= 0 ? images[(size_t)i * N2d + p] : 0.0f;
    const bool nz = v != 0.0f;
    const unsigned long long m = __ballot(nz);
    if (nz) {
      const int pos = at + __popcll(m & ((1ull << lane) - 1ull));
      px[pos] = p;
      cnt[pos] = v;
    }
    at += __popcll(m);
  }
}

// ---- collection ----
// Level thresholds of a pattern's entries: every entry >= T[0] (the
// option's tau) is refined, unless a pattern has more than kRefineCap of
// them; then its threshold rises to T[1] (1e-6), then to T[2] (1e-3: at most
// 1,000 entries of a distribution reach it).  lvl == nullptr: T[0] for
// every pattern (first pass, which also counts each pattern's entries per
// level into cnt[level * N + i]).
constexpr int kRefineCap = 1024;
struct RefineLevels {
  float T[3];
  const unsigned char* lvl;
  int* cnt;
  int N;
};

__device__ __forceinline__ void push(RefineEntry* out, unsigned* count, unsigned cap, bool take, float* src, int r,
                                     int i, float p) {
  const unsigned long long m = __ballot(take);
  if (!m) return;
  const int lane = threadIdx.x & 63;
  unsigned base = 0;
  if (lane == 0) base = atomicAdd(count, (unsigned)__popcll(m));
  base = __shfl(base, 0, 64);
  if (take) {
    const unsigned at = base + (unsigned)__popcll(m & ((1ull << lane) - 1ull));
    if (at < cap) out[at] = RefineEntry{src, r, i, p, 0, 0.0, 0.0};
  }
}

// A [rows][ld] block of responsibilities: row -> rotation (row_rot, or the
// row itself), column -> pattern (col_pat, or the column itself); columns
// whose pattern has skip set are left out.  Thread = column, workgroup row
// = a chunk of rows_per rows: the loads are coalesced along the columns and
// the level counts stay in registers until the chunk ends.
__global__ __launch_bounds__(256) void k_refine_rows(float* __restrict__ P, int rows, int ld, int cols,
                                                     const int* __restrict__ row_rot,
                                                     const int* __restrict__ col_pat,
                                                     const unsigned char* __restrict__ skip, int own, int rows_per,
                                                     RefineLevels lv, RefineEntry* __restrict__ out,
                                                     unsigned* __restrict__ count, unsigned cap) {
  const int col = blockIdx.x * 256 + threadIdx.x;
  const int r0 = blockIdx.y * rows_per, r1 = min(rows, r0 + rows_per);
  bool ok = col < cols;
  int i = 0;
  if (ok) {
    i = col_pat ? col_pat[col] : col;
    ok = !(skip && skip[i]);
  }
  const float thr = ok ? lv.T[lv.lvl ? lv.lvl[i] : 0] : 3.0e38f;
  int c0 = 0, c1 = 0, c2 = 0;
  constexpr int U = 8;  // rows loaded together, so that U loads are in flight
  for (int row0 = r0; row0 < r1; row0 += U) {  // uniform over the workgroup
    float pv[U];
#pragma unroll
    for (int u = 0; u < U; u++) pv[u] = ok && row0 + u < r1 ? P[(size_t)(row0 + u) * ld + col] : 0.0f;
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int row = row0 + u;
      if (row >= r1) break;  // uniform
      const float p = pv[u];
      c0 += p >= lv.T[0];
      c1 += p >= lv.T[1];
      c2 += p >= lv.T[2];
      push(out, count, cap, ok && p >= thr, own ? P + (size_t)row * ld + col : nullptr, row_rot ? row_rot[row] : row,
           i, p);
    }
  }
  if (ok && !lv.lvl && c0) {
    atomicAdd(&lv.cnt[i], c0);
    if (c1) atomicAdd(&lv.cnt[lv.N + i], c1);
    if (c2) atomicAdd(&lv.cnt[2 * lv.N + i], c2);
  }
}

// Candidate slots [N][topk][32] of the patterns outside the whole-table
// subset: slot s of candidate j is fine rotation cell_rot[cell_off[c] + s].
// One wave per pattern.
__global__ __launch_bounds__(256) void k_refine_cand(float* __restrict__ Lc, int N, int topk, int slots,
                                                     const int* __restrict__ cand, const int* __restrict__ cell_off,
                                                     const int* __restrict__ cell_rot,
                                                     const unsigned char* __restrict__ insub, RefineLevels lv,
                                                     RefineEntry* __restrict__ out, unsigned* __restrict__ count,
                                                     unsigned cap) {
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= N || (insub && insub[i])) return;  // uniform over the wave
  const float thr = lv.T[lv.lvl ? lv.lvl[i] : 0];
  float* row = Lc + (size_t)i * topk * slots;
  int c0 = 0, c1 = 0, c2 = 0;
  constexpr int U = 4;  // slot groups loaded together
  for (int t0 = 0; t0 < topk * slots; t0 += 64 * U) {
    int rv[U];
    float pv[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int t = t0 + u * 64 + lane;
      rv[u] = 0;
      pv[u] = 0.0f;
      if (t < topk * slots) {
        const int j = t / slots, s = t - j * slots;
        const int c = cand[(size_t)i * topk + j];
        const int o = cell_off[c];
        if (s < cell_off[c + 1] - o) {
          pv[u] = row[t];
          rv[u] = cell_rot[o + s];
        }
      }
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int t = t0 + u * 64 + lane;
      const float p = pv[u];
      c0 += p >= lv.T[0];
      c1 += p >= lv.T[1];
      c2 += p >= lv.T[2];
      push(out, count, cap, p >= thr, row + (t < topk * slots ? t : 0), rv[u], i, p);
    }
  }
  c0 = wave_sum(c0);
  c1 = wave_sum(c1);
  c2 = wave_sum(c2);
  if (lane == 0 && !lv.lvl && c0) {
    atomicAdd(&lv.cnt[i], c0);
    if (c1) atomicAdd(&lv.cnt[lv.N + i], c1);
    if (c2) atomicAdd(&lv.cnt[2 * lv.N + i], c2);
  }
}

// Each pattern's level and entry count from the first pass's counts;
// info[1] += entries kept, info[2] += patterns above level 0, info[3] +=
// patterns whose last level still holds more than kRefineCap entries.
__global__ __launch_bounds__(256) void k_refine_levels(const int* __restrict__ cnt, int N,
                                                       unsigned char* __restrict__ lvl, int* __restrict__ npat,
                                                       unsigned* __restrict__ info) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  unsigned kept = 0, over = 0, bad = 0;
  if (i < N) {
    const int c0 = cnt[i], c1 = cnt[N + i], c2 = cnt[2 * N + i];
    const int l = c0 <= kRefineCap ? 0 : (c1 <= kRefineCap ? 1 : 2);
    const int k = l == 0 ? c0 : (l == 1 ? c1 : c2);
    lvl[i] = (unsigned char)l;
    npat[i] = k;
    kept = (unsigned)k;
    over = l > 0;
    bad = k > kRefineCap;
  }
  kept = wave_sum(kept);
  over = wave_sum(over);
  bad = wave_sum(bad);
  if ((threadIdx.x & 63) == 0) {
    if (kept) atomicAdd(&info[1], kept);
    if (over) atomicAdd(&info[2], over);
    if (bad) atomicAdd(&info[3], bad);
  }
}

// Entries grouped by pattern: slot[pat_off[i] .. pat_off[i + 1]) holds the
// indices of pattern i's entries, in no particular order (the per-pattern
// pass sorts them).
__global__ __launch_bounds__(256) void k_refine_scatter(const RefineEntry* __restrict__ ent,
                                                        const unsigned* __restrict__ count,
                                                        const int* __restrict__ pat_off, int* __restrict__ fill,
                                                        int* __restrict__ slot) {
  const unsigned e = blockIdx.x * 256 + threadIdx.x;
  if (e >= *count) return;
  const int i = ent[e].i;
  slot[pat_off[i] + atomicAdd(&fill[i], 1)] = (int)e;
}

// One workgroup per pattern (at most kRefineCap entries).  (1) Each wave
// takes entries in turn: C_ir over the pattern's nonzero counts in FP64, then
// L_ir.  (2) The entries are ranked by (rotation, old value) so that every
// sum below runs in one fixed order whatever order the atomics filled the
// slots in.  (3) Wave 0: the largest L, the renormalised responsibilities
// written back, the correction of the scaling denominator and of
// sum P log P (dtot[i]), each a lane-strided sum in rank order followed by
// the fixed butterfly of wave_sum.
__global__ __launch_bounds__(256) void k_refine_group(RefineEntry* __restrict__ ent, const int* __restrict__ pat_off,
                                                      const int* __restrict__ slot, const float* __restrict__ slices,
                                                      int N2d, const int* __restrict__ off,
                                                      const int* __restrict__ px, const float* __restrict__ cnt,
                                                      const double* __restrict__ A, const double* __restrict__ B,
                                                      const float* __restrict__ phi, double inv_sigma2x, int poisson,
                                                      float* __restrict__ respons, int N,
                                                      double* __restrict__ phiden, double* __restrict__ dtot) {
  __shared__ unsigned long long key[kRefineCap];
  __shared__ int sorted[kRefineCap];
  const int i = blockIdx.x;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int s0 = pat_off[i];
  const int n = min(pat_off[i + 1] - s0, kRefineCap);  // the host checked the cap
  if (n == 0) {
    if (t == 0) dtot[i] = 0.0;
    return;
  }
  const double f = (double)phi[i];
  const int k0 = off[i], k1 = off[i + 1];
  for (int j = wave; j < n; j += 4) {
    RefineEntry& en = ent[slot[s0 + j]];
    const float* srow = slices + (size_t)en.r * N2d;
    double c = 0.0;
    for (int k = k0 + lane; k < k1; k += 64) {
      const double s = (double)srow[px[k]];
      c += (double)cnt[k] * (poisson ? log(s > (double)kSlogFloor ? s : (double)kSlogFloor) : s);
    }
    c = wave_sum(c);
    if (lane == 0) {
      const int r = en.r;
      en.C = c;
      en.L = poisson ? c + B[i] * log(f) - f * A[r] : -(A[r] - 2.0 * c / f + B[i] / (f * f)) * inv_sigma2x;
    }
  }
  for (int j = t; j < n; j += 256) {
    const RefineEntry& en = ent[slot[s0 + j]];
    key[j] = ((unsigned long long)(unsigned)en.r << 32) | (unsigned long long)__float_as_uint(en.p);
  }
  __syncthreads();
  for (int j = t; j < n; j += 256) {
    const unsigned long long kj = key[j];
    int rank = 0;
    for (int k = 0; k < n; k++) {
      const unsigned long long kk = key[k];
      rank += kk < kj || (kk == kj && k < j);
    }
    sorted[rank] = slot[s0 + j];
  }
  __syncthreads();
  if (wave != 0) return;
  double m = -INFINITY, pold = 0.0;
  for (int j = lane; j < n; j += 64) {
    const RefineEntry& en = ent[sorted[j]];
    m = fmax(m, en.L);
    pold += (double)en.p;
  }
  m = wave_max(m);
  pold = wave_sum(pold);
  double w = 0.0;
  for (int j = lane; j < n; j += 64) w += exp(ent[sorted[j]].L - m);
  w = wave_sum(w);
  const double keep = 1.0 - fmax(0.0, 1.0 - pold);  // mass of the refined entries
  double dphi = 0.0, d = 0.0;
  for (int j = lane; j < n; j += 64) {
    const RefineEntry en = ent[sorted[j]];
    const float p = (float)(keep * exp(en.L - m) / w);
    respons[(size_t)en.r * N + i] = p;
    if (en.src) *en.src = p;
    dphi += ((double)p - (double)en.p) * (poisson ? A[en.r] : en.C);
    // sum P log P as the softmax forms it (FP32 product and log)
    d += (p > 0.0f ? (double)(p * logf(p)) : 0.0) - (en.p > 0.0f ? (double)(en.p * logf(en.p)) : 0.0);
  }
  dphi = wave_sum(dphi);
  d = wave_sum(d);
  if (lane == 0) {
    if (phiden) phiden[i] += dphi;
    dtot[i] = d;
  }
}

// ---- host side ----
static bool ensure_counts(const float* images, int N, int N2d) {
  Engine& e = engine();
  if (e.cnt_images == images && e.cnt_N == N && e.cnt_pix_gen == e.pix_gen && e.cnt_stamp == e.invalidate_stamp)
    return true;
  std::vector<int> nnz(N), off(N + 1, 0);
  if (!e.d_cnt_off.ensure(sizeof(int) * (N + 1))) return false;
  launch_csr_row_nnz(images, e.d_pix.as<int>(), N, N2d, e.n_mp, e.d_cnt_off.as<int>(), e.stream);
  EMC_CHECK(hipMemcpyAsync(nnz.data(), e.d_cnt_off.p, sizeof(int) * N, hipMemcpyDeviceToHost, e.stream));
  EMC_CHECK(hipStreamSynchronize(e.stream));
  if (failed()) return false;
  for (int i = 0; i < N; i++) off[i + 1] = off[i] + nnz[i];
  const size_t nz = std::max<size_t>((size_t)off[N], 1);
  if (!e.d_cnt_px.ensure(sizeof(int) * nz) || !e.d_cnt_val.ensure(sizeof(float) * nz)) return false;
  EMC_CHECK(hipMemcpyAsync(e.d_cnt_off.p, off.data(), sizeof(int) * (N + 1), hipMemcpyHostToDevice, e.stream));
  hipLaunchKernelGGL(k_cnt_fill, dim3((N + 3) / 4), dim3(256), 0, e.stream, images, e.d_pix.as<int>(), N, N2d,
                     e.n_mp, e.d_cnt_off.as<int>(), e.d_cnt_px.as<int>(), e.d_cnt_val.as<float>());
  EMC_LAUNCHED();
  EMC_CHECK(hipStreamSynchronize(e.stream));
  if (failed()) return false;
  e.cnt_images = images;
  e.cnt_N = N;
  e.cnt_pix_gen = e.pix_gen;
  e.cnt_stamp = e.invalidate_stamp;
  return true;
}

static void launch_collect(const RefineInputs& in, const RefineLevels& lv, unsigned cap) {
  Engine& e = engine();
  RefineEntry* out = e.d_ref_ent.as<RefineEntry>();
  unsigned* count = e.d_ref_count.as<unsigned>();
  constexpr int kRowsPer = 64;
  for (int k = 0; k < in.nrows; k++) {
    const RefineRows& s = in.rows[k];
    if (!s.P || s.rows <= 0 || s.cols <= 0) continue;
    const dim3 grid((unsigned)((s.cols + 255) / 256), (unsigned)((s.rows + kRowsPer - 1) / kRowsPer));
    hipLaunchKernelGGL(k_refine_rows, grid, dim3(256), 0, e.stream, s.P, s.rows, s.ld, s.cols, s.row_rot, s.col_pat,
                       s.skip, (int)s.own, kRowsPer, lv, out, count, cap);
    EMC_LAUNCHED();
  }
  if (in.Lc) {
    hipLaunchKernelGGL(k_refine_cand, dim3((in.N + 3) / 4), dim3(256), 0, e.stream, in.Lc, in.N, in.topk, in.slots,
                       in.cand, in.cell_off, in.cell_rot, in.insub, lv, out, count, cap);
    EMC_LAUNCHED();
  }
}

bool refine_posterior(const RefineInputs& in, double* dtotal) {
  Engine& e = engine();
  hipStream_t st = e.stream;
  *dtotal = 0.0;
  e.ref_entries = 0;
  if (in.N <= 0) return true;
  if (!ensure_counts(in.images, in.N, in.N2d)) return false;
  static unsigned* h_info = nullptr;
  static double* h_tot = nullptr;
  if (!h_info) {
    if (hipHostMalloc((void**)&h_info, sizeof(unsigned) * 4, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&h_tot, sizeof(double) * 2, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      set_error("refinement: page-locked host memory");
      return false;
    }
  }
  const int N = in.N;
  unsigned cap = (unsigned)(e.d_ref_ent.bytes / sizeof(RefineEntry));
  const unsigned want = (unsigned)std::min<size_t>((size_t)32 * N + 65536, 1u << 26);
  if (cap < want) {
    if (!e.d_ref_ent.ensure(sizeof(RefineEntry) * (size_t)want)) return false;
    cap = want;
  }
  // per pattern: entry count (N + 1, the last stays 0), offsets (N + 1),
  // fill counters (N), level counts (3N); then the sum P log P changes (N
  // doubles) and the levels (N bytes)
  const size_t ints = 6 * (size_t)N + 2, dbl_at = round_up(sizeof(int) * ints, 16);
  if (!e.d_ref_count.ensure(16) ||
      !e.d_ref_pat.ensure(dbl_at + sizeof(double) * (size_t)N + round_up((size_t)N, 16)) ||
      !e.d_ref_tot.ensure(sizeof(double) * 2))
    return false;
  int* npat = e.d_ref_pat.as<int>();
  int* pat_off = npat + (N + 1);
  int* fill = pat_off + (N + 1);
  int* cnt = fill + N;
  double* dtot = reinterpret_cast<double*>(e.d_ref_pat.as<char>() + dbl_at);
  unsigned char* lvl = reinterpret_cast<unsigned char*>(dtot + N);
  unsigned* info = e.d_ref_count.as<unsigned>();
  RefineLevels lv;
  lv.T[0] = in.tau;
  lv.T[1] = std::max(in.tau, 1e-6f);
  lv.T[2] = std::max(in.tau, 1e-3f);
  lv.lvl = nullptr;
  lv.cnt = cnt;
  lv.N = N;
  launch_fill(e.d_ref_pat.as<float>(), ints, 0.0f, st);
  launch_fill(e.d_ref_count.as<float>(), 4, 0.0f, st);
  // first pass: every entry >= tau, counted per pattern and level
  launch_collect(in, lv, cap);
  hipLaunchKernelGGL(k_refine_levels, dim3((N + 255) / 256), dim3(256), 0, st, cnt, N, lvl, npat, info);
  EMC_LAUNCHED();
  EMC_CHECK(hipMemcpyAsync(h_info, info, sizeof(unsigned) * 4, hipMemcpyDeviceToHost, st));
  EMC_CHECK(hipStreamSynchronize(st));
  if (failed()) return false;
  if (h_info[3]) {
    set_error("refinement: %u patterns hold more than %d responsibilities >= %g", h_info[3], kRefineCap,
              (double)lv.T[2]);
    return false;
  }
  if (h_info[2] || h_info[0] > cap) {
    // second pass: each pattern at its level, into a list of the exact size
    const unsigned need = h_info[1];
    if (need > cap) {
      if (!e.d_ref_ent.ensure(sizeof(RefineEntry) * (size_t)need)) return false;
      cap = need;
    }
    lv.lvl = lvl;
    launch_fill(e.d_ref_count.as<float>(), 1, 0.0f, st);
    launch_collect(in, lv, cap);
    EMC_CHECK(hipMemcpyAsync(h_info, info, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    EMC_CHECK(hipStreamSynchronize(st));
    if (failed()) return false;
    if (h_info[0] != need) {
      set_error("refinement: second pass collected %u entries, the counts said %u", h_info[0], need);
      return false;
    }
  }
  const unsigned n = h_info[0];
  e.ref_entries = n;
  if (n == 0) return true;
  if (!e.d_ref_slot.ensure(sizeof(int) * (size_t)n)) return false;
  const unsigned* count = info;
  RefineEntry* ent = e.d_ref_ent.as<RefineEntry>();
  int* slot = e.d_ref_slot.as<int>();
  size_t tbytes = 0;
  EMC_CHECK(rocprim::exclusive_scan(nullptr, tbytes, npat, pat_off, 0, (size_t)N + 1, rocprim::plus<int>(), st));
  if (!e.d_ref_tmp.ensure(std::max<size_t>(tbytes, 16))) return false;
  EMC_CHECK(rocprim::exclusive_scan(e.d_ref_tmp.p, tbytes, npat, pat_off, 0, (size_t)N + 1, rocprim::plus<int>(),
                                    st));
  hipLaunchKernelGGL(k_refine_scatter, dim3((n + 255) / 256), dim3(256), 0, st, ent, count, pat_off, fill, slot);
  EMC_LAUNCHED();
  hipLaunchKernelGGL(k_refine_group, dim3(N), dim3(256), 0, st, ent, pat_off, slot, in.slices, in.N2d,
                     e.d_cnt_off.as<int>(), e.d_cnt_px.as<int>(), e.d_cnt_val.as<float>(), e.d_A.as<double>(),
                     e.d_B.as<double>(), e.d_phi_used.as<float>(), in.sigma2x > 0.0 ? 1.0 / in.sigma2x : 0.0,
                     (int)in.poisson, in.respons, N, in.phiden, dtot);
  EMC_LAUNCHED();
  launch_sum_f64(dtot, N, e.d_ref_tot.as<double>(), st);
  EMC_CHECK(hipMemcpyAsync(h_tot, e.d_ref_tot.p, sizeof(double), hipMemcpyDeviceToHost, st));
  EMC_CHECK(hipStreamSynchronize(st));
  if (failed()) return false;
  *dtotal = *h_tot;
  return true;
}

}  // namespace emc
