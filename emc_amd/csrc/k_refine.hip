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
// a few per pattern under the Poisson likelihood) is recomputed: C_ir over
// the pattern's nonzero
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
// whose pattern has skip set are left out.
__global__ __launch_bounds__(256) void k_refine_collect_rows(float* __restrict__ P, int rows, int ld, int cols,
                                                             const int* __restrict__ row_rot,
                                                             const int* __restrict__ col_pat,
                                                             const unsigned char* __restrict__ skip, int own,
                                                             float tau, RefineEntry* __restrict__ out,
                                                             unsigned* __restrict__ count, unsigned cap) {
  const size_t n = (size_t)rows * cols;
  for (size_t base = (size_t)blockIdx.x * 256; base < n; base += (size_t)gridDim.x * 256) {
    const size_t t = base + threadIdx.x;
    bool take = false;
    int r = 0, i = 0;
    float p = 0.0f;
    float* src = nullptr;
    if (t < n) {
      const int row = (int)(t / cols), col = (int)(t % cols);
      float* q = P + (size_t)row * ld + col;
      p = *q;
      i = col_pat ? col_pat[col] : col;
      r = row_rot ? row_rot[row] : row;
      take = p >= tau && !(skip && skip[i]);
      src = own ? q : nullptr;
    }
    push(out, count, cap, take, src, r, i, p);
  }
}

// Candidate slots [N][topk][32] of the patterns outside the whole-table
// subset: slot s of candidate j is fine rotation cell_rot[cell_off[c] + s].
__global__ __launch_bounds__(256) void k_refine_collect_cand(float* __restrict__ Lc, int N, int topk, int slots,
                                                             const int* __restrict__ cand,
                                                             const int* __restrict__ cell_off,
                                                             const int* __restrict__ cell_rot,
                                                             const unsigned char* __restrict__ insub, float tau,
                                                             RefineEntry* __restrict__ out,
                                                             unsigned* __restrict__ count, unsigned cap) {
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= N || (insub && insub[i])) return;
  float* row = Lc + (size_t)i * topk * slots;
  for (int t0 = 0; t0 < topk * slots; t0 += 64) {
    const int t = t0 + lane;
    bool take = false;
    int r = 0;
    float p = 0.0f;
    if (t < topk * slots) {
      const int j = t / slots, s = t - j * slots;
      const int c = cand[(size_t)i * topk + j];
      if (s < cell_off[c + 1] - cell_off[c]) {
        p = row[t];
        r = cell_rot[cell_off[c] + s];
        take = p >= tau;
      }
    }
    push(out, count, cap, take, row + (t < topk * slots ? t : 0), r, i, p);
  }
}

// ---- FP64 likelihood of each entry; entries per pattern ----
__global__ __launch_bounds__(256) void k_refine_like(RefineEntry* __restrict__ ent, const unsigned* __restrict__ count,
                                                     const float* __restrict__ slices, int N2d,
                                                     const int* __restrict__ off, const int* __restrict__ px,
                                                     const float* __restrict__ cnt, const double* __restrict__ A,
                                                     const double* __restrict__ B, const float* __restrict__ phi,
                                                     double inv_sigma2x, int poisson, int* __restrict__ npat) {
  const int lane = threadIdx.x & 63;
  const unsigned e = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (e >= *count) return;
  RefineEntry& en = ent[e];
  const int r = en.r, i = en.i;
  const float* srow = slices + (size_t)r * N2d;
  double c = 0.0;
  for (int k = off[i] + lane; k < off[i + 1]; k += 64) {
    const double s = (double)srow[px[k]];
    c += (double)cnt[k] * (poisson ? log(s > (double)kSlogFloor ? s : (double)kSlogFloor) : s);
  }
  c = wave_sum(c);
  if (lane == 0) {
    const double f = (double)phi[i];
    const double L = poisson ? c + B[i] * log(f) - f * A[r] : -(A[r] - 2.0 * c / f + B[i] / (f * f)) * inv_sigma2x;
    en.C = c;
    en.L = L;
    atomicAdd(&npat[i], 1);
  }
}

// Entries grouped by pattern: slot[pat_off[i] .. pat_off[i + 1]) holds the
// indices of pattern i's entries (in no particular order; the per-pattern
// pass orders them by rotation, so every sum below is formed in one fixed
// order and the result does not depend on the order of the atomics).
__global__ __launch_bounds__(256) void k_refine_scatter(const RefineEntry* __restrict__ ent,
                                                        const unsigned* __restrict__ count,
                                                        const int* __restrict__ pat_off, int* __restrict__ fill,
                                                        int* __restrict__ slot) {
  const unsigned e = blockIdx.x * 256 + threadIdx.x;
  if (e >= *count) return;
  const int i = ent[e].i;
  slot[pat_off[i] + atomicAdd(&fill[i], 1)] = (int)e;
}

// One thread per pattern: its entries in rotation order, the largest L, the
// renormalised responsibilities written back, the correction of the
// scaling denominator and of sum P log P (dtot[i]).
__global__ __launch_bounds__(256) void k_refine_pattern(const RefineEntry* __restrict__ ent,
                                                        const int* __restrict__ pat_off, int* __restrict__ slot,
                                                        int N, float* __restrict__ respons,
                                                        const double* __restrict__ A, int poisson,
                                                        double* __restrict__ phiden, double* __restrict__ dtot) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= N) return;
  const int s0 = pat_off[i], s1 = pat_off[i + 1];
  double d = 0.0;
  if (s1 > s0) {
    for (int a = s0 + 1; a < s1; a++) {  // insertion sort by rotation
      const int v = slot[a];
      const int rv = ent[v].r;
      int b = a - 1;
      while (b >= s0 && ent[slot[b]].r > rv) {
        slot[b + 1] = slot[b];
        b--;
      }
      slot[b + 1] = v;
    }
    double m = -INFINITY, pold = 0.0;
    for (int a = s0; a < s1; a++) {
      m = fmax(m, ent[slot[a]].L);
      pold += (double)ent[slot[a]].p;
    }
    double w = 0.0;
    for (int a = s0; a < s1; a++) w += exp(ent[slot[a]].L - m);
    const double keep = 1.0 - fmax(0.0, 1.0 - pold);  // mass of the refined entries
    double dphi = 0.0;
    for (int a = s0; a < s1; a++) {
      const RefineEntry en = ent[slot[a]];
      const float p = (float)(keep * exp(en.L - m) / w);
      respons[(size_t)en.r * N + i] = p;
      if (en.src) *en.src = p;
      dphi += ((double)p - (double)en.p) * (poisson ? A[en.r] : en.C);
      // sum P log P as the softmax forms it (FP32 product and log)
      d += (p > 0.0f ? (double)(p * logf(p)) : 0.0) - (en.p > 0.0f ? (double)(en.p * logf(en.p)) : 0.0);
    }
    if (phiden) phiden[i] += dphi;
  }
  dtot[i] = d;
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

static void launch_collect(const RefineInputs& in, unsigned cap) {
  Engine& e = engine();
  RefineEntry* out = e.d_ref_ent.as<RefineEntry>();
  unsigned* count = e.d_ref_count.as<unsigned>();
  for (int k = 0; k < in.nrows; k++) {
    const RefineRows& s = in.rows[k];
    if (!s.P || s.rows <= 0 || s.cols <= 0) continue;
    const size_t n = (size_t)s.rows * s.cols;
    const unsigned nb = (unsigned)std::min<size_t>(8192, (n + 255) / 256);
    hipLaunchKernelGGL(k_refine_collect_rows, dim3(nb), dim3(256), 0, e.stream, s.P, s.rows, s.ld, s.cols,
                       s.row_rot, s.col_pat, s.skip, (int)s.own, in.tau, out, count, cap);
    EMC_LAUNCHED();
  }
  if (in.Lc) {
    hipLaunchKernelGGL(k_refine_collect_cand, dim3((in.N + 3) / 4), dim3(256), 0, e.stream, in.Lc, in.N, in.topk,
                       in.slots, in.cand, in.cell_off, in.cell_rot, in.insub, in.tau, out, count, cap);
    EMC_LAUNCHED();
  }
}

bool refine_posterior(const RefineInputs& in, double* dtotal) {
  Engine& e = engine();
  hipStream_t st = e.stream;
  *dtotal = 0.0;
  if (!ensure_counts(in.images, in.N, in.N2d)) return false;
  static unsigned* h_count = nullptr;
  static double* h_tot = nullptr;
  if (!h_count) {
    if (hipHostMalloc((void**)&h_count, sizeof(unsigned) * 4, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&h_tot, sizeof(double) * 2, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      set_error("refinement: page-locked host memory");
      return false;
    }
  }
  unsigned cap = (unsigned)std::max<size_t>(e.d_ref_ent.bytes / sizeof(RefineEntry), 0);
  const unsigned want = (unsigned)std::min<size_t>((size_t)8 * in.N + 65536, 1u << 28);
  if (cap < want) {
    if (!e.d_ref_ent.ensure(sizeof(RefineEntry) * (size_t)want)) return false;
    cap = want;
  }
  if (!e.d_ref_count.ensure(16) ||
      !e.d_ref_pat.ensure(round_up(sizeof(int) * (3 * (size_t)in.N + 2), 16) + sizeof(double) * (size_t)in.N) ||
      !e.d_ref_tot.ensure(sizeof(double) * 2))
    return false;
  for (int pass = 0; pass < 2; pass++) {
    launch_fill(e.d_ref_count.as<float>(), 1, 0.0f, st);
    launch_collect(in, cap);
    EMC_CHECK(hipMemcpyAsync(h_count, e.d_ref_count.p, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    EMC_CHECK(hipStreamSynchronize(st));
    if (failed()) return false;
    if (*h_count <= cap) break;
    cap = *h_count + 1024;  // room for every entry: collect again
    if (!e.d_ref_ent.ensure(sizeof(RefineEntry) * (size_t)cap)) return false;
  }
  if (!e.d_ref_slot.ensure(sizeof(int) * (size_t)std::max(*h_count, 1u))) return false;
  if (*h_count > cap) {
    set_error("refinement: %u entries exceed the list of %u", *h_count, cap);
    return false;
  }
  const unsigned n = *h_count;
  e.ref_entries = n;
  if (n == 0) return true;
  const int N = in.N;
  // per pattern: entry count, offsets (N + 1), fill counters, sum P log P change
  int* npat = e.d_ref_pat.as<int>();  // N + 1 (the last stays 0)
  int* pat_off = npat + (N + 1);
  int* fill = pat_off + (N + 1);
  double* dtot = reinterpret_cast<double*>(e.d_ref_pat.as<char>() + round_up(sizeof(int) * (3 * (size_t)N + 2), 16));
  int* slot = e.d_ref_slot.as<int>();
  launch_fill(e.d_ref_pat.as<float>(), 3 * (size_t)N + 2, 0.0f, st);
  const unsigned* count = e.d_ref_count.as<unsigned>();
  RefineEntry* ent = e.d_ref_ent.as<RefineEntry>();
  hipLaunchKernelGGL(k_refine_like, dim3((n + 3) / 4), dim3(256), 0, st, ent, count, in.slices, in.N2d,
                     e.d_cnt_off.as<int>(), e.d_cnt_px.as<int>(), e.d_cnt_val.as<float>(), e.d_A.as<double>(),
                     e.d_B.as<double>(), e.d_phi_used.as<float>(), in.sigma2x > 0.0 ? 1.0 / in.sigma2x : 0.0,
                     (int)in.poisson, npat);
  EMC_LAUNCHED();
  size_t tbytes = 0;
  EMC_CHECK(rocprim::exclusive_scan(nullptr, tbytes, npat, pat_off, 0, (size_t)N + 1, rocprim::plus<int>(), st));
  if (!e.d_ref_tmp.ensure(std::max<size_t>(tbytes, 16))) return false;
  EMC_CHECK(rocprim::exclusive_scan(e.d_ref_tmp.p, tbytes, npat, pat_off, 0, (size_t)N + 1, rocprim::plus<int>(),
                                    st));
  hipLaunchKernelGGL(k_refine_scatter, dim3((n + 255) / 256), dim3(256), 0, st, ent, count, pat_off, fill, slot);
  EMC_LAUNCHED();
  hipLaunchKernelGGL(k_refine_pattern, dim3((N + 255) / 256), dim3(256), 0, st, ent, pat_off, slot, N, in.respons,
                     e.d_A.as<double>(), (int)in.poisson, in.phiden, dtot);
  EMC_LAUNCHED();
  launch_sum_f64(dtot, N, e.d_ref_tot.as<double>(), st);
  EMC_CHECK(hipMemcpyAsync(h_tot, e.d_ref_tot.p, sizeof(double), hipMemcpyDeviceToHost, st));
  EMC_CHECK(hipStreamSynchronize(st));
  if (failed()) return false;
  *dtotal = *h_tot;
  return true;
}

}  // namespace emc
