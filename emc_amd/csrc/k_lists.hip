// Candidate lists of the rotation proposal, built on the GPU every
// iteration: the shared-cell support, the re-centred candidate cells of
// every pattern, and the cell-major entry lists (a stable radix sort of the
// (cell, pattern, rank) entries by cell, so that every cell lists its
// patterns in ascending order, ranks ascending).
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include "engine.hpp"
#include "proposal_kernels.hpp"
#include "reduce.hpp"

namespace emc {

// mark[c] = 1 for every cell that is some pattern's best cell or holds at
// least share_mass of its posterior (mark zeroed beforehand).
__global__ __launch_bounds__(256) void k_mark_support(const int* __restrict__ top, const float* __restrict__ topm,
                                                      long n, int ktop, float share_mass, int C,
                                                      unsigned char* __restrict__ mark) {
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  const int c = top[t];
  if (c >= 0 && c < C && (t % ktop == 0 || topm[t] >= share_mass)) mark[c] = 1;
}

void launch_mark_support(const int* top, const float* topm, int N, int ktop, float share_mass, int C,
                         unsigned char* mark, hipStream_t st) {
  launch_fill(reinterpret_cast<float*>(mark), (size_t)(C + 3) / 4, 0.0f, st);
  const long n = (long)N * ktop;
  hipLaunchKernelGGL(k_mark_support, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, top, topm, n, ktop,
                     share_mass, C, mark);
  EMC_LAUNCHED();
}

// One thread per pattern: its topk cells, in this order and skipping shared
// and repeated cells: the (up to ktop) cells holding most of its last
// posterior (a cell after the first only with > 1e-6 of the first's mass),
// the nearest cells of each of those, the network's cells in rank order,
// the nearest cells of the cells taken so far, the lowest-numbered free
// cells.  moved counts the patterns whose list changed.
//
// Here one wave per pattern: lane q holds list entry q (topk <= 64), and
// the offered cells come 64 at a time, one a lane, in offer order.  A lane's
// cell passes the range and shared tests in parallel; the passing cells are
// then taken one by one in lane order, each after one ballot over the list
// for repeats, until topk are held.  (At least topk cells are free: the
// shared set leaves them.)
__device__ __forceinline__ void recentre_take(int c, bool offer, int C, const unsigned char* __restrict__ shared,
                                              int topk, int lane, int& lst, int& n) {
  const bool ok = offer && c >= 0 && c < C && !shared[c];
  unsigned long long m = __ballot(ok);
  while (m && n < topk) {
    const int l = __ffsll((long long)m) - 1;
    m &= m - 1;
    const int cl = __builtin_amdgcn_readlane(c, l);
    if (__ballot(lane < n && lst == cl) == 0) {
      if (lane == n) lst = cl;
      n++;
    }
  }
}

__global__ __launch_bounds__(256) void k_recentre_cand(const int* __restrict__ top, const float* __restrict__ topm,
                                                       int N, int ktop, const int* __restrict__ nbr, int nnbr,
                                                       const int* __restrict__ net,
                                                       const unsigned char* __restrict__ shared, int C, int topk,
                                                       int* __restrict__ cand, int* __restrict__ moved) {
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= N) return;
  int lst = -1, n = 0;
  const int* ti = top + (size_t)i * ktop;
  const float* mi = topm + (size_t)i * ktop;
  const float m0 = mi[0];
  int used = 0;
  for (int t0 = 0; t0 < ktop; t0 += 64) {
    const int t = t0 + lane;
    const int c = t < ktop ? ti[t] : -1;
    const bool pick = t < ktop && c >= 0 && c < C && (t == 0 || mi[t] > 1e-6f * m0);
    const unsigned long long b = __ballot(pick);
    if (b) used = t0 + 64 - __clzll((long long)b);
    recentre_take(c, pick, C, shared, topk, lane, lst, n);
  }
  for (int f0 = 0; f0 < used * nnbr; f0 += 64) {
    const int f = f0 + lane;
    const int t = f / nnbr;
    const int c = f < used * nnbr ? ti[t] : -1;
    const bool has = c >= 0 && c < C;
    const int d = has ? nbr[(size_t)c * nnbr + (f - t * nnbr)] : -1;
    recentre_take(d, has, C, shared, topk, lane, lst, n);
  }
  for (int j0 = 0; j0 < topk && n < topk; j0 += 64) {
    const int j = j0 + lane;
    const int c = j < topk ? net[(size_t)i * topk + j] : -1;
    recentre_take(c, j < topk, C, shared, topk, lane, lst, n);
  }
  // nearest cells of the held cells, the held list growing meanwhile: a
  // batch offers the cells of entries below the count at its start, and
  // entries taken in it are offered in a later batch
  for (int f0 = 0; n < topk;) {
    const int fend = n * nnbr;
    if (f0 >= fend) break;
    const int f = f0 + lane;
    const int q = f / nnbr;
    const bool has = f < fend;
    const int cq = __shfl(lst, has ? q : 0);
    const int d = has ? nbr[(size_t)cq * nnbr + (f - q * nnbr)] : -1;
    recentre_take(d, has, C, shared, topk, lane, lst, n);
    f0 = min(f0 + 64, fend);
  }
  for (int c0 = 0; c0 < C && n < topk; c0 += 64) recentre_take(c0 + lane, c0 + lane < C, C, shared, topk, lane, lst, n);
  int* ci = cand + (size_t)i * topk;
  const bool differ = lane < topk && ci[lane] != lst;
  if (__ballot(differ)) {
    if (lane < topk) ci[lane] = lst;
    if (lane == 0) atomicAdd(moved, 1);
  }
}

void launch_recentre_cand(const int* top, const float* topm, int N, int ktop, const int* nbr, int nnbr,
                          const int* net, const unsigned char* shared, int C, int topk, int* cand, int* moved,
                          hipStream_t st) {
  hipLaunchKernelGGL(k_recentre_cand, dim3((N + 3) / 4), dim3(256), 0, st, top, topm, N, ktop, nbr, nnbr, net,
                     shared, C, topk, cand, moved);
  EMC_LAUNCHED();
}

// One thread per pattern: its candidate rotation count; count[C] (zeroed
// beforehand) counts entries naming no valid cell (c == C is allowed as the
// no-entry key and counted there too).
__global__ __launch_bounds__(256) void k_cand_hist(const int* __restrict__ cand, int N, int topk, int C,
                                                   const int* __restrict__ cell_off, int* __restrict__ count,
                                                   int* __restrict__ ncand) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= N) return;
  int s = 0, bad = 0;
  for (int j = 0; j < topk; j++) {
    const int c = cand[(size_t)i * topk + j];
    if (c < 0 || c >= C) {
      bad++;
      continue;
    }
    s += cell_off[c + 1] - cell_off[c];
  }
  ncand[i] = s;
  if (bad) atomicAdd(&count[C], bad);
}

void launch_cand_hist(const int* cand, int N, int topk, int C, const int* cell_off, int* count, int* ncand,
                      hipStream_t st) {
  launch_fill(reinterpret_cast<float*>(count), (size_t)C + 1, 0.0f, st);
  hipLaunchKernelGGL(k_cand_hist, dim3((N + 255) / 256), dim3(256), 0, st, cand, N, topk, C, cell_off, count, ncand);
  EMC_LAUNCHED();
}

// count[c] for c < C from the sorted keys (first index with key > c minus
// first index with key >= c); the entries were sorted already
__global__ __launch_bounds__(256) void k_counts_sorted(const int* __restrict__ keys, long n, int C,
                                                       int* __restrict__ count) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= C) return;
  auto lower = [&](int v) {
    long lo = 0, hi = n;
    while (lo < hi) {
      const long mid = (lo + hi) >> 1;
      if (keys[mid] < v) lo = mid + 1;
      else hi = mid;
    }
    return lo;
  };
  count[c] = (int)(lower(c + 1) - lower(c));
}

void launch_counts_sorted(const int* keys, long n, int C, int* count, hipStream_t st) {
  hipLaunchKernelGGL(k_counts_sorted, dim3((C + 255) / 256), dim3(256), 0, st, keys, n, C, count);
  EMC_LAUNCHED();
}

// Entry e of the cell-sorted list: pattern v / topk, rank v % topk.
__global__ __launch_bounds__(256) void k_split_entries(const int* __restrict__ v, long n, int topk,
                                                       int* __restrict__ pat, int* __restrict__ jj) {
  const long e = (long)blockIdx.x * 256 + threadIdx.x;
  if (e >= n) return;
  pat[e] = v[e] / topk;
  jj[e] = v[e] % topk;
}

bool sort_entries(const int* cand, int N, int topk, int C, DBuf& keys_sorted, DBuf& vals_sorted, DBuf& temp,
                  int* ent_pat, int* ent_j, hipStream_t st) {
  const long n = (long)N * topk;
  int bits = 1;
  while ((1 << bits) < C) bits++;
  if (!keys_sorted.ensure(sizeof(int) * std::max(n, 1L)) || !vals_sorted.ensure(sizeof(int) * std::max(n, 1L)))
    return false;
  rocprim::counting_iterator<int> ids(0);
  size_t tbytes = 0;
  EMC_CHECK(rocprim::radix_sort_pairs(nullptr, tbytes, cand, keys_sorted.as<int>(), ids, vals_sorted.as<int>(),
                                      (size_t)n, 0, bits, st));
  if (!temp.ensure(std::max(tbytes, (size_t)1))) return false;
  EMC_CHECK(rocprim::radix_sort_pairs(temp.p, tbytes, cand, keys_sorted.as<int>(), ids, vals_sorted.as<int>(),
                                      (size_t)n, 0, bits, st));
  hipLaunchKernelGGL(k_split_entries, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, vals_sorted.as<int>(), n,
                     topk, ent_pat, ent_j);
  EMC_LAUNCHED();
  return !failed();
}

// ---------------------------------------------------------------------------
// Nonzero responsibilities of a proposal iteration, per rotation, from where
// they were computed (candidate slots, shared rows, whole-table subset)
// instead of a scan of the dense [R][N] matrix.  Pass 1 counts per rotation;
// pass 2 appends (key = r * N + i, P) in any order; a radix sort by key then
// gives every rotation's list in pattern order, independent of that order.
// ---------------------------------------------------------------------------
// mode 0: count into rownnz; mode 1: append to keys / vals at *tail.  Called
// by every thread of the workgroup (256) for its kNzItems entries, entry q
// of thread t being the workgroup's entry q * 256 + t (take[q] = it exists):
// one atomic per workgroup for the append position, and for the count when
// r is the same for all its entries (r_uniform), else one per nonzero.  The
// append order is irrelevant: the keys are unique and sorted afterwards.
constexpr int kNzItems = 4;

__device__ __forceinline__ void nz_take(int mode, const bool (&take)[kNzItems], const float (&p)[kNzItems],
                                        const int (&r)[kNzItems], const int (&i)[kNzItems], int N, bool r_uniform,
                                        int* rownnz, unsigned* tail, unsigned* keys, float* vals) {
  __shared__ unsigned s_cnt[kNzItems][4];
  __shared__ unsigned s_base;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  unsigned long long m[kNzItems];
#pragma unroll
  for (int q = 0; q < kNzItems; q++) m[q] = __ballot(take[q] && p[q] != 0.0f);
  if (mode == 0 && !r_uniform) {
#pragma unroll
    for (int q = 0; q < kNzItems; q++)
      if (m[q] >> lane & 1) atomicAdd(&rownnz[r[q]], 1);
    return;
  }
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < kNzItems; q++) s_cnt[q][w] = (unsigned)__popcll(m[q]);
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned tot = 0;
#pragma unroll
    for (int q = 0; q < kNzItems; q++) tot += s_cnt[q][0] + s_cnt[q][1] + s_cnt[q][2] + s_cnt[q][3];
    if (mode == 0) {
      if (tot) atomicAdd(&rownnz[r[0]], (int)tot);
    } else {
      s_base = tot ? atomicAdd(tail, tot) : 0u;
    }
  }
  if (mode == 0) return;
  __syncthreads();
  unsigned at = s_base;
  const unsigned long long below = (1ull << lane) - 1ull;
#pragma unroll
  for (int q = 0; q < kNzItems; q++) {
    unsigned mine = at;
    for (int v = 0; v < 4; v++) {
      if (v == w) mine = at;
      at += s_cnt[q][v];
    }
    if (m[q] >> lane & 1) {
      const unsigned k = mine + (unsigned)__popcll(m[q] & below);
      keys[k] = (unsigned)r[q] * (unsigned)N + (unsigned)i[q];
      vals[k] = p[q];
    }
  }
}

__global__ __launch_bounds__(256) void k_nz_cand(NzSources a, int mode, int* rownnz, unsigned* tail, unsigned* keys,
                                                 float* vals) {
  const long nc = (long)a.N * a.topk * kSlot;
  bool take[kNzItems];
  int r[kNzItems], i[kNzItems];
  float p[kNzItems];
#pragma unroll
  for (int q = 0; q < kNzItems; q++) {
    const long t = (long)blockIdx.x * (256 * kNzItems) + q * 256 + threadIdx.x;  // (i, j, slot)
    take[q] = t < nc;
    r[q] = 0;
    i[q] = 0;
    p[q] = 0.0f;
    if (take[q]) {
      const int s = (int)(t % kSlot);
      const int ij = (int)(t / kSlot);  // < N * topk, the size of cand
      i[q] = ij / a.topk;
      const int c = a.cand[ij];
      const int o = a.cell_off[c];
      take[q] = !a.insub[i[q]] && s < a.cell_off[c + 1] - o;
      if (take[q]) {
        p[q] = a.Pc[t];
        r[q] = a.cell_rot[o + s];
      }
    }
  }
  nz_take(mode, take, p, r, i, a.N, false, rownnz, tail, keys, vals);
}

__global__ __launch_bounds__(256) void k_nz_shared(NzSources a, int mode, int* rownnz, unsigned* tail,
                                                   unsigned* keys, float* vals) {
  const int rs = blockIdx.y;
  const int rot = a.sh_rot[rs];
  bool take[kNzItems];
  int r[kNzItems], i[kNzItems];
  float p[kNzItems];
#pragma unroll
  for (int q = 0; q < kNzItems; q++) {
    i[q] = blockIdx.x * (256 * kNzItems) + q * 256 + threadIdx.x;
    take[q] = i[q] < a.N && !a.insub[i[q]];
    p[q] = take[q] ? a.Psh[(size_t)rs * a.N + i[q]] : 0.0f;
    r[q] = rot;
  }
  nz_take(mode, take, p, r, i, a.N, true, rownnz, tail, keys, vals);
}

__global__ __launch_bounds__(256) void k_nz_sub(NzSources a, int mode, int* rownnz, unsigned* tail, unsigned* keys,
                                                float* vals) {
  const int rot = blockIdx.x;
  bool take[kNzItems];
  int r[kNzItems], i[kNzItems];
  float p[kNzItems];
#pragma unroll
  for (int q = 0; q < kNzItems; q++) {
    const int k = blockIdx.y * (256 * kNzItems) + q * 256 + threadIdx.x;
    take[q] = k < a.n_sub;
    p[q] = take[q] ? a.Psub[(size_t)rot * a.n_sub + k] : 0.0f;
    i[q] = take[q] ? a.sub[k] : 0;
    r[q] = rot;
  }
  nz_take(mode, take, p, r, i, a.N, true, rownnz, tail, keys, vals);
}

static void nz_pass(const NzSources& a, int mode, int* rownnz, unsigned* tail, unsigned* keys, float* vals,
                    hipStream_t st) {
  constexpr int span = 256 * kNzItems;
  const long nc = (long)a.N * a.topk * kSlot;
  hipLaunchKernelGGL(k_nz_cand, dim3((unsigned)((nc + span - 1) / span)), dim3(256), 0, st, a, mode, rownnz, tail,
                     keys, vals);
  EMC_LAUNCHED();
  if (a.R_sh > 0) {
    hipLaunchKernelGGL(k_nz_shared, dim3((a.N + span - 1) / span, a.R_sh), dim3(256), 0, st, a, mode, rownnz, tail,
                       keys, vals);
    EMC_LAUNCHED();
  }
  if (a.n_sub > 0) {
    hipLaunchKernelGGL(k_nz_sub, dim3(a.R, (a.n_sub + span - 1) / span), dim3(256), 0, st, a, mode, rownnz, tail,
                       keys, vals);
    EMC_LAUNCHED();
  }
}

void launch_nz_count(const NzSources& a, int* rownnz, hipStream_t st) {
  launch_fill(reinterpret_cast<float*>(rownnz), (size_t)a.R, 0.0f, st);
  nz_pass(a, 0, rownnz, nullptr, nullptr, nullptr, st);
}

// Sorted lists: col[e] = pattern, val[e] = P * opscale[pattern], and per
// rotation T_r = sum P (FP64), Tf, and with phi Tphi_r = sum P phi.
__global__ __launch_bounds__(256) void k_nz_split(const unsigned* __restrict__ keys, const float* __restrict__ P,
                                                  long n, int N, const float* __restrict__ opscale,
                                                  int* __restrict__ col, float* __restrict__ val) {
  const long e = (long)blockIdx.x * 256 + threadIdx.x;
  if (e >= n) return;
  const int i = (int)(keys[e] % (unsigned)N);
  col[e] = i;
  val[e] = P[e] * opscale[i];
}

// one wave per rotation (lanes along its list, fixed-order wave sums)
__global__ __launch_bounds__(256) void k_nz_rowsums(const int* __restrict__ off, const int* __restrict__ col,
                                                    const float* __restrict__ P, int R, const float* __restrict__ phi,
                                                    double* __restrict__ T, float* __restrict__ Tf,
                                                    float* __restrict__ Tphi) {
  const int lane = threadIdx.x & 63;
  const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= R) return;
  double t = 0.0, tp = 0.0;
  for (int e = off[r] + lane; e < off[r + 1]; e += 64) {
    t += P[e];
    if (phi) tp += (double)P[e] * (double)phi[col[e]];
  }
  t = wave_sum(t);
  if (phi) tp = wave_sum(tp);
  if (lane == 0) {
    T[r] = t;
    Tf[r] = (float)t;
    if (phi) Tphi[r] = (float)tp;
  }
}

__global__ __launch_bounds__(256) void k_zero_keys(const unsigned* __restrict__ keys, long n,
                                                   float* __restrict__ respons) {
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t < n) respons[keys[t]] = 0.0f;
}

void launch_zero_keys(const unsigned* keys, long n, float* respons, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_zero_keys, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, keys, n, respons);
  EMC_LAUNCHED();
}

bool build_nz_lists(const NzSources& a, long nnz, const int* off, const float* opscale, const float* phi,
                    DBuf& keys, DBuf& keys_sorted, DBuf& vals, DBuf& vals_sorted, DBuf& tail, DBuf& temp, int* col,
                    float* val, double* T, float* Tf, float* Tphi, hipStream_t st) {
  const long n = std::max(nnz, 1L);
  if (!keys.ensure(4 * n) || !keys_sorted.ensure(4 * n) || !vals.ensure(4 * n) || !vals_sorted.ensure(4 * n) ||
      !tail.ensure(4))
    return false;
  launch_fill(tail.as<float>(), 1, 0.0f, st);
  nz_pass(a, 1, nullptr, tail.as<unsigned>(), keys.as<unsigned>(), vals.as<float>(), st);
  int bits = 1;
  while (bits < 32 && ((uint64_t)1 << bits) < (uint64_t)a.R * (uint64_t)a.N) bits++;
  size_t tbytes = 0;
  EMC_CHECK(rocprim::radix_sort_pairs(nullptr, tbytes, keys.as<unsigned>(), keys_sorted.as<unsigned>(),
                                      vals.as<float>(), vals_sorted.as<float>(), (size_t)nnz, 0, bits, st));
  if (!temp.ensure(std::max(tbytes, (size_t)1))) return false;
  EMC_CHECK(rocprim::radix_sort_pairs(temp.p, tbytes, keys.as<unsigned>(), keys_sorted.as<unsigned>(),
                                      vals.as<float>(), vals_sorted.as<float>(), (size_t)nnz, 0, bits, st));
  if (nnz > 0) {
    hipLaunchKernelGGL(k_nz_split, dim3((unsigned)((nnz + 255) / 256)), dim3(256), 0, st, keys_sorted.as<unsigned>(),
                       vals_sorted.as<float>(), nnz, a.N, opscale, col, val);
    EMC_LAUNCHED();
  }
  hipLaunchKernelGGL(k_nz_rowsums, dim3((a.R + 3) / 4), dim3(256), 0, st, off, col, vals_sorted.as<float>(), a.R,
                     phi, T, Tf, Tphi);
  EMC_LAUNCHED();
  return !failed();
}

}  // namespace emc
