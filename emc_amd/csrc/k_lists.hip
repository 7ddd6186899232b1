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
__global__ __launch_bounds__(256) void k_recentre_cand(const int* __restrict__ top, const float* __restrict__ topm,
                                                       int N, int ktop, const int* __restrict__ nbr, int nnbr,
                                                       const int* __restrict__ net,
                                                       const unsigned char* __restrict__ shared, int C, int topk,
                                                       int* __restrict__ cand, int* __restrict__ moved) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= N) return;
  int next[kMaxTopk];
  int n = 0;
  auto add = [&](int c) {
    if (c < 0 || c >= C || n >= topk || shared[c]) return;
    for (int q = 0; q < n; q++)
      if (next[q] == c) return;
    next[n++] = c;
  };
  const int* ti = top + (size_t)i * ktop;
  const float* mi = topm + (size_t)i * ktop;
  int used = 0;
  for (int t = 0; t < ktop; t++)
    if (ti[t] >= 0 && ti[t] < C && (t == 0 || mi[t] > 1e-6f * mi[0])) {
      add(ti[t]);
      used = t + 1;
    }
  for (int t = 0; t < used; t++)
    if (ti[t] >= 0 && ti[t] < C)
      for (int g = 0; g < nnbr; g++) add(nbr[(size_t)ti[t] * nnbr + g]);
  for (int j = 0; j < topk; j++) add(net[(size_t)i * topk + j]);
  for (int q = 0; q < n && n < topk; q++)
    for (int g = 0; g < nnbr; g++) add(nbr[(size_t)next[q] * nnbr + g]);
  for (int c = 0; c < C && n < topk; c++) add(c);
  int* ci = cand + (size_t)i * topk;
  bool same = true;
  for (int j = 0; j < topk; j++) same = same && ci[j] == next[j];
  if (!same) {
    for (int j = 0; j < topk; j++) ci[j] = next[j];
    atomicAdd(moved, 1);
  }
}

void launch_recentre_cand(const int* top, const float* topm, int N, int ktop, const int* nbr, int nnbr,
                          const int* net, const unsigned char* shared, int C, int topk, int* cand, int* moved,
                          hipStream_t st) {
  hipLaunchKernelGGL(k_recentre_cand, dim3((N + 255) / 256), dim3(256), 0, st, top, topm, N, ktop, nbr, nnbr, net,
                     shared, C, topk, cand, moved);
  EMC_LAUNCHED();
}

// One thread per pattern: entry counts of every cell (count[C], zeroed
// beforehand; count[C] counts invalid cells) and the pattern's candidate
// rotation count.
__global__ __launch_bounds__(256) void k_cand_hist(const int* __restrict__ cand, int N, int topk, int C,
                                                   const int* __restrict__ cell_off, int* __restrict__ count,
                                                   int* __restrict__ ncand) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= N) return;
  int s = 0;
  for (int j = 0; j < topk; j++) {
    const int c = cand[(size_t)i * topk + j];
    if (c < 0 || c >= C) {
      atomicAdd(&count[C], 1);
      continue;
    }
    atomicAdd(&count[c], 1);
    s += cell_off[c + 1] - cell_off[c];
  }
  ncand[i] = s;
}

void launch_cand_hist(const int* cand, int N, int topk, int C, const int* cell_off, int* count, int* ncand,
                      hipStream_t st) {
  launch_fill(reinterpret_cast<float*>(count), (size_t)C + 1, 0.0f, st);
  hipLaunchKernelGGL(k_cand_hist, dim3((N + 255) / 256), dim3(256), 0, st, cand, N, topk, C, cell_off, count, ncand);
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

}  // namespace emc
