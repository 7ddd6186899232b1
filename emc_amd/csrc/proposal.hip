// Learned rotation proposal: the network's C-ABI, training on EMC's own
// responsibilities, prediction of candidate coarse cells, and the
// candidate-restricted iteration (E-step, softmax, audit, scaling, M-step).
// None of this has a counterpart in the reference.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

#include "cnn.hpp"
#include "emc_host.h"
#include "engine.hpp"
#include "proposal.hpp"
#include "proposal_kernels.hpp"

namespace emc {

Proposal& proposal() {
  static Proposal* p = new Proposal();
  return *p;
}

void Proposal::release() {
  DBuf* bufs[] = {&d_cell_off, &d_cell_rot, &d_cand, &d_pmass, &d_top, &d_topm, &d_ent_off, &d_ent_pat, &d_ent_j, &d_work,
                  &d_chunk_off, &d_chunk_cell, &d_seg_cell, &d_seg_chunk0, &d_seg_nch, &d_seg_part, &d_mcell,
                  &d_mcell_part, &d_part, &d_part_T, &d_part_Tphi, &d_part_max, &d_Lc, &d_panels, &d_crowscale, &d_plogp, &d_zero, &d_sub, &d_insub, &d_Ksub,
                  &d_Ksublo, &d_Ktsub, &d_Ktsublo, &d_Bsub, &d_phisub, &d_invphisub, &d_Lsub, &d_phiden_sub,
                  &d_Psub16, &d_Psub16lo, &d_mass, &d_tot, &d_iota, &d_part_c, &d_part_m, &d_sh_cells, &d_sh_off, &d_sh_rot,
                  &d_Ssh, &d_Sshlo, &d_Ash, &d_sscale_sh, &d_Lsh, &d_top_s, &d_topm_s, &d_net_cand,
                  &d_nbr, &d_mark, &d_shared, &d_count, &d_ncand, &d_moved, &d_keys_s, &d_vals_s, &d_sort_tmp, &d_colmap,
                  &d_gkey, &d_gcount, &d_gpat, &d_gkey_s, &d_ghist, &d_gcore, &d_gncore, &d_gnrot, &d_gcofs, &d_gtrunc, &d_crot,
                  &d_ccell, &d_cslot, &d_rank, &d_rkey, &d_gwork, &d_rcount, &d_rncand, &d_rent_off, &d_rent_pat,
                  &d_rent_j, &d_rwork, &d_rkeys_s, &d_rvals_s, &d_targets, &d_perm};
  for (DBuf* b : bufs) b->release();
  cnn.release();
  configured = predicted = have_p = false;
  post = POST_NONE;
  clear_shared();
}

// Pinned staging for the host -> device uploads of the list building: a
// copy from pinned memory is queued without blocking the host (one from
// pageable memory is staged by the runtime first), so consecutive uploads
// and the kernels after them follow each other on the GPU.  The space is
// reused after the next stream synchronisation.
struct PinnedArena {
  char* base = nullptr;
  size_t cap = 0, used = 0;
};
static PinnedArena g_pin;

static void sync_stream() {
  EMC_CHECK(hipStreamSynchronize(engine().stream));
  g_pin.used = 0;
}

static void* pinned_take(size_t bytes) {
  bytes = (bytes + 255) / 256 * 256;
  if (g_pin.used + bytes > g_pin.cap) {
    sync_stream();  // every queued copy from the arena is done
    if (bytes > g_pin.cap) {
      if (g_pin.base) (void)hipHostFree(g_pin.base);
      g_pin.base = nullptr;
      g_pin.cap = 0;
      const size_t cap = std::max<size_t>(bytes, 4u << 20);
      if (hipHostMalloc((void**)&g_pin.base, cap, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        g_pin.base = nullptr;
        return nullptr;
      }
      g_pin.cap = cap;
    }
  }
  void* at = g_pin.base + g_pin.used;
  g_pin.used += bytes;
  return at;
}

template <class V>
static bool upload(DBuf& d, const V& h, hipStream_t st) {
  using T = typename V::value_type;
  if (!d.ensure(sizeof(T) * std::max<size_t>(h.size(), 1))) return false;
  if (h.empty()) return !failed();
  const size_t bytes = sizeof(T) * h.size();
  void* pin = pinned_take(bytes);
  if (pin) {
    std::memcpy(pin, h.data(), bytes);
    EMC_CHECK(hipMemcpyAsync(d.p, pin, bytes, hipMemcpyHostToDevice, st));
  } else {
    // no pinned memory: copy from the caller's vector and wait for it
    EMC_CHECK(hipMemcpyAsync(d.p, h.data(), bytes, hipMemcpyHostToDevice, st));
    EMC_CHECK(hipStreamSynchronize(st));
  }
  return !failed();
}

// Several host arrays to their device buffers in one host->device copy:
// each hipMemcpyAsync costs tens of microseconds of host time, during which
// the GPU, waiting for these lists, idles.  The arrays go back to back into
// the pinned arena, followed by a table of (destination, offset, bytes);
// one copy moves the block to a staging buffer and k_scatter_segments puts
// each array in place (stream order keeps the staging buffer's uses apart).
struct UploadSeg {
  void* dst;
  unsigned long long src, bytes;
};

__global__ __launch_bounds__(256) void k_scatter_segments(const unsigned char* __restrict__ stage,
                                                          unsigned long long table) {
  const UploadSeg s = reinterpret_cast<const UploadSeg*>(stage + table)[blockIdx.x];
  const unsigned* src = reinterpret_cast<const unsigned*>(stage + s.src);
  unsigned* dst = reinterpret_cast<unsigned*>(s.dst);
  const unsigned long long words = s.bytes / 4;
  for (unsigned long long k = threadIdx.x; k < words; k += 256) dst[k] = src[k];
  if (threadIdx.x < (s.bytes & 3))
    reinterpret_cast<unsigned char*>(s.dst)[words * 4 + threadIdx.x] = stage[s.src + words * 4 + threadIdx.x];
}

class UploadBatch {
 public:
  template <class V>
  bool add(DBuf& d, const V& h) {
    using T = typename V::value_type;
    if (!d.ensure(sizeof(T) * std::max<size_t>(h.size(), 1))) return false;
    if (!h.empty()) items_.push_back({&d, h.data(), sizeof(T) * h.size()});
    return true;
  }
  bool flush(hipStream_t st) {
    if (items_.empty()) return !failed();
    std::vector<UploadSeg> segs;
    size_t total = 0;
    for (const Item& it : items_) {
      segs.push_back({it.d->p, total, it.bytes});
      total += (it.bytes + 255) / 256 * 256;
    }
    const size_t table = total;
    total += segs.size() * sizeof(UploadSeg);
    DBuf& stage = proposal().d_stage;
    char* pin = static_cast<char*>(pinned_take(total));
    if (!pin || !stage.ensure(total)) {
      // no pinned memory: one copy each from the callers' arrays, waited for
      (void)hipGetLastError();
      for (const Item& it : items_) EMC_CHECK(hipMemcpyAsync(it.d->p, it.src, it.bytes, hipMemcpyHostToDevice, st));
      EMC_CHECK(hipStreamSynchronize(st));
      items_.clear();
      return !failed();
    }
    for (size_t k = 0; k < items_.size(); k++) std::memcpy(pin + segs[k].src, items_[k].src, items_[k].bytes);
    std::memcpy(pin + table, segs.data(), segs.size() * sizeof(UploadSeg));
    EMC_CHECK(hipMemcpyAsync(stage.p, pin, total, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_scatter_segments, dim3((unsigned)segs.size()), dim3(256), 0, st,
                       stage.as<unsigned char>(), (unsigned long long)table);
    EMC_LAUNCHED();
    items_.clear();
    return !failed();
  }

 private:
  struct Item {
    DBuf* d;
    const void* src;
    size_t bytes;
  };
  std::vector<Item> items_;
};

// Entries of each cell (patterns ascending), E-step work items, 32-entry
// chunks of the M-step, candidate rotation count of each pattern, from the
// candidate cells in d_cand: per-cell counts on the GPU, the entry lists by
// a stable sort on the GPU, the small per-cell arrays on the host.
static bool build_groups(Proposal& p);

static bool build_lists(Proposal& p) {
  Engine& e = engine();
  hipStream_t st = e.stream;
  const int N = p.N, K = p.topk, C = p.C;
  const long n = (long)N * K;
  if (!p.d_count.ensure(sizeof(int) * (C + 1)) || !p.d_ncand.ensure(sizeof(int) * N) ||
      !p.d_ent_pat.ensure(sizeof(int) * std::max(n, 1L)) || !p.d_ent_j.ensure(sizeof(int) * std::max(n, 1L)))
    return false;
  launch_cand_hist(p.d_cand.as<int>(), N, K, C, p.d_cell_off.as<int>(), p.d_count.as<int>(), p.d_ncand.as<int>(), st);
  if (!sort_entries(p.d_cand.as<int>(), N, K, C, p.d_keys_s, p.d_vals_s, p.d_sort_tmp, p.d_ent_pat.as<int>(),
                    p.d_ent_j.as<int>(), st))
    return false;
  launch_counts_sorted(p.d_keys_s.as<int>(), n, C, p.d_count.as<int>(), st);
  p.h_count.resize(C + 1);
  p.h_ncand.resize(N);
  EMC_CHECK(hipMemcpyAsync(p.h_count.data(), p.d_count.p, sizeof(int) * (C + 1), hipMemcpyDeviceToHost, st));
  EMC_CHECK(hipMemcpyAsync(p.h_ncand.data(), p.d_ncand.p, sizeof(int) * N, hipMemcpyDeviceToHost, st));
  sync_stream();
  if (failed()) return false;
  if (p.h_count[C] > 0) {
    set_error("proposal: %d candidate entries name no valid cell (network output not finite?)", p.h_count[C]);
    return false;
  }
  p.h_ent_off.assign(C + 1, 0);
  for (int c = 0; c < C; c++) p.h_ent_off[c + 1] = p.h_ent_off[c] + p.h_count[c];
  p.n_entries = p.h_ent_off[C];
  p.st_cand_pairs = 0;
  for (int i = 0; i < N; i++) p.st_cand_pairs += p.h_ncand[i];
  std::vector<int2> work;
  p.h_chunk_off.assign(C + 1, 0);
  for (int c = 0; c < C; c++) {
    for (int e0 = p.h_ent_off[c]; e0 < p.h_ent_off[c + 1]; e0 += kCandRows) work.push_back(make_int2(c, e0));
    p.h_chunk_off[c + 1] = p.h_chunk_off[c] + (p.h_ent_off[c + 1] - p.h_ent_off[c] + 31) / 32;
  }
  p.n_work = (int)work.size();
  p.n_chunks = p.h_chunk_off[C];
  std::vector<int> chunk_cell(std::max<long>(p.n_chunks, 1), 0);
  // M-step segments of at most kSegChunks chunks; a cell with one segment
  // (or none: no entries) finishes its slices in the M-step kernel
  std::vector<int> seg_cell, seg_nch, seg_part, mcell, mcell_part(1, 0);
  std::vector<long> seg_chunk0;
  int parts = 0;
  for (int c = 0; c < C; c++) {
    const long q0 = p.h_chunk_off[c], nq = p.h_chunk_off[c + 1] - q0;
    for (long q = q0; q < q0 + nq; q++) chunk_cell[q] = c;
    const bool split = nq > kSegChunks;
    for (long s0 = 0; s0 < std::max<long>(nq, 1); s0 += kSegChunks) {
      seg_cell.push_back(c);
      seg_chunk0.push_back(q0 + s0);
      seg_nch.push_back((int)std::min<long>(kSegChunks, nq - s0 > 0 ? nq - s0 : 0));
      seg_part.push_back(split ? parts++ : -1);
    }
    if (split) {
      mcell.push_back(c);
      mcell_part.push_back(parts);
    }
  }
  p.n_seg = (int)seg_cell.size();
  p.n_mcell = (int)mcell.size();
  p.n_parts = parts;
  UploadBatch ub;
  if (!ub.add(p.d_ent_off, p.h_ent_off) || !ub.add(p.d_work, work) || !ub.add(p.d_chunk_off, p.h_chunk_off) ||
      !ub.add(p.d_chunk_cell, chunk_cell) || !ub.add(p.d_seg_cell, seg_cell) || !ub.add(p.d_seg_chunk0, seg_chunk0) ||
      !ub.add(p.d_seg_nch, seg_nch) || !ub.add(p.d_seg_part, seg_part) || !ub.add(p.d_mcell, mcell) ||
      !ub.add(p.d_mcell_part, mcell_part) || !ub.flush(st))
    return false;
  const size_t nq = (size_t)std::max<long>(p.n_chunks, 1) * kSlot;
  if (!p.d_part_T.ensure(8 * nq) || !p.d_part_Tphi.ensure(8 * nq) || !p.d_part_max.ensure(4 * nq)) return false;
  if (!p.d_Lc.ensure(sizeof(float) * (size_t)N * K * kSlot) ||
      !p.d_panels.ensure(2 * (size_t)std::max<long>(p.n_chunks, 1) * 2 * kSlot * 32) ||
      !p.d_plogp.ensure(sizeof(double) * N) || !p.d_crowscale.ensure(sizeof(float) * p.R))
    return false;
  return !failed() && build_groups(p);
}

// Pattern groups and their dense blocks; the per-cell E-step's entries are
// those left outside the blocks.  With no group large enough, no blocks.
static bool build_groups(Proposal& p) {
  Engine& e = engine();
  hipStream_t st = e.stream;
  const int N = p.N, K = p.topk, C = p.C;
  p.n_gwork = 0;
  p.n_rwork = 0;
  p.st_group_pairs = 0;
  p.st_core_trunc = 0;
  p.st_resid_pairs = p.st_cand_pairs;
  if (p.group_min <= 0) return true;
  if (!p.d_gkey.ensure(sizeof(int) * N) || !p.d_gcount.ensure(sizeof(int) * C) || !p.d_gpat.ensure(sizeof(int) * N) ||
      !p.d_ghist.ensure(sizeof(int) * (size_t)C * C) || !p.d_gcore.ensure(sizeof(int) * (size_t)C * kMaxCore) ||
      !p.d_gncore.ensure(sizeof(int) * C) || !p.d_gnrot.ensure(sizeof(int) * C) || !p.d_gtrunc.ensure(sizeof(int)))
    return false;
  if (!build_group_keys(p.d_cand.as<int>(), N, K, C, p.d_gkey.as<int>(), p.d_gcount.as<int>(), p.d_gpat.as<int>(),
                        p.d_gkey_s, p.d_sort_tmp, st))
    return false;
  launch_group_core(p.d_cand.as<int>(), p.d_gkey.as<int>(), p.d_gcount.as<int>(), N, K, C, p.group_min, p.core_div,
                    p.d_cell_off.as<int>(), p.d_ghist.as<int>(), p.d_gcore.as<int>(), p.d_gncore.as<int>(),
                    p.d_gnrot.as<int>(), p.d_gtrunc.as<int>(), st);
  p.h_gcount.resize(C);
  p.h_gncore.resize(C);
  p.h_gnrot.resize(C);
  EMC_CHECK(hipMemcpyAsync(p.h_gcount.data(), p.d_gcount.p, sizeof(int) * C, hipMemcpyDeviceToHost, st));
  EMC_CHECK(hipMemcpyAsync(p.h_gncore.data(), p.d_gncore.p, sizeof(int) * C, hipMemcpyDeviceToHost, st));
  EMC_CHECK(hipMemcpyAsync(p.h_gnrot.data(), p.d_gnrot.p, sizeof(int) * C, hipMemcpyDeviceToHost, st));
  int* gtrunc = e.host_scalar<int>(0);
  EMC_CHECK(hipMemcpyAsync(gtrunc, p.d_gtrunc.p, sizeof(int), hipMemcpyDeviceToHost, st));
  sync_stream();
  if (failed()) return false;
  p.h_gtrunc = *gtrunc;
  p.st_core_trunc = p.h_gtrunc;
  // block work items: (first pattern in the sorted list, count, first core
  // rotation, count), 128 x 128
  p.h_gcofs.assign(C + 1, 0);
  p.h_gwork.clear();
  long pat0 = 0;
  for (int g = 0; g < C; g++) {
    p.h_gcofs[g + 1] = p.h_gcofs[g] + (p.h_gncore[g] > 0 ? p.h_gnrot[g] : 0);
    if (p.h_gncore[g] > 0)
      for (int a = 0; a < p.h_gcount[g]; a += kGroupTile)
        for (int b = 0; b < p.h_gnrot[g]; b += kGroupTile) {
          p.h_gwork.push_back(make_int4((int)(pat0 + a), std::min(kGroupTile, p.h_gcount[g] - a),
                                        p.h_gcofs[g] + b, std::min(kGroupTile, p.h_gnrot[g] - b)));
          p.st_group_pairs += (double)std::min(kGroupTile, p.h_gcount[g] - a) * std::min(kGroupTile, p.h_gnrot[g] - b);
        }
    pat0 += p.h_gcount[g];
  }
  p.n_gwork = (int)p.h_gwork.size();
  if (p.n_gwork == 0) return true;
  const int ncr = p.h_gcofs[C];
  const long n = (long)N * K;
  UploadBatch ub;
  if (!ub.add(p.d_gcofs, p.h_gcofs) || !ub.add(p.d_gwork, p.h_gwork) || !ub.flush(st) ||
      !p.d_crot.ensure(sizeof(int) * ncr) ||
      !p.d_ccell.ensure(ncr) || !p.d_cslot.ensure(ncr) || !p.d_rank.ensure((size_t)N * kMaxCore) ||
      !p.d_rkey.ensure(sizeof(int) * n) || !p.d_rcount.ensure(sizeof(int) * (C + 1)) ||
      !p.d_rncand.ensure(sizeof(int) * N) || !p.d_rent_pat.ensure(sizeof(int) * n) ||
      !p.d_rent_j.ensure(sizeof(int) * n))
    return false;
  launch_group_rots(p.d_gcore.as<int>(), p.d_gncore.as<int>(), p.d_gcofs.as<int>(), p.d_cell_off.as<int>(),
                    p.d_cell_rot.as<int>(), C, p.d_crot.as<int>(), p.d_ccell.as<unsigned char>(),
                    p.d_cslot.as<unsigned char>(), st);
  launch_group_rank(p.d_cand.as<int>(), p.d_gkey.as<int>(), N, K, C, p.d_gcore.as<int>(), p.d_gncore.as<int>(),
                    p.d_rank.as<signed char>(), p.d_rkey.as<int>(), st);
  // the remaining entries, cell-major (key C = inside a block: sorted last)
  launch_cand_hist(p.d_rkey.as<int>(), N, K, C, p.d_cell_off.as<int>(), p.d_rcount.as<int>(), p.d_rncand.as<int>(), st);
  if (!sort_entries(p.d_rkey.as<int>(), N, K, C + 1, p.d_rkeys_s, p.d_rvals_s, p.d_sort_tmp, p.d_rent_pat.as<int>(),
                    p.d_rent_j.as<int>(), st))
    return false;
  launch_counts_sorted(p.d_rkeys_s.as<int>(), n, C, p.d_rcount.as<int>(), st);
  p.h_rcount.resize(C + 1);
  HostVec<int>& rn = p.h_rn;
  rn.resize(N);
  EMC_CHECK(hipMemcpyAsync(p.h_rcount.data(), p.d_rcount.p, sizeof(int) * (C + 1), hipMemcpyDeviceToHost, st));
  EMC_CHECK(hipMemcpyAsync(rn.data(), p.d_rncand.p, sizeof(int) * N, hipMemcpyDeviceToHost, st));
  sync_stream();
  if (failed()) return false;
  p.st_resid_pairs = 0;
  for (int i = 0; i < N; i++) p.st_resid_pairs += rn[i];
  p.h_rent_off.assign(C + 1, 0);
  for (int c = 0; c < C; c++) p.h_rent_off[c + 1] = p.h_rent_off[c] + p.h_rcount[c];
  std::vector<int2> rwork;
  for (int c = 0; c < C; c++)
    for (int e0 = p.h_rent_off[c]; e0 < p.h_rent_off[c + 1]; e0 += kCandRows) rwork.push_back(make_int2(c, e0));
  p.n_rwork = (int)rwork.size();
  if (!ub.add(p.d_rent_off, p.h_rent_off) || !ub.add(p.d_rwork, rwork) || !ub.flush(st)) return false;
  return !failed();
}

static void reset_modes(Proposal& p) {
  p.h_mode.assign(p.N, 0);
  p.st_gated = 0;
  if (p.gate_predicted)
    for (int i = 0; i < p.N; i++)
      if (p.h_pmass[i] < p.min_mass) {
        p.h_mode[i] = 1;
        p.st_gated++;
      }
}

bool proposal_applies(int N, int R) {
  const Proposal& p = proposal();
  return p.active && p.predicted && p.N == N && p.R == R && engine().precision == PREC_FP16;
}

static bool ensure_iota(Proposal& p, int n) {
  if (p.d_iota.bytes >= sizeof(int) * (size_t)n) return true;
  std::vector<int> iota(n);
  for (int i = 0; i < n; i++) iota[i] = i;
  return upload(p.d_iota, iota, engine().stream);
}

// Re-centre every pattern's candidates on its last posterior: the kTop cells
// holding the most mass (from its candidates and the shared cells, or from
// the whole table for the patterns scored there), then their nearest cells
// in mass order, then the network's cells in rank order, up to topk distinct
// cells.  Before that, the shared cells of the next iteration: every cell
// that is some pattern's best cell or holds >= kShareMass of its posterior,
// if there are at most max_shared of them and they leave topk cells free;
// candidate cells exclude them.
static bool recentre(Proposal& p) {
  Engine& e = engine();
  hipStream_t st = e.stream;
  const int N = p.N, K = Proposal::kTop, ng = (p.C + kTopGroup - 1) / kTopGroup;
  if (!p.d_top.ensure(sizeof(int) * (size_t)N * K) || !p.d_topm.ensure(sizeof(float) * (size_t)N * K)) return false;
  if (p.post == Proposal::POST_CAND) {
    const size_t np = (size_t)std::max(p.n_sub, 1) * ng * 4;
    if (!p.d_part_c.ensure(sizeof(int) * np) || !p.d_part_m.ensure(sizeof(float) * np)) return false;
    launch_topcells(p.d_Lc.as<float>(), N, p.topk, p.d_insub.as<unsigned char>(), p.d_cand.as<int>(),
                    p.d_Lsub.as<float>(), p.n_sub, p.d_sub.as<int>(), p.C, p.d_cell_off.as<int>(),
                    p.d_cell_rot.as<int>(), K, p.d_top.as<int>(), p.d_topm.as<float>(), p.d_part_c.as<int>(),
                    p.d_part_m.as<float>(), st);
    if (p.R_sh > 0) {
      // the shared cells of the last iteration: their top cells from the
      // shared responsibilities, merged with the candidates' top cells
      const int gs = (p.n_shc + kTopGroup - 1) / kTopGroup;
      const size_t ns = (size_t)N * gs * 4;
      if (!ensure_iota(p, std::max(N, p.R_sh)) || !p.d_part_c.ensure(std::max(p.d_part_c.bytes, sizeof(int) * ns)) ||
          !p.d_part_m.ensure(std::max(p.d_part_m.bytes, sizeof(float) * ns)) ||
          !p.d_top_s.ensure(sizeof(int) * (size_t)N * K) || !p.d_topm_s.ensure(sizeof(float) * (size_t)N * K))
        return false;
      launch_topcells(nullptr, N, p.topk, nullptr, nullptr, p.d_Lsh.as<float>(), N, p.d_iota.as<int>(), p.n_shc,
                      p.d_sh_off.as<int>(), p.d_iota.as<int>(), K, p.d_top_s.as<int>(), p.d_topm_s.as<float>(),
                      p.d_part_c.as<int>(), p.d_part_m.as<float>(), st);
      launch_merge_top(N, K, p.d_insub.as<unsigned char>(), p.d_top.as<int>(), p.d_topm.as<float>(),
                       p.d_top_s.as<int>(), p.d_topm_s.as<float>(), p.d_sh_cells.as<int>(), st);
    }
  } else if (p.post != Proposal::POST_TOP || p.post_N != N) {
    p.clear_shared();
    return true;
  }
  // shared cells of the next iteration, from the support marks
  p.clear_shared();
  const size_t cb = ((size_t)p.C + 3) / 4 * 4;
  if (!p.d_mark.ensure(cb) || !p.d_shared.ensure(cb) || !p.d_moved.ensure(sizeof(int))) return false;
  p.h_shared.assign(cb, 0);
  UploadBatch ub;
  if (p.max_shared > 0) {
    launch_mark_support(p.d_top.as<int>(), p.d_topm.as<float>(), N, K, Proposal::kShareMass, p.C,
                        p.d_mark.as<unsigned char>(), st);
    p.h_mark.resize(cb);
    EMC_CHECK(hipMemcpyAsync(p.h_mark.data(), p.d_mark.p, cb, hipMemcpyDeviceToHost, st));
    sync_stream();
    if (failed()) return false;
    int n = 0;
    for (int c = 0; c < p.C; c++) n += p.h_mark[c] != 0;
    if (n <= p.max_shared && n <= p.C - p.topk) {
      for (int c = 0; c < p.C; c++)
        if (p.h_mark[c]) {
          p.h_shared[c] = 1;
          p.h_sh_cells.push_back(c);
          for (int j = p.h_cell_off[c]; j < p.h_cell_off[c + 1]; j++) p.h_sh_rot.push_back(p.h_cell_rot[j]);
          p.h_sh_off.push_back((int)p.h_sh_rot.size());
        }
      p.n_shc = (int)p.h_sh_cells.size();
      p.R_sh = (int)p.h_sh_rot.size();
      p.R_sh_pad = (int)round_up(std::max(p.R_sh, 1), kPadR);
      if (!ub.add(p.d_sh_cells, p.h_sh_cells) || !ub.add(p.d_sh_off, p.h_sh_off) || !ub.add(p.d_sh_rot, p.h_sh_rot))
        return false;
    }
  }
  if (!ub.add(p.d_shared, p.h_shared) || !ub.flush(st)) return false;
  // every pattern's candidates, re-centred (GPU), then the lists if any changed
  launch_fill(p.d_moved.as<float>(), 1, 0.0f, st);
  launch_recentre_cand(p.d_top.as<int>(), p.d_topm.as<float>(), N, K, p.d_nbr.as<int>(), Proposal::kNbr,
                       p.d_net_cand.as<int>(), p.d_shared.as<unsigned char>(), p.C, p.topk, p.d_cand.as<int>(),
                       p.d_moved.as<int>(), st);
  // the lists are rebuilt whether or not a pattern moved (almost every
  // iteration some do), so the count is read after the list building's own
  // synchronisation instead of with one of its own
  static int* moved = nullptr;
  if (!moved && hipHostMalloc((void**)&moved, sizeof(int), hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    moved = nullptr;
  }
  int moved_now = 0;
  EMC_CHECK(hipMemcpyAsync(moved ? moved : &moved_now, p.d_moved.p, sizeof(int), hipMemcpyDeviceToHost, st));
  if (!moved) sync_stream();
  if (failed() || !build_lists(p)) return false;  // it synchronises after the copy above
  p.st_moved = moved ? *moved : moved_now;
  return !failed();
}

bool proposal_estep(int N, int R, float sigma, const float* phi_used, bool* dense) {
  Engine& e = engine();
  Proposal& p = proposal();
  hipStream_t st = e.stream;
  *dense = false;
  const auto t0 = std::chrono::steady_clock::now();
  // (re-centring builds lists of at most kMaxTopk cells per pattern; longer
  // caller-given lists stay as given)
  if (p.track && p.post != Proposal::POST_NONE && p.topk <= kMaxTopk) {
    if (!recentre(p)) return false;
  } else {
    p.clear_shared();
  }
  p.st_recentre_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  p.post = Proposal::POST_NONE;
  p.st_shared = p.n_shc;
  p.iteration++;
  p.have_p = false;
  // whole-table subset: the audit sample and the patterns in full search
  p.h_sub.clear();
  p.h_insub.assign(N, 0);
  p.h_audit.assign(N, 0);
  p.st_full = 0;
  p.st_audited = 0;
  p.st_dense = 0;
  for (int i = 0; i < N; i++) {
    const bool aud = emc_uniform(p.seed, 0xA0D17ULL, (uint64_t)p.iteration, (uint64_t)i) < (double)p.audit_frac;
    const bool full = p.h_mode[i] != 0;
    if (aud || full) {
      p.h_insub[i] = 1;
      p.h_audit[i] = aud && !full;
      p.h_sub.push_back(i);
      p.st_full += full;
      p.st_audited += aud && !full;
    }
  }
  p.n_sub = (int)p.h_sub.size();
  p.n_sub_pad = (int)round_up(std::max(p.n_sub, 1), kPadN);
  UploadBatch ub;
  if (!ub.add(p.d_sub, p.h_sub) || !ub.add(p.d_insub, p.h_insub) || !ub.flush(st)) return false;
  if (!p.d_zero.ensure(2 * (size_t)e.n_mp) || !e.d_phiden.ensure(sizeof(double) * N)) return false;
  launch_fill(p.d_zero.as<float>(), (size_t)e.n_mp / 2, 0.0f, st);
  const bool split = !e.k_exact;
  const bool poisson = e.likelihood == LIKE_POISSON;
  const double inv_2nms2 = 1.0 / (2.0 * (double)(float)e.n_m * ((double)sigma * (double)sigma));
  const double sigma2x = 2.0 * (double)(float)e.n_m * ((double)sigma * (double)sigma);
  // the subset first: its posterior measures the candidates' captured mass
  // before the candidates are scored
  p.sub_total = 0.0;
  p.st_fail = 0;
  p.st_min_mass = p.st_mean_mass = 1.0;
  if (p.n_sub > 0) {
    const int ns = p.n_sub;
    const size_t rows = (size_t)p.n_sub_pad * e.n_mp;
    const int chunks = softmax_chunks(ns, R);
    const size_t cn = (size_t)chunks * ns;
    const int nb = (ns + 255) / 256;
    if (!p.d_Ksub.ensure(2 * rows) || (split && !p.d_Ksublo.ensure(2 * rows)) ||
        !p.d_Bsub.ensure(sizeof(double) * ns) || !p.d_phisub.ensure(sizeof(float) * ns) ||
        !p.d_Lsub.ensure(sizeof(float) * (size_t)R * ns) || !p.d_mass.ensure(sizeof(float) * ns) ||
        !e.d_smax.ensure(std::max(e.d_smax.bytes, cn * 4)) || !e.d_ssum.ensure(std::max(e.d_ssum.bytes, cn * 8)) ||
        !e.d_sacc.ensure(std::max(e.d_sacc.bytes, 4 * cn * 8)) || !p.d_phiden_sub.ensure(sizeof(double) * ns) ||
        !e.d_red.ensure(std::max(e.d_red.bytes, sizeof(double) * (nb + 1))))
      return false;
    {
      PhaseTimer gt(PH_AUDIT_GEMM);
      launch_gather_rows(e.d_K16.as<_Float16>(), split ? e.d_K16lo.as<_Float16>() : nullptr, p.d_sub.as<int>(), ns,
                         p.n_sub_pad, e.n_mp, p.d_Ksub.as<_Float16>(), split ? p.d_Ksublo.as<_Float16>() : nullptr,
                         st);
      launch_gather_f64(e.d_B.as<double>(), p.d_sub.as<int>(), ns, p.d_Bsub.as<double>(), st);
      launch_gather_f32(phi_used, p.d_sub.as<int>(), ns, p.d_phisub.as<float>(), st);
      LikeEpilogue ep{};
      ep.L = p.d_Lsub.as<float>();
      ep.A = e.d_A.as<double>();
      ep.B = p.d_Bsub.as<double>();
      ep.phi = p.d_phisub.as<float>();
      ep.cscale = e.d_Sscale.as<float>();
      ep.kscale_inv = 1.0f / e.k_scale;
      ep.inv_2nms2 = inv_2nms2;
      ep.N = ns;
      ep.R = R;
      ep.poisson = poisson;
      ep.vecL = ns % 4 == 0;
      launch_estep16(split, p.d_Ksub.p, p.d_Ksublo.p, e.d_S16.p, e.d_S16lo.p, p.n_sub_pad, e.R_pad, e.n_mp, ep, st);
    }
    PhaseTimer pt(PH_AUDIT);
    launch_softmax(p.d_Lsub.as<float>(), R, ns, chunks, e.d_smax.as<float>(), e.d_ssum.as<double>(),
                   e.d_sacc.as<double>(), e.d_A.as<double>(), true, st);
    launch_softmax_reduce(ns, chunks, e.d_sacc.as<double>(), true, p.d_Bsub.as<double>(), p.d_phisub.as<float>(),
                          sigma2x, p.d_phiden_sub.as<double>(), e.d_red.as<double>(), st, poisson);
    launch_sub_finish(p.d_Lsub.as<float>(), ns, p.d_sub.as<int>(), p.d_phiden_sub.as<double>(),
                      e.d_phiden.as<double>(), p.d_cand.as<int>(), p.topk, p.d_cell_off.as<int>(),
                      p.d_cell_rot.as<int>(), p.d_sh_rot.as<int>(), p.R_sh, p.d_mass.as<float>(), st);
    HostVec<float>& mass = p.h_mass;
    mass.resize(ns);
    double* sub_total = e.host_scalar<double>(1);
    EMC_CHECK(hipMemcpyAsync(sub_total, e.d_red.as<double>() + nb, sizeof(double), hipMemcpyDeviceToHost, st));
    EMC_CHECK(hipMemcpyAsync(mass.data(), p.d_mass.p, sizeof(float) * ns, hipMemcpyDeviceToHost, st));
    sync_stream();
    if (failed()) return false;
    p.sub_total = *sub_total;
    // search mode of every pattern scored on the whole table: full search
    // while its candidates hold less than min_mass of its posterior
    double fails = 0, mn = 1.0, sum = 0.0;
    for (int k = 0; k < ns; k++) {
      const int i = p.h_sub[k];
      const bool low = mass[k] < p.min_mass;
      if (p.h_audit[i]) {
        fails += low;
        mn = std::min(mn, (double)mass[k]);
        sum += mass[k];
      }
      p.h_mode[i] = low ? 1 : 0;
    }
    p.st_fail = fails;
    p.st_min_mass = p.st_audited > 0 ? mn : 1.0;
    p.st_mean_mass = p.st_audited > 0 ? sum / p.st_audited : 1.0;
  }
  if (p.st_audited > 0 && p.st_fail > (double)p.max_fail * p.st_audited) {
    // the audit says the candidates miss too much of the posterior: this
    // iteration runs on the whole table
    p.st_dense = 1;
    p.st_pairs = (double)N * R;
    p.st_cand_pairs = 0;
    *dense = true;
    return true;
  }
  {
    CandEstepArgs a{};
    a.K = e.d_K16.as<_Float16>();
    a.Klo = split ? e.d_K16lo.as<_Float16>() : nullptr;
    a.S = e.d_S16.as<_Float16>();
    a.Slo = e.d_S16lo.as<_Float16>();
    a.zero = p.d_zero.as<_Float16>();
    // with group blocks: the blocks, then the entries outside them
    const bool groups = p.n_gwork > 0;
    a.work = groups ? p.d_rwork.as<int2>() : p.d_work.as<int2>();
    a.n_work = groups ? p.n_rwork : p.n_work;
    a.cell_off = p.d_cell_off.as<int>();
    a.cell_rot = p.d_cell_rot.as<int>();
    a.ent_off = groups ? p.d_rent_off.as<int>() : p.d_ent_off.as<int>();
    a.ent_pat = groups ? p.d_rent_pat.as<int>() : p.d_ent_pat.as<int>();
    a.ent_j = groups ? p.d_rent_j.as<int>() : p.d_ent_j.as<int>();
    a.A = e.d_A.as<double>();
    a.B = e.d_B.as<double>();
    a.phi = phi_used;
    a.sscale = e.d_Sscale.as<float>();
    a.kscale_inv = 1.0f / e.k_scale;
    a.inv_2nms2 = inv_2nms2;
    a.n_mp = e.n_mp;
    a.topk = p.topk;
    a.poisson = poisson;
    a.Lc = p.d_Lc.as<float>();
    if (groups) {
      GroupEstepArgs ga{};
      ga.K = a.K;
      ga.Klo = a.Klo;
      ga.S = a.S;
      ga.Slo = a.Slo;
      ga.zero = a.zero;
      ga.work = p.d_gwork.as<int4>();
      ga.n_work = p.n_gwork;
      ga.gpat = p.d_gpat.as<int>();
      ga.crot = p.d_crot.as<int>();
      ga.ccell = p.d_ccell.as<unsigned char>();
      ga.cslot = p.d_cslot.as<unsigned char>();
      ga.rank = p.d_rank.as<signed char>();
      ga.A = a.A;
      ga.B = a.B;
      ga.phi = a.phi;
      ga.sscale = a.sscale;
      ga.kscale_inv = a.kscale_inv;
      ga.inv_2nms2 = a.inv_2nms2;
      ga.n_mp = a.n_mp;
      ga.topk = a.topk;
      ga.poisson = a.poisson;
      ga.Lc = a.Lc;
      PhaseTimer gt(PH_GROUPS);
      launch_group_estep(split, ga, st);
    }
    PhaseTimer pt(PH_LIKE);
    launch_cand_estep(split, a, st);
  }
  if (p.R_sh > 0) {
    {
      // every pattern on the shared rotations: gathered slice rows, dense kernel
      PhaseTimer pt(PH_SHARED);
      const size_t rows = (size_t)p.R_sh_pad * e.n_mp;
      if (!p.d_Ssh.ensure(2 * rows) || !p.d_Sshlo.ensure(2 * rows) || !p.d_Ash.ensure(sizeof(double) * p.R_sh) ||
          !p.d_sscale_sh.ensure(sizeof(float) * p.R_sh) || !p.d_Lsh.ensure(sizeof(float) * (size_t)p.R_sh * N))
        return false;
      launch_gather_rows(e.d_S16.as<_Float16>(), e.d_S16lo.as<_Float16>(), p.d_sh_rot.as<int>(), p.R_sh, p.R_sh_pad,
                         e.n_mp, p.d_Ssh.as<_Float16>(), p.d_Sshlo.as<_Float16>(), st);
      launch_gather_f64(e.d_A.as<double>(), p.d_sh_rot.as<int>(), p.R_sh, p.d_Ash.as<double>(), st);
      launch_gather_f32(e.d_Sscale.as<float>(), p.d_sh_rot.as<int>(), p.R_sh, p.d_sscale_sh.as<float>(), st);
      LikeEpilogue ep{};
      ep.L = p.d_Lsh.as<float>();
      ep.A = p.d_Ash.as<double>();
      ep.B = e.d_B.as<double>();
      ep.phi = phi_used;
      ep.cscale = p.d_sscale_sh.as<float>();
      ep.kscale_inv = 1.0f / e.k_scale;
      ep.inv_2nms2 = inv_2nms2;
      ep.N = N;
      ep.R = p.R_sh;
      ep.poisson = poisson;
      ep.vecL = N % 4 == 0;
      launch_estep16(split, e.d_K16.p, e.d_K16lo.p, p.d_Ssh.p, p.d_Sshlo.p, e.N_pad, p.R_sh_pad, e.n_mp, ep, st);
    }
  }
  return !failed();
}

bool proposal_softmax(float* respons, int N, int R, double sigma2x, double* total) {
  Engine& e = engine();
  Proposal& p = proposal();
  hipStream_t st = e.stream;
  if (!p.d_tot.ensure(sizeof(double) * 2)) return false;
  // The matrix still holds the last proposal iteration's posterior, whose
  // nonzeros the M-step listed: zero those entries and write the subset's
  // nonzeros, instead of rewriting all N * R entries.
  auto& rec = e.resp_nz;
  const bool delta = e.respons_delta && rec.valid && rec.ptr == respons && rec.N == N && rec.R == R &&
                     rec.stamp == e.invalidate_stamp && e.d_nz_keys_s.bytes >= sizeof(unsigned) * (size_t)rec.nnz;
  rec.valid = false;  // the keys describe the contents until this call writes them
  const bool vec_fill = !delta && N % 4 == 0 && (uintptr_t)respons % 16 == 0 && p.n_sub > 0;
  if (vec_fill) {
    p.h_colmap.assign(N, -1);
    for (int k = 0; k < p.n_sub; k++) p.h_colmap[p.h_sub[k]] = k;
    if (!upload(p.d_colmap, p.h_colmap, st)) return false;
  }
  const int chunks = p.R_sh > 0 ? softmax_chunks(N, p.R_sh) : 0;
  const int nmerge = chunks + 1, nb = (N + 255) / 256;
  if (p.R_sh > 0) {
    const size_t cn = (size_t)nmerge * N;
    if (!e.d_smax.ensure(std::max(e.d_smax.bytes, cn * 4)) || !e.d_ssum.ensure(std::max(e.d_ssum.bytes, cn * 8)) ||
        !e.d_sacc.ensure(std::max(e.d_sacc.bytes, 4 * cn * 8)) ||
        !e.d_red.ensure(std::max(e.d_red.bytes, sizeof(double) * (nb + 1))))
      return false;
  }
  {
    PhaseTimer pt(PH_SOFTMAX);
    // zero, with the whole-table subset's columns written in the same pass
    // when the rows allow 16-byte stores
    p.st_cleared = delta ? (double)rec.nnz : (double)N * R;
    if (delta) {
      launch_zero_keys(e.d_nz_keys_s.as<unsigned>(), rec.nnz, respons, st);
      launch_sub_scatter(p.d_Lsub.as<float>(), R, p.n_sub, p.d_sub.as<int>(), N, respons, true, st);
    } else if (vec_fill) {
      launch_fill_subset(respons, R, N, p.d_colmap.as<int>(), p.d_Lsub.as<float>(), p.n_sub, st);
    } else {
      launch_fill(respons, (size_t)N * R, 0.0f, st);
    }
    CandSoftmaxArgs a{};
    a.Lc = p.d_Lc.as<float>();
    a.cand = p.d_cand.as<int>();
    a.cell_off = p.d_cell_off.as<int>();
    a.cell_rot = p.d_cell_rot.as<int>();
    a.insub = p.d_insub.as<unsigned char>();
    a.A = e.d_A.as<double>();
    a.B = e.d_B.as<double>();
    a.phi = e.d_phi_used.as<float>();
    a.sigma2x = sigma2x;
    a.N = N;
    a.topk = p.topk;
    a.poisson = e.likelihood == LIKE_POISSON;
    a.respons = respons;
    a.phiden = e.d_phiden.as<double>();
    a.plogp = p.d_plogp.as<double>();
    if (p.R_sh == 0) {
      a.mode = 0;
      launch_cand_softmax(a, st);
      launch_sum_f64(p.d_plogp.as<double>(), N, p.d_tot.as<double>(), st);
    } else {
      // candidates and shared rotations normalised together: the shared
      // rows' chunk partials plus the candidates' as partial `chunks`
      a.smax = e.d_smax.as<float>();
      a.ssum = e.d_ssum.as<double>();
      a.sacc = e.d_sacc.as<double>();
      a.nmerge = nmerge;
      a.slot = chunks;
      launch_softmax_partial(p.d_Lsh.as<float>(), p.R_sh, N, chunks, a.smax, a.ssum, st);
      a.mode = 1;
      launch_cand_softmax(a, st);
      launch_softmax_finish(p.d_Lsh.as<float>(), p.R_sh, N, chunks, nmerge, a.smax, a.ssum, a.sacc,
                            p.d_Ash.as<double>(), true, st);
      a.mode = 2;
      launch_cand_softmax(a, st);
      launch_softmax_reduce(N, nmerge, a.sacc, true, e.d_B.as<double>(), e.d_phi_used.as<float>(), sigma2x,
                            e.d_phiden.as<double>(), e.d_red.as<double>(), st, a.poisson != 0,
                            p.d_insub.as<unsigned char>());
      launch_shared_scatter(p.d_Lsh.as<float>(), p.R_sh, N, p.d_sh_rot.as<int>(), p.d_insub.as<unsigned char>(),
                            respons, st);
      EMC_CHECK(hipMemcpyAsync(p.d_tot.p, e.d_red.as<double>() + nb, sizeof(double), hipMemcpyDeviceToDevice, st));
    }
  }
  if (p.n_sub > 0 && !vec_fill && !delta) {
    // the subset's posterior (computed with the E-step) into its columns
    PhaseTimer pt(PH_AUDIT);
    launch_sub_scatter(p.d_Lsub.as<float>(), R, p.n_sub, p.d_sub.as<int>(), N, respons, false, st);
  }
  double* h_tot = e.host_scalar<double>(2);
  EMC_CHECK(hipMemcpyAsync(h_tot, p.d_tot.p, sizeof(double), hipMemcpyDeviceToHost, st));
  sync_stream();
  if (failed()) return false;
  const double tot = *h_tot;
  double dtot = 0.0;
  e.ref_entries = 0;
  if (e.refine && e.like.slices && e.like.images) {
    // the significant responsibilities from FP64 likelihoods, in the matrix
    // and in the candidate, shared and subset buffers the M-step reads
    PhaseTimer pt(PH_REFINE);
    RefineInputs in;
    in.respons = respons;
    in.N = N;
    in.R = R;
    in.N2d = e.like.N2d;
    in.slices = e.like.slices;
    in.images = e.like.images;
    if (p.R_sh > 0) {
      RefineRows& s = in.rows[in.nrows++];
      s.P = p.d_Lsh.as<float>();
      s.rows = p.R_sh;
      s.ld = N;
      s.cols = N;
      s.row_rot = p.d_sh_rot.as<int>();
      s.skip = p.d_insub.as<unsigned char>();
      s.own = true;
    }
    if (p.n_sub > 0) {
      RefineRows& s = in.rows[in.nrows++];
      s.P = p.d_Lsub.as<float>();
      s.rows = R;
      s.ld = p.n_sub;
      s.cols = p.n_sub;
      s.col_pat = p.d_sub.as<int>();
      s.own = true;
    }
    in.Lc = p.d_Lc.as<float>();
    in.cand = p.d_cand.as<int>();
    in.cell_off = p.d_cell_off.as<int>();
    in.cell_rot = p.d_cell_rot.as<int>();
    in.insub = p.d_insub.as<unsigned char>();
    in.topk = p.topk;
    in.slots = kSlot;
    in.phiden = e.d_phiden.as<double>();
    in.poisson = e.likelihood == LIKE_POISSON;
    in.sigma2x = sigma2x;
    in.tau = e.refine_tau;
    if (!refine_posterior(in, &dtot)) return false;
  }
  double pairs = (double)p.n_sub * R + (double)(N - p.n_sub) * p.R_sh;
  for (int i = 0; i < N; i++)
    if (!p.h_insub[i]) pairs += p.h_ncand[i];
  p.st_pairs = pairs;
  *total = tot + p.sub_total + dtot;
  p.have_p = true;
  p.post = Proposal::POST_CAND;
  p.respons = respons;
  return true;
}

bool proposal_has_shared() { return proposal().R_sh > 0; }

bool proposal_nz_sources(NzSources& a, int N, int R) {
  const Proposal& p = proposal();
  if (!p.have_p || p.N != N || p.R != R) {
    set_error("proposal: no candidate posterior for %d patterns x %d rotations", N, R);
    return false;
  }
  a = {};
  a.Pc = p.d_Lc.as<float>();
  a.cand = p.d_cand.as<int>();
  a.cell_off = p.d_cell_off.as<int>();
  a.cell_rot = p.d_cell_rot.as<int>();
  a.insub = p.d_insub.as<unsigned char>();
  a.N = N;
  a.topk = p.topk;
  a.Psh = p.d_Lsh.as<float>();
  a.sh_rot = p.d_sh_rot.as<int>();
  a.R_sh = p.R_sh;
  a.Psub = p.d_Lsub.as<float>();
  a.sub = p.d_sub.as<int>();
  a.R = R;
  a.n_sub = p.n_sub;
  return true;
}

bool proposal_tracks(int N, int R) {
  const Proposal& p = proposal();
  return p.configured && p.track && p.R == R && N > 0 && engine().precision == PREC_FP16;
}

bool proposal_dense_posterior(const float* respons, int N, int R) {
  Engine& e = engine();
  Proposal& p = proposal();
  hipStream_t st = e.stream;
  const int K = Proposal::kTop, ng = (p.C + kTopGroup - 1) / kTopGroup;
  p.post = Proposal::POST_NONE;
  if (!ensure_iota(p, N)) return false;
  const size_t np = (size_t)N * ng * 4;
  if (!p.d_top.ensure(sizeof(int) * (size_t)N * K) || !p.d_topm.ensure(sizeof(float) * (size_t)N * K) ||
      !p.d_part_c.ensure(sizeof(int) * np) || !p.d_part_m.ensure(sizeof(float) * np))
    return false;
  {
    PhaseTimer pt(PH_AUDIT);
    launch_topcells(nullptr, N, p.topk, nullptr, nullptr, respons, N, p.d_iota.as<int>(), p.C,
                    p.d_cell_off.as<int>(), p.d_cell_rot.as<int>(), K, p.d_top.as<int>(), p.d_topm.as<float>(),
                    p.d_part_c.as<int>(), p.d_part_m.as<float>(), st);
    sync_stream();
  }
  if (failed()) return false;
  (void)R;
  p.post = Proposal::POST_TOP;
  p.post_N = N;
  return true;
}

bool proposal_mstep(int N, int R, float* slices, int N2d, const float* scaling) {
  (void)N;
  Engine& e = engine();
  Proposal& p = proposal();
  hipStream_t st = e.stream;
  const bool split = !e.k_exact;
  const bool has_sub = p.n_sub > 0;
  // Gaussian: operand P/phi, divisor T; Poisson: operand P, divisor sum P phi
  const bool poisson = e.likelihood == LIKE_POISSON;
  const float* opscale = poisson ? e.d_ones.as<float>() : e.d_invphi.as<float>();
  if (has_sub) {
    PhaseTimer pt(PH_AUDIT);
    const size_t pk = (size_t)e.R_pad * p.n_sub_pad, kt = (size_t)e.n_mp * p.n_sub_pad;
    if (!p.d_invphisub.ensure(sizeof(float) * p.n_sub) || !p.d_Psub16.ensure(2 * pk) || !p.d_Psub16lo.ensure(2 * pk) ||
        !p.d_Ktsub.ensure(2 * kt) || (split && !p.d_Ktsublo.ensure(2 * kt)))
      return false;
    launch_gather_f32(opscale, p.d_sub.as<int>(), p.n_sub, p.d_invphisub.as<float>(), st);
    if (poisson) launch_gather_f32(scaling, p.d_sub.as<int>(), p.n_sub, p.d_phisub.as<float>(), st);
    launch_rowsum(p.d_Lsub.as<float>(), R, p.n_sub, p.d_invphisub.as<float>(), e.d_T.as<double>(), e.d_Tf.as<float>(),
                  e.d_Pscale.as<float>(), true, st, poisson ? p.d_phisub.as<float>() : nullptr,
                  poisson ? e.d_Tphi.as<float>() : nullptr);
    MstepArgs ms{};
    ms.P = p.d_Lsub.as<float>();
    ms.invphi = p.d_invphisub.as<float>();
    ms.rowscale = e.d_Pscale.as<float>();
    ms.Tf = e.d_Tf.as<float>();
    ms.pix = e.d_pix.as<int>();
    ms.slices = slices;
    ms.Sc = e.d_Sc.as<float>();
    ms.kscale_inv = 1.0f / e.k_scale;
    ms.N = p.n_sub;
    ms.R = R;
    ms.N_pad = p.n_sub_pad;
    ms.R_pad = e.R_pad;
    ms.n_mp = e.n_mp;
    ms.n_m = e.n_m;
    ms.N2d = N2d;
    ms.vecP = p.n_sub % 4 == 0;
    ms.raw = 1;
    launch_convert_p(ms, p.d_Psub16.as<_Float16>(), p.d_Psub16lo.as<_Float16>(), st);
    launch_gather_rows_t(e.d_K16.as<_Float16>(), p.d_sub.as<int>(), p.n_sub, p.n_sub_pad, e.n_mp,
                         p.d_Ktsub.as<_Float16>(), st);
    if (split)
      launch_gather_rows_t(e.d_K16lo.as<_Float16>(), p.d_sub.as<int>(), p.n_sub, p.n_sub_pad, e.n_mp,
                           p.d_Ktsublo.as<_Float16>(), st);
    launch_mstep16(split, p.d_Psub16.p, p.d_Psub16lo.p, p.d_Ktsub.p, p.d_Ktsublo.p, ms, st);
  }
  {
    PhaseTimer pt(PH_ROWSUM);
    CandPrepArgs a{};
    a.Pc = p.d_Lc.as<float>();
    a.ent_off = p.d_ent_off.as<int>();
    a.ent_pat = p.d_ent_pat.as<int>();
    a.ent_j = p.d_ent_j.as<int>();
    a.cell_off = p.d_cell_off.as<int>();
    a.cell_rot = p.d_cell_rot.as<int>();
    a.chunk_off = p.d_chunk_off.as<long>();
    a.invphi = opscale;
    a.phi = poisson ? scaling : nullptr;
    a.Tsub = has_sub ? e.d_T.as<double>() : nullptr;
    a.Tphisub = has_sub && poisson ? e.d_Tphi.as<float>() : nullptr;
    a.T = e.d_T.as<double>();
    a.Tf = e.d_Tf.as<float>();
    a.Tphi = e.d_Tphi.as<float>();
    a.rowscale = p.d_crowscale.as<float>();
    a.panels = p.d_panels.as<_Float16>();
    a.chunk_cell = p.d_chunk_cell.as<int>();
    a.part_T = p.d_part_T.as<double>();
    a.part_Tphi = p.d_part_Tphi.as<double>();
    a.part_max = p.d_part_max.as<float>();
    a.C = p.C;
    a.topk = p.topk;
    launch_cand_prep(a, p.n_chunks, st);
  }
  // partial sums of the split cells: [parts][32 slots][n_mp]
  if (!p.d_part.ensure(sizeof(float) * (size_t)std::max(p.n_parts, 1) * kSlot * e.n_mp)) return false;
  {
    PhaseTimer pt(PH_MSTEP);
    CandMstepArgs a{};
    a.K = e.d_K16.as<_Float16>();
    a.Klo = split ? e.d_K16lo.as<_Float16>() : nullptr;
    a.zero = p.d_zero.as<_Float16>();
    a.panels = p.d_panels.as<_Float16>();
    a.chunk_off = p.d_chunk_off.as<long>();
    a.seg_cell = p.d_seg_cell.as<int>();
    a.seg_chunk0 = p.d_seg_chunk0.as<long>();
    a.seg_nch = p.d_seg_nch.as<int>();
    a.seg_part = p.d_seg_part.as<int>();
    a.n_seg = p.n_seg;
    a.part = p.d_part.as<float>();
    a.mcell = p.d_mcell.as<int>();
    a.mcell_part = p.d_mcell_part.as<int>();
    a.n_mcell = p.n_mcell;
    a.ent_off = p.d_ent_off.as<int>();
    a.ent_pat = p.d_ent_pat.as<int>();
    a.cell_off = p.d_cell_off.as<int>();
    a.cell_rot = p.d_cell_rot.as<int>();
    a.pix = e.d_pix.as<int>();
    a.rowscale = p.d_crowscale.as<float>();
    a.Tf = poisson ? e.d_Tphi.as<float>() : e.d_Tf.as<float>();
    a.poisson = poisson;
    a.Sc = e.d_Sc.as<float>();
    a.slices = slices;
    a.kscale_inv = 1.0f / e.k_scale;
    a.has_sub = has_sub;
    a.C = p.C;
    a.n_ptiles = e.n_mp / 256;
    a.n_mp = e.n_mp;
    a.n_m = e.n_m;
    a.N2d = N2d;
    launch_cand_mstep(split, a, st);
  }
  p.have_p = false;
  return !failed();
}

}  // namespace emc

using namespace emc;

extern "C" {

// ---------------------------------------------------------------------------
// network C-ABI (tests and tools)
// ---------------------------------------------------------------------------
int emc_hip_cnn_init(int side, int ncell, int batch, unsigned long long seed) {
  if (!ready()) return 1;
  return proposal().cnn.init(side, ncell, batch, seed) ? 0 : 1;
}

long emc_hip_cnn_param_count(void) { return proposal().cnn.nparam; }

int emc_hip_cnn_get(float* params_host, float* grads_host) {
  if (!ready()) return 1;
  Cnn& n = proposal().cnn;
  if (!n.nparam) return 1;
  if (params_host) EMC_CHECK(hipMemcpy(params_host, n.params.p, n.nparam * 4, hipMemcpyDeviceToHost));
  if (grads_host) EMC_CHECK(hipMemcpy(grads_host, n.grads.p, n.nparam * 4, hipMemcpyDeviceToHost));
  return failed() ? 1 : 0;
}

int emc_hip_cnn_set(const float* params_host) {
  if (!ready()) return 1;
  Cnn& n = proposal().cnn;
  if (!n.nparam || !params_host) return 1;
  EMC_CHECK(hipMemcpy(n.params.p, params_host, n.nparam * 4, hipMemcpyHostToDevice));
  return failed() ? 1 : 0;
}

int emc_hip_cnn_forward(const real* d_images, const int* d_mask, int N2d, const int* idx_host, int B,
                        float* logits_host) {
  if (!ready()) return 1;
  Cnn& n = proposal().cnn;
  Engine& e = engine();
  if (!n.nparam || B <= 0 || B > n.batch || N2d != n.side * n.side) {
    set_error("emc_hip_cnn_forward: invalid argument");
    return 1;
  }
  EMC_CHECK(hipMemcpyAsync(n.idx.p, idx_host, sizeof(int) * B, hipMemcpyHostToDevice, e.stream));
  n.forward(d_images, d_mask, N2d, n.idx.as<int>(), B, e.stream);
  if (logits_host)
    EMC_CHECK(hipMemcpyAsync(logits_host, n.logits.p, sizeof(float) * (size_t)n.ncell * B, hipMemcpyDeviceToHost,
                             e.stream));
  EMC_CHECK(hipStreamSynchronize(e.stream));
  return failed() ? 1 : 0;
}

// One training step on B patterns: targets_host is [C][B]; lr = 0 leaves the
// weights unchanged (gradients are still computed and readable).
int emc_hip_cnn_step(const real* d_images, const int* d_mask, int N2d, const int* idx_host, int B,
                     const float* targets_host, float lr, double* loss) {
  if (!ready()) return 1;
  Cnn& n = proposal().cnn;
  Engine& e = engine();
  if (!n.nparam || B <= 0 || B > n.batch || N2d != n.side * n.side || !targets_host) {
    set_error("emc_hip_cnn_step: invalid argument");
    return 1;
  }
  DBuf t, ti;
  if (!t.ensure(sizeof(float) * (size_t)n.ncell * B) || !ti.ensure(sizeof(int) * B)) return 1;
  std::vector<int> iota(B);
  std::iota(iota.begin(), iota.end(), 0);
  // the caller's [n_cells][B] targets, pattern-major for the loss kernel
  std::vector<float> tt((size_t)n.ncell * B);
  for (int c = 0; c < n.ncell; c++)
    for (int b = 0; b < B; b++) tt[(size_t)b * n.ncell + c] = targets_host[(size_t)c * B + b];
  EMC_CHECK(hipMemcpyAsync(t.p, tt.data(), sizeof(float) * tt.size(), hipMemcpyHostToDevice, e.stream));
  EMC_CHECK(hipStreamSynchronize(e.stream));
  EMC_CHECK(hipMemcpyAsync(ti.p, iota.data(), sizeof(int) * B, hipMemcpyHostToDevice, e.stream));
  EMC_CHECK(hipMemcpyAsync(n.idx.p, idx_host, sizeof(int) * B, hipMemcpyHostToDevice, e.stream));
  n.forward(d_images, d_mask, N2d, n.idx.as<int>(), B, e.stream);
  launch_fill(n.loss.as<float>(), 2, 0.0f, e.stream);
  softmax_xent(n.logits.as<float>(), n.ncell, B, t.as<float>(), B, ti.as<int>(), 1.0f / B, n.dlogits.as<float>(),
               n.loss.as<double>(), e.stream);
  n.backward_and_step(B, lr, e.stream);
  double l = 0;
  EMC_CHECK(hipMemcpyAsync(&l, n.loss.p, 8, hipMemcpyDeviceToHost, e.stream));
  EMC_CHECK(hipStreamSynchronize(e.stream));
  if (loss) *loss = l / B;
  t.release();
  ti.release();
  return failed() ? 1 : 0;
}


// ---------------------------------------------------------------------------
// proposal workflow
// ---------------------------------------------------------------------------
int emc_hip_proposal_init(const real* q_fine, int n_rot, int n_coarse, int side, int topk, int batch,
                          unsigned long long seed) {
  if (!ready()) return 1;
  Proposal& p = proposal();
  Engine& e = engine();
  if (!q_fine || n_rot <= 0 || n_coarse < 1 || topk <= 0 || batch <= 0) {
    set_error("emc_hip_proposal_init: invalid argument");
    return 1;
  }
  const int C = emc_n_to_samples(n_coarse) / 2;
  std::vector<float> qc(4 * (size_t)C), wc(C);
  if (emc_rotation_table(n_coarse, 1, qc.data(), wc.data()) != C) {
    set_error("emc_hip_proposal_init: coarse table n=%d failed", n_coarse);
    return 1;
  }
  if (topk > C) {
    set_error("emc_hip_proposal_init: topk %d exceeds the %d coarse cells", topk, C);
    return 1;
  }
  std::vector<int> cell(n_rot);
  emc_assign_cells(q_fine, n_rot, qc.data(), C, cell.data());
  // Number the cells along a nearest-neighbour walk through the coarse
  // rotations: cells with nearby numbers are nearby in orientation, so the
  // cell-major contractions work on cells that share patterns at the same
  // time (their count rows stay in cache).
  {
    std::vector<int> walk_id(C, -1);
    int cur = 0;
    for (int pos = 0; pos < C; pos++) {
      walk_id[cur] = pos;
      const float* a = &qc[4 * (size_t)cur];
      double best = -1.0;
      int next = -1;
      for (int c = 0; c < C; c++) {
        if (walk_id[c] >= 0) continue;
        const float* b = &qc[4 * (size_t)c];
        const double d = std::fabs((double)a[0] * b[0] + (double)a[1] * b[1] + (double)a[2] * b[2] +
                                   (double)a[3] * b[3]);
        if (d > best) {
          best = d;
          next = c;
        }
      }
      if (next < 0) break;
      cur = next;
    }
    for (int r = 0; r < n_rot; r++) cell[r] = walk_id[cell[r]];
    // nearest cells of every cell (by rotation angle), in the new numbering
    std::vector<float> qw(4 * (size_t)C);
    for (int t = 0; t < C; t++) std::copy(&qc[4 * (size_t)t], &qc[4 * (size_t)t] + 4, &qw[4 * (size_t)walk_id[t]]);
    const int G = std::min(Proposal::kNbr, C - 1);
    std::vector<int> nbr((size_t)C * Proposal::kNbr, -1);
    std::vector<std::pair<double, int>> d(C);
    for (int c = 0; c < C; c++) {
      const float* a = &qw[4 * (size_t)c];
      for (int o = 0; o < C; o++) {
        const float* b = &qw[4 * (size_t)o];
        d[o] = {-std::fabs((double)a[0] * b[0] + (double)a[1] * b[1] + (double)a[2] * b[2] + (double)a[3] * b[3]), o};
      }
      d[c].first = 1e30;  // not its own neighbour
      std::partial_sort(d.begin(), d.begin() + G, d.end());
      for (int g = 0; g < G; g++) nbr[(size_t)c * Proposal::kNbr + g] = d[g].second;
    }
    proposal().h_nbr.swap(nbr);
  }
  std::vector<int> off(C + 1, 0), rot(n_rot);
  for (int r = 0; r < n_rot; r++) off[cell[r] + 1]++;
  int largest = 0;
  for (int c = 0; c < C; c++) {
    largest = std::max(largest, off[c + 1]);
    off[c + 1] += off[c];
  }
  if (largest > kSlot) {
    set_error("emc_hip_proposal_init: a coarse cell holds %d fine rotations (at most %d); use a finer coarse table",
              largest, kSlot);
    return 1;
  }
  std::vector<int> fill(off.begin(), off.end() - 1);
  for (int r = 0; r < n_rot; r++) rot[fill[cell[r]]++] = r;
  p.release();
  p.C = C;
  p.R = n_rot;
  p.topk = topk;
  p.batch = batch;
  p.seed = seed;
  p.iteration = 0;
  p.h_cell_off = off;
  p.h_cell_rot = rot;
  if (!upload(p.d_cell_off, off, e.stream) || !upload(p.d_cell_rot, rot, e.stream) ||
      !upload(p.d_nbr, p.h_nbr, e.stream))
    return 1;
  sync_stream();
  if (!p.cnn.init(side, C, batch, seed)) return 1;
  p.configured = true;
  return failed() ? 1 : 0;
}

int emc_hip_proposal_cells(int* cell_off, int* cell_rot) {
  const Proposal& p = proposal();
  if (!p.configured) return -1;
  if (cell_off) std::copy(p.h_cell_off.begin(), p.h_cell_off.end(), cell_off);
  if (cell_rot) std::copy(p.h_cell_rot.begin(), p.h_cell_rot.end(), cell_rot);
  return p.C;
}

int emc_hip_proposal_train(const real* d_images, const int* d_mask, int N_2d, const real* d_respons, int N_images,
                           int N_slices, int epochs, float lr, double* loss) {
  if (!ready()) return 1;
  Proposal& p = proposal();
  Engine& e = engine();
  Cnn& n = p.cnn;
  if (!p.configured || !d_images || !d_mask || !d_respons || N_images <= 0 || N_slices != p.R || epochs < 0 ||
      N_2d != n.side * n.side) {
    set_error("emc_hip_proposal_train: invalid argument (proposal configured for %d rotations, side %d)", p.R,
              n.side);
    return 1;
  }
  hipStream_t st = e.stream;
  const auto t0 = std::chrono::steady_clock::now();
  const int N = N_images, B = p.batch;
  DBuf& T = p.d_targets;
  DBuf& perm = p.d_perm;
  if (!T.ensure(sizeof(float) * (size_t)p.C * N) || !perm.ensure(sizeof(int) * N)) return 1;
  coarse_targets(d_respons, N, p.d_cell_off.as<int>(), p.d_cell_rot.as<int>(), p.C, T.as<float>(), st);
  std::vector<int> order(N);
  std::iota(order.begin(), order.end(), 0);
  std::mt19937_64 gen(p.seed ^ (0x9e3779b97f4a7c15ULL * (uint64_t)(n.step + 1)));
  double last = 0.0;
  for (int ep = 0; ep < epochs; ep++) {
    std::shuffle(order.begin(), order.end(), gen);
    EMC_CHECK(hipMemcpyAsync(perm.p, order.data(), sizeof(int) * N, hipMemcpyHostToDevice, st));
    launch_fill(n.loss.as<float>(), 2, 0.0f, st);
    for (int b0 = 0; b0 < N; b0 += B) {
      const int nb = std::min(B, N - b0);
      const int* idx = perm.as<int>() + b0;
      n.forward(d_images, d_mask, N_2d, idx, nb, st);
      softmax_xent(n.logits.as<float>(), p.C, nb, T.as<float>(), N, idx, 1.0f / nb, n.dlogits.as<float>(),
                   n.loss.as<double>(), st);
      n.backward_and_step(nb, lr, st);
    }
    EMC_CHECK(hipMemcpyAsync(&last, n.loss.p, sizeof(double), hipMemcpyDeviceToHost, st));
    sync_stream();
    if (failed()) break;
    last /= N;
  }
  p.st_train_loss = last;
  p.st_train_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (loss) *loss = last;
  return failed() ? 1 : 0;
}

int emc_hip_proposal_predict(const real* d_images, const int* d_mask, int N_2d, int N_images) {
  if (!ready()) return 1;
  Proposal& p = proposal();
  Engine& e = engine();
  Cnn& n = p.cnn;
  if (!p.configured || !d_images || !d_mask || N_images <= 0 || N_2d != n.side * n.side) {
    set_error("emc_hip_proposal_predict: invalid argument");
    return 1;
  }
  hipStream_t st = e.stream;
  const auto t0 = std::chrono::steady_clock::now();
  const int N = N_images, B = p.batch;
  std::vector<int> iota(N);
  std::iota(iota.begin(), iota.end(), 0);
  DBuf idx;
  if (!upload(idx, iota, st) || !p.d_cand.ensure(sizeof(int) * (size_t)N * p.topk) ||
      !p.d_pmass.ensure(sizeof(float) * N))
    return 1;
  for (int b0 = 0; b0 < N; b0 += B) {
    const int nb = std::min(B, N - b0);
    n.forward(d_images, d_mask, N_2d, idx.as<int>() + b0, nb, st);
    topk(n.logits.as<float>(), p.C, nb, p.topk, nullptr, p.d_cand.as<int>() + (size_t)b0 * p.topk,
         p.d_pmass.as<float>() + b0, st);
  }
  p.N = N;
  p.h_pmass.resize(N);
  if (!p.d_net_cand.ensure(sizeof(int) * (size_t)N * p.topk)) return 1;
  EMC_CHECK(hipMemcpyAsync(p.d_net_cand.p, p.d_cand.p, sizeof(int) * (size_t)N * p.topk, hipMemcpyDeviceToDevice, st));
  EMC_CHECK(hipMemcpyAsync(p.h_pmass.data(), p.d_pmass.p, sizeof(float) * N, hipMemcpyDeviceToHost, st));
  sync_stream();
  idx.release();
  if (failed()) return 1;
  reset_modes(p);
  if (p.post == Proposal::POST_CAND) p.post = Proposal::POST_NONE;
  if (!build_lists(p)) return 1;
  p.predicted = true;
  p.iteration = 0;
  p.st_predict_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return 0;
}

int emc_hip_proposal_set_candidates(const int* cand, int N_images, int topk) {
  if (!ready()) return 1;
  Proposal& p = proposal();
  if (!p.configured || !cand || N_images <= 0 || topk <= 0 || topk > p.C) {
    set_error("emc_hip_proposal_set_candidates: invalid argument");
    return 1;
  }
  std::vector<char> seen(p.C, 0);
  for (int i = 0; i < N_images; i++) {
    for (int j = 0; j < topk; j++) {
      const int c = cand[(size_t)i * topk + j];
      if (c < 0 || c >= p.C || seen[c]) {
        set_error("emc_hip_proposal_set_candidates: pattern %d: cell %d out of range or repeated", i, c);
        return 1;
      }
      seen[c] = 1;
    }
    for (int j = 0; j < topk; j++) seen[cand[(size_t)i * topk + j]] = 0;
  }
  p.N = N_images;
  p.topk = topk;
  const std::vector<int> h(cand, cand + (size_t)N_images * topk);
  if (!upload(p.d_cand, h, engine().stream) || !upload(p.d_net_cand, h, engine().stream)) return 1;
  p.h_pmass.assign(N_images, 1.0f);
  reset_modes(p);
  if (p.post == Proposal::POST_CAND) p.post = Proposal::POST_NONE;
  if (!build_lists(p)) return 1;
  p.predicted = true;
  p.iteration = 0;
  return 0;
}

int emc_hip_proposal_get(int* cand, int* mode, float* mass, int N_images) {
  const Proposal& p = proposal();
  if (!p.predicted || N_images != p.N) return 1;
  if (cand) {
    EMC_CHECK(hipMemcpyAsync(cand, p.d_cand.p, sizeof(int) * (size_t)p.N * p.topk, hipMemcpyDeviceToHost,
                             engine().stream));
    sync_stream();
    if (failed()) return 1;
  }
  if (mode)
    for (int i = 0; i < p.N; i++) mode[i] = p.h_mode[i];
  if (mass) std::copy(p.h_pmass.begin(), p.h_pmass.end(), mass);
  return 0;
}

int emc_hip_proposal_shared(int* cells, int max_cells) {
  const Proposal& p = proposal();
  if (cells)
    for (int k = 0; k < p.n_shc && k < max_cells; k++) cells[k] = p.h_sh_cells[k];
  return p.n_shc;
}

int emc_hip_proposal_stats(double* out, int n) {
  const Proposal& p = proposal();
  constexpr int kStats = 22;
  const double v[kStats] = {(double)p.iteration, p.st_full,     p.st_audited,    p.st_fail,
                            p.st_min_mass,       p.st_mean_mass, p.st_pairs,      p.st_gated,
                            p.st_train_loss,     p.st_train_ms,  p.st_predict_ms, (double)p.n_sub,
                            p.st_cand_pairs,     p.st_moved,     p.st_dense,      p.st_shared,
                            p.st_recentre_ms,    p.st_group_pairs, p.st_resid_pairs, (double)p.R_sh,
                            p.st_cleared,        p.st_core_trunc};
  int k = 0;
  for (; k < n && k < kStats; k++) out[k] = v[k];
  return k;
}

}  // extern "C"
