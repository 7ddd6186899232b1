// The EMC C-ABI (emc.h, emc_hip.h) on top of the HIP kernels.
//
// Each entry point keeps the reference's signature and observable behaviour
// (emc_cuda.cu:154-505): the same device buffers are read and written, the
// same host arrays are filled, and every call returns after its work is
// complete.  Inside, the engine keeps derived operands of caller buffers
// (masked-pixel packing of the patterns, packed slices, the per-pattern sums
// that let the scaling update reuse the likelihood's contraction) and
// validates them against the buffers they came from.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "engine.hpp"
#include "proposal.hpp"
#include "proposal_kernels.hpp"

namespace emc {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static char g_err[1024];
static int g_err_code = 0;
static bool g_err_printed = false;

void set_error(const char* fmt, ...) {
  if (g_err_code) return;  // keep the first error
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  g_err_code = 1;
  fprintf(stderr, "emc_hip error: %s\n", g_err);
  g_err_printed = true;
  const char* ab = getenv("EMC_HIP_ABORT_ON_ERROR");
  if (ab && ab[0] == '1') abort();
}

bool failed() { return g_err_code != 0; }

bool DBuf::ensure(size_t b) {
  if (b == 0) b = 1;
  if (p && bytes >= b) return true;
  release();
  if (hipMalloc(&p, b) != hipSuccess) {
    p = nullptr;
    bytes = 0;
    (void)hipGetLastError();
    set_error("out of device memory allocating %zu bytes", b);
    return false;
  }
  bytes = b;
  return true;
}

void DBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
}

// ---------------------------------------------------------------------------
// engine
// ---------------------------------------------------------------------------
static Engine* g_engine = nullptr;

bool Engine::init() {
  if (hipGetDevice(&device) != hipSuccess) {
    set_error("no HIP device available");
    return false;
  }
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0) n_cu = cus;
  if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess) {
    set_error("hipStreamCreate failed");
    return false;
  }
  // the allocation-free normalisation's scratch address on this device
  return noalloc_scratch(device) != nullptr;
}

void Engine::release_all() {
  DBuf* bufs[] = {&d_pix, &d_B, &d_K32, &d_Kt32, &d_K16, &d_K16lo, &d_Kt16, &d_Kt16lo, &d_flags,
                  &d_S32, &d_S16, &d_S16lo, &d_Sscale, &d_A, &d_phi_used, &d_smax, &d_ssum, &d_sacc,
                  &d_phiden, &d_red, &d_invphi, &d_T, &d_Tf, &d_Tphi, &d_ones, &d_Pscale, &d_P16, &d_P16lo,
                  &d_Sc, &d_wts, &d_wT, &d_csr_off, &d_csr_col, &d_csr_cnt, &d_LSt, &d_ins_tiled, &d_ins_pix,
                  &d_ins_col, &d_map_off, &d_map_ent, &d_map_wt, &d_map_heavy, &d_act, &d_zrange, &d_cnt_off, &d_cnt_px,
                  &d_cnt_val, &d_ref_ent, &d_ref_count, &d_ref_pat, &d_ref_tot, &d_ref_slot, &d_ref_tmp,
                  &d_rot_order};
  for (DBuf* b : bufs) b->release();
  order_key = {};
  zrange_key = {};
  map_heavy = 0;
  cnt_images = nullptr;
  pack_key = {};
  map_key = {};
  like = {};
}

Engine& engine() {
  if (!g_engine) {
    g_engine = new Engine();
    if (!g_engine->init()) g_engine->stream = nullptr;
    const char* v = getenv("EMC_HIP_VERBOSE");
    if (v && v[0] == '1') g_engine->verbose = true;
  }
  return *g_engine;
}

bool ready() {
  Engine& e = engine();
  if (failed()) return false;
  return e.stream != nullptr;
}

hipEvent_t Engine::next_event() {
  if (ev_used == ev_pool.size()) {
    hipEvent_t ev = nullptr;
    EMC_CHECK(hipEventCreate(&ev));
    if (failed() || !ev) return nullptr;
    ev_pool.push_back(ev);
  }
  return ev_pool[ev_used++];
}

// Adds the recorded pairs to phase_ms (waits for the last one) and hands the
// pool's events out again.
void Engine::resolve_phase_times() {
  if (!ev_pending.empty()) {
    EMC_CHECK(hipEventSynchronize(ev_pending.back().b));
    for (const PhaseEvents& pe : ev_pending) {
      float ms = 0.0f;
      EMC_CHECK(hipEventElapsedTime(&ms, pe.a, pe.b));
      phase_ms[pe.phase] += ms;
      phase_launches[pe.phase] += 1;
    }
    ev_pending.clear();
  }
  ev_used = 0;
}

PhaseTimer::PhaseTimer(int ph) : phase(ph), on(false) {
  Engine& e = engine();
  if (!e.timing || failed()) return;
  // at most 512 pairs outstanding: resolve (one wait) when the list is full
  if (e.ev_pending.size() >= 512) e.resolve_phase_times();
  start = e.next_event();
  if (!start) return;
  EMC_CHECK(hipEventRecord(start, e.stream));
  on = !failed();
}

PhaseTimer::~PhaseTimer() {
  if (!on) return;
  Engine& e = engine();
  hipEvent_t end = e.next_event();
  if (!end) return;
  EMC_CHECK(hipEventRecord(end, e.stream));
  if (!failed()) e.ev_pending.push_back({start, end, phase});
}

static void sync() { EMC_CHECK(hipStreamSynchronize(engine().stream)); }

// ---------------------------------------------------------------------------
// packing of the patterns and the mask
// ---------------------------------------------------------------------------
static bool ensure_pack(const float* images, const int* mask, int N, int N2d) {
  Engine& e = engine();
  auto& k = e.pack_key;
  if (k.images == images && k.mask == mask && k.N == N && k.N2d == N2d && k.precision == e.precision &&
      k.likelihood == e.likelihood && k.sparse == (int)e.sparse && k.stamp == e.invalidate_stamp)
    return true;
  k = {};
  if (N <= 0 || N2d <= 0) {
    set_error("invalid sizes N_images=%d N_2d=%d", N, N2d);
    return false;
  }
  const int S = (int)std::lround(std::sqrt((double)N2d));
  if (S * S != N2d) {
    set_error("N_2d=%d is not a square detector", N2d);
    return false;
  }
  std::vector<int> hmask(N2d);
  EMC_CHECK(hipMemcpyAsync(hmask.data(), mask, sizeof(int) * N2d, hipMemcpyDeviceToHost, e.stream));
  sync();
  if (failed()) return false;
  std::vector<int> pix, col_of(N2d, -1), ins;
  for (int p = 0; p < N2d; p++)
    if (hmask[p] != 0) {
      col_of[p] = (int)pix.size();
      pix.push_back(p);
    }
  // inserted pixels (mask == 1) in the reference's x-outer, y-inner order
  for (int x = 0; x < S; x++)
    for (int y = 0; y < S; y++)
      if (hmask[y * S + x] == 1) ins.push_back(y * S + x);
  e.n_m = (int)pix.size();
  e.n_mp = (int)round_up(std::max(e.n_m, 1), kPadK);
  e.n_ins = (int)ins.size();
  e.N_pad = (int)round_up(N, kPadN);
  std::vector<int> pixp(e.n_mp, -1), inscol(std::max(e.n_ins, 1)), tiled;
  std::copy(pix.begin(), pix.end(), pixp.begin());
  for (int t = 0; t < e.n_ins; t++) inscol[t] = col_of[ins[t]];
  // the same pixels in 8 x 8 detector tiles (atomic insertion order)
  if (S % 8 == 0) {
    for (int ty = 0; ty < S / 8; ty++)
      for (int tx = 0; tx < S / 8; tx++)
        for (int ly = 0; ly < 8; ly++)
          for (int lx = 0; lx < 8; lx++) {
            const int p = (ty * 8 + ly) * S + tx * 8 + lx;
            if (hmask[p] == 1) tiled.push_back(col_of[p]);
          }
  } else {
    for (int p = 0; p < N2d; p++)
      if (hmask[p] == 1) tiled.push_back(col_of[p]);
  }
  if (!e.d_pix.ensure(sizeof(int) * e.n_mp) || !e.d_kidx.ensure(sizeof(int) * N2d) ||
      !e.d_ins_pix.ensure(sizeof(int) * std::max(e.n_ins, 1)) ||
      !e.d_ins_tiled.ensure(sizeof(int) * std::max(e.n_ins, 1)) ||
      !e.d_ins_col.ensure(sizeof(int) * std::max(e.n_ins, 1)) || !e.d_B.ensure(sizeof(double) * N))
    return false;
  EMC_CHECK(hipMemcpyAsync(e.d_pix.p, pixp.data(), sizeof(int) * e.n_mp, hipMemcpyHostToDevice, e.stream));
  EMC_CHECK(hipMemcpyAsync(e.d_kidx.p, col_of.data(), sizeof(int) * N2d, hipMemcpyHostToDevice, e.stream));
  e.pix_gen++;
  if (e.n_ins) {
    EMC_CHECK(hipMemcpyAsync(e.d_ins_pix.p, ins.data(), sizeof(int) * e.n_ins, hipMemcpyHostToDevice, e.stream));
    EMC_CHECK(hipMemcpyAsync(e.d_ins_col.p, inscol.data(), sizeof(int) * e.n_ins, hipMemcpyHostToDevice, e.stream));
    EMC_CHECK(hipMemcpyAsync(e.d_ins_tiled.p, tiled.data(), sizeof(int) * e.n_ins, hipMemcpyHostToDevice, e.stream));
  }
  const size_t nk = (size_t)e.N_pad * e.n_mp;
  PhaseTimer pt(PH_PACK);
  if (e.precision == PREC_FP32) {
    e.d_K16.release(); e.d_K16lo.release(); e.d_Kt16.release(); e.d_Kt16lo.release();
    if (!e.d_K32.ensure(nk * 4) || !e.d_Kt32.ensure(nk * 4)) return false;
    launch_pack_images_f32(images, e.d_pix.as<int>(), N, N2d, e.n_mp, e.N_pad, e.d_K32.as<float>(),
                           e.d_Kt32.as<float>(), e.d_B.as<double>(), e.stream, e.likelihood == LIKE_POISSON);
    e.k_exact = true;
    e.k_scale = 1.0f;
  } else {
    e.d_K32.release(); e.d_Kt32.release();
    if (!e.d_flags.ensure(16)) return false;
    unsigned int zero[4] = {0, 0, 0, 0}, flags[4];
    EMC_CHECK(hipMemcpyAsync(e.d_flags.p, zero, 16, hipMemcpyHostToDevice, e.stream));
    launch_images_stats(images, e.d_pix.as<int>(), N, N2d, e.n_mp, e.d_flags.as<unsigned int>(), e.stream);
    EMC_CHECK(hipMemcpyAsync(flags, e.d_flags.p, 16, hipMemcpyDeviceToHost, e.stream));
    sync();
    if (failed()) return false;
    float kmax;
    memcpy(&kmax, &flags[1], 4);
    e.k_exact = flags[0] == 0;
    e.k_scale = 1.0f;
    if (!e.k_exact && kmax > 0.0f && std::isfinite(kmax)) {
      int ex;
      std::frexp(kmax, &ex);
      e.k_scale = std::ldexp(1.0f, std::max(-120, std::min(120, 14 - ex)));
    }
    if (!e.d_K16.ensure(nk * 2) || !e.d_Kt16.ensure(nk * 2)) return false;
    if (!e.k_exact) {
      if (!e.d_K16lo.ensure(nk * 2) || !e.d_Kt16lo.ensure(nk * 2)) return false;
    } else {
      e.d_K16lo.release();
      e.d_Kt16lo.release();
    }
    launch_pack_images_f16(images, e.d_pix.as<int>(), N, N2d, e.n_mp, e.N_pad, e.k_scale, !e.k_exact,
                           e.d_K16.as<_Float16>(), e.d_K16lo.as<_Float16>(), e.d_Kt16.as<_Float16>(),
                           e.d_Kt16lo.as<_Float16>(), e.d_B.as<double>(), e.stream, e.likelihood == LIKE_POISSON);
    if (e.sparse && e.likelihood == LIKE_POISSON) {
      // CSR copy of the counts for the sparse E-step; the dense FP16 operands
      // used beside it (M-step, candidates) need the counts exact as well
      if (!e.k_exact) {
        set_error("sparse mode needs integer photon counts <= 2048");
        return false;
      }
      std::vector<int> nnz(N), off(N + 1, 0);
      if (!e.d_csr_off.ensure(sizeof(int) * (N + 1))) return false;
      launch_csr_row_nnz(images, e.d_pix.as<int>(), N, N2d, e.n_mp, e.d_csr_off.as<int>(), e.stream);
      EMC_CHECK(hipMemcpyAsync(nnz.data(), e.d_csr_off.p, sizeof(int) * N, hipMemcpyDeviceToHost, e.stream));
      sync();
      if (failed()) return false;
      for (int i = 0; i < N; i++) off[i + 1] = off[i] + nnz[i];
      e.csr_nnz = (size_t)off[N];
      if (!e.d_csr_col.ensure(sizeof(int) * std::max<size_t>(e.csr_nnz, 1)) ||
          !e.d_csr_cnt.ensure(sizeof(unsigned short) * std::max<size_t>(e.csr_nnz, 1)))
        return false;
      EMC_CHECK(hipMemcpyAsync(e.d_csr_off.p, off.data(), sizeof(int) * (N + 1), hipMemcpyHostToDevice, e.stream));
      launch_csr_fill(images, e.d_pix.as<int>(), N, N2d, e.n_mp, e.d_csr_off.as<int>(), e.d_csr_col.as<int>(),
                      e.d_csr_cnt.as<unsigned short>(), e.stream);
      sync();
    }
  }
  if (failed()) return false;
  k.images = images;
  k.mask = mask;
  k.N = N;
  k.N2d = N2d;
  k.precision = e.precision;
  k.likelihood = e.likelihood;
  k.sparse = (int)e.sparse;
  k.stamp = e.invalidate_stamp;
  e.like.valid = false;
  return true;
}

// The Poisson E-step reads the CSR counts (FP16 log-slice operand).
static bool sparse_estep() {
  const Engine& e = engine();
  return e.sparse && e.likelihood == LIKE_POISSON && e.precision == PREC_FP16;
}

// Packed slice operand for the current precision from an [R][N2d] buffer.
static bool pack_slices(const float* slices, int R, int N2d) {
  Engine& e = engine();
  e.R_pad = (int)round_up(R, kPadR);
  const size_t ns = (size_t)e.R_pad * e.n_mp;
  if (!e.d_A.ensure(sizeof(double) * R)) return false;
  PhaseTimer pt(PH_PACK);
  if (e.precision == PREC_FP32) {
    if (!e.d_S32.ensure(ns * 4)) return false;
    launch_pack_slices_f32(slices, e.d_pix.as<int>(), R, N2d, e.n_mp, e.R_pad, e.d_S32.as<float>(),
                           e.d_A.as<double>(), e.stream, e.likelihood == LIKE_POISSON);
  } else {
    if (!e.d_S16.ensure(ns * 2) || !e.d_S16lo.ensure(ns * 2) || !e.d_Sscale.ensure(sizeof(float) * R)) return false;
    launch_pack_slices_f16(slices, e.d_pix.as<int>(), R, N2d, e.n_mp, e.R_pad, e.d_S16.as<_Float16>(),
                           e.d_S16lo.as<_Float16>(), e.d_Sscale.as<float>(), e.d_A.as<double>(), e.stream,
                           e.likelihood == LIKE_POISSON);
    e.s16_gen++;
  }
  return !failed();
}

static uint64_t g_slices_epoch = 1;   // bumped on every library write to a slices buffer
static uint64_t g_respons_epoch = 1;  // bumped on every library write of L into a respons buffer

// The E-step contraction with either the likelihood epilogue (P == nullptr)
// or the sum_r P_ri C_ir epilogue into D.
static void run_estep(int N, int R, float sigma, const float* phi, float* L, const float* P, double* D) {
  Engine& e = engine();
  LikeEpilogue ep{};
  ep.L = L;
  ep.P = P;
  ep.D = D;
  ep.A = e.d_A.as<double>();
  ep.B = e.d_B.as<double>();
  ep.phi = phi;
  ep.cscale = e.precision == PREC_FP16 ? e.d_Sscale.as<float>() : nullptr;
  ep.kscale_inv = 1.0f / e.k_scale;
  ep.inv_2nms2 = 1.0 / (2.0 * (double)(float)e.n_m * ((double)sigma * (double)sigma));
  ep.N = N;
  ep.R = R;
  ep.poisson = e.likelihood == LIKE_POISSON;
  ep.vecL = (N % 4 == 0) && L && ((uintptr_t)L % 16 == 0);
  const bool csr = L && !P && sparse_estep();
  if (csr && e.lst_gen != e.s16_gen) {
    // the CSR E-step reads the hi/lo operand transposed; built from the
    // packed slices when they changed since (proposal iterations that stay
    // on their candidates never need it)
    PhaseTimer pp(PH_PACK);
    if (!e.d_LSt.ensure(4 * (size_t)e.R_pad * e.n_mp)) return;
    launch_transpose_hilo(e.d_S16.as<_Float16>(), e.d_S16lo.as<_Float16>(), e.R_pad, e.n_mp, e.d_LSt.p, e.stream);
    e.lst_gen = e.s16_gen;
  }
  PhaseTimer pt(PH_LIKE);
  if (csr) {
    launch_estep_csr(e.d_csr_off.as<int>(), e.d_csr_col.as<int>(), e.d_csr_cnt.as<unsigned short>(), e.d_LSt.p,
                     e.R_pad, ep, e.stream);
    return;
  }
  if (e.precision == PREC_FP32)
    launch_estep(PREC_FP32, false, e.d_K32.p, nullptr, e.d_S32.p, nullptr, e.N_pad, e.R_pad, e.n_mp, ep, e.stream);
  else
    launch_estep(PREC_FP16, !e.k_exact, e.d_K16.p, e.d_K16lo.p, e.d_S16.p, e.d_S16lo.p, e.N_pad, e.R_pad, e.n_mp, ep,
                 e.stream);
}

}  // namespace emc

using namespace emc;

// ---------------------------------------------------------------------------
// emc.h
// ---------------------------------------------------------------------------
extern "C" {

void cuda_get_slices(sp_3matrix* model, real* d_model, real* d_slices, real* d_rot, real* d_x_coordinates,
                     real* d_y_coordinates, real* d_z_coordinates, int N_slices) {
  if (!ready()) return;
  Engine& e = engine();
  if (!model || !d_model || !d_slices || !d_rot || N_slices < 0) {
    set_error("cuda_get_slices: invalid argument");
    return;
  }
  if (model->x != model->y || model->x != model->z) {
    set_error("cuda_get_slices: model must be a cube (got %u x %u x %u)", model->x, model->y, model->z);
    return;
  }
  const int S = (int)model->x;
  // FP16 form with a compact pixel list from an earlier E-step of this
  // detector: pack the slice operands in the same pass (the E-step then
  // skips its pack while the slices, pixel list and likelihood are these)
  const bool fuse = e.precision == PREC_FP16 && e.pix_gen > 0 && e.pack_key.N2d == S * S &&
                    expand_pack_fits(S, e.n_mp) && N_slices > 0;
  e.slice_pack = {};
  // gathers from a bricked copy of the model (made here: the model changes
  // every iteration; 2 x S^3 x 4 bytes of traffic)
  const bool bricked = S % 4 == 0 && e.d_model_brick.ensure(sizeof(float) * (size_t)S * S * S);
  const float* gm = bricked ? e.d_model_brick.as<float>() : d_model;
  if (bricked) {
    PhaseTimer pt(PH_EXPAND);
    launch_brick_model(d_model, S, e.d_model_brick.as<float>(), e.stream);
  }
  if (fuse) {
    e.R_pad = (int)round_up(N_slices, kPadR);
    const size_t ns = (size_t)e.R_pad * e.n_mp;
    if (!e.d_A.ensure(sizeof(double) * N_slices) || !e.d_S16.ensure(ns * 2) || !e.d_S16lo.ensure(ns * 2) ||
        !e.d_Sscale.ensure(sizeof(float) * N_slices))
      return;
    PhaseTimer pt(PH_EXPAND);
    launch_expand_pack_f16(gm, S, d_rot, N_slices, e.R_pad, d_x_coordinates, d_y_coordinates, d_z_coordinates,
                           d_slices, e.d_pix.as<int>(), e.d_kidx.as<int>(), e.n_m, e.n_mp, e.d_S16.as<_Float16>(),
                           e.d_S16lo.as<_Float16>(), e.d_Sscale.as<float>(), e.d_A.as<double>(), e.interp,
                           e.likelihood == LIKE_POISSON, e.stream, bricked);
    e.s16_gen++;
  } else {
    PhaseTimer pt(PH_EXPAND);
    launch_expand(gm, S, d_rot, N_slices, d_x_coordinates, d_y_coordinates, d_z_coordinates, d_slices, e.interp,
                  e.stream, bricked);
  }
  g_slices_epoch++;
  if (fuse) e.slice_pack = {d_slices, g_slices_epoch, e.pix_gen, N_slices, e.likelihood};
  sync();
}

void cuda_calculate_responsabilities(real* slices, real* images, int* d_mask, real sigma, real* d_scaling,
                                     real* d_respons, int N_2d, int N_images, int N_slices, real* respons) {
  if (!ready()) return;
  Engine& e = engine();
  if (!slices || !images || !d_mask || !d_scaling || !d_respons || N_slices <= 0) {
    set_error("cuda_calculate_responsabilities: invalid argument");
    return;
  }
  if (!ensure_pack(images, d_mask, N_images, N_2d)) return;
  const auto& sp = e.slice_pack;
  const bool packed = e.precision == PREC_FP16 && sp.slices == slices && sp.epoch == g_slices_epoch &&
                      sp.pix_gen == e.pix_gen && sp.R == N_slices && sp.likelihood == e.likelihood;
  if (!packed && !pack_slices(slices, N_slices, N_2d)) return;
  if (!e.d_phi_used.ensure(sizeof(float) * N_images)) return;
  EMC_CHECK(hipMemcpyAsync(e.d_phi_used.p, d_scaling, sizeof(float) * N_images, hipMemcpyDeviceToDevice, e.stream));
  bool prop = proposal_applies(N_images, N_slices);
  if (prop) {
    if (respons) {
      set_error("cuda_calculate_responsabilities: no host copy of the log-likelihoods in proposal mode");
      return;
    }
    bool dense = false;
    if (!proposal_estep(N_images, N_slices, sigma, e.d_phi_used.as<float>(), &dense)) return;
    prop = !dense;
  }
  if (!prop) {
    run_estep(N_images, N_slices, sigma, e.d_phi_used.as<float>(), d_respons, nullptr, nullptr);
    e.resp_nz.valid = false;
  }
  g_respons_epoch++;
  auto& lk = e.like;
  lk.valid = !failed();
  lk.respons = d_respons;
  lk.slices = slices;
  lk.images = images;
  lk.slices_gen = g_slices_epoch;
  lk.N = N_images;
  lk.R = N_slices;
  lk.N2d = N_2d;
  lk.sigma2x = 2.0 * (double)(float)e.n_m * ((double)sigma * (double)sigma);
  lk.normalized = false;
  lk.have_total = false;
  lk.proposal = prop;
  if (respons) {
    EMC_CHECK(hipMemcpyAsync(respons, d_respons, sizeof(float) * (size_t)N_images * N_slices,
                             hipMemcpyDeviceToHost, e.stream));
    sync();
    if (e.verbose) {
      double s = 0;
      for (size_t k = 0; k < (size_t)N_images * N_slices; k++) s += respons[k];
      printf("respons_sum = %f\n", s);
    }
  }
  sync();
}

void cuda_normalize_responsabilities(real* d_respons, int N_slices, int N_images) {
  if (!ready()) return;
  Engine& e = engine();
  if (!d_respons || N_slices <= 0 || N_images <= 0) {
    set_error("cuda_normalize_responsabilities: invalid argument");
    return;
  }
  auto& lk = e.like;
  if (lk.valid && lk.proposal && !lk.normalized && lk.respons == d_respons && lk.N == N_images &&
      lk.R == N_slices) {
    // responsibilities over each pattern's candidates (whole table for the
    // audited and fallback patterns), written densely into d_respons
    double total = 0.0;
    if (!e.d_phiden.ensure(sizeof(double) * N_images)) return;
    if (!proposal_softmax(d_respons, N_images, N_slices, lk.sigma2x, &total)) return;
    lk.normalized = true;
    lk.have_total = true;
    lk.total = total;
    return;
  }
  e.resp_nz.valid = false;  // normalised in place below
  const bool fuse = lk.valid && !lk.normalized && lk.respons == d_respons && lk.N == N_images && lk.R == N_slices &&
                    e.d_A.p && e.d_B.p;
  const int chunks = softmax_chunks(N_images, N_slices);
  const size_t cn = (size_t)chunks * N_images;
  const int nb = (N_images + 255) / 256;
  if (!e.d_smax.ensure(cn * 4) || !e.d_ssum.ensure(cn * 8) || !e.d_sacc.ensure(4 * cn * 8) ||
      !e.d_phiden.ensure(sizeof(double) * N_images) || !e.d_red.ensure(sizeof(double) * (nb + 1 + 1200)))
    return;
  {
    PhaseTimer pt(PH_SOFTMAX);
    launch_softmax(d_respons, N_slices, N_images, chunks, e.d_smax.as<float>(), e.d_ssum.as<double>(),
                   e.d_sacc.as<double>(), fuse ? e.d_A.as<double>() : nullptr, fuse, e.stream);
    launch_softmax_reduce(N_images, chunks, e.d_sacc.as<double>(), fuse, fuse ? e.d_B.as<double>() : nullptr,
                          fuse ? e.d_phi_used.as<float>() : nullptr, lk.sigma2x, e.d_phiden.as<double>(),
                          e.d_red.as<double>(), e.stream, e.likelihood == LIKE_POISSON);
  }
  double* h_total = e.host_scalar<double>(3);
  EMC_CHECK(hipMemcpyAsync(h_total, e.d_red.as<double>() + nb, sizeof(double), hipMemcpyDeviceToHost, e.stream));
  sync();
  if (failed()) return;
  double total = *h_total;
  // the significant responsibilities from FP64 likelihoods, when the
  // likelihood this buffer holds is the library's own of these slices
  e.ref_entries = 0;
  if (e.refine && fuse && lk.slices && lk.images) {
    PhaseTimer pt(PH_REFINE);
    RefineInputs in;
    in.respons = d_respons;
    in.N = N_images;
    in.R = N_slices;
    in.N2d = lk.N2d;
    in.slices = lk.slices;
    in.images = lk.images;
    in.rows[0].P = d_respons;
    in.rows[0].rows = N_slices;
    in.rows[0].ld = N_images;
    in.rows[0].cols = N_images;
    in.nrows = 1;
    in.phiden = e.d_phiden.as<double>();
    in.poisson = e.likelihood == LIKE_POISSON;
    in.sigma2x = lk.sigma2x;
    in.tau = e.refine_tau;
    double dt = 0.0;
    if (!refine_posterior(in, &dt)) return;
    total += dt;
  }
  if (lk.respons == d_respons && lk.N == N_images && lk.R == N_slices) {
    lk.normalized = fuse;
    lk.have_total = true;
    lk.total = total;
    // a whole-table iteration while the proposal tracks the posterior
    if (fuse && proposal_tracks(N_images, N_slices)) proposal_dense_posterior(d_respons, N_images, N_slices);
  } else {
    lk.valid = false;
  }
}

real emc_hip_total_respons_n(const real* d_respons, size_t n) {
  if (!ready()) return 0.0f;
  Engine& e = engine();
  auto& lk = e.like;
  if (lk.have_total && lk.respons == d_respons && n == (size_t)lk.N * lk.R) return (real)lk.total;
  if (!d_respons) {
    set_error("cuda_total_respons: invalid argument");
    return 0.0f;
  }
  if (!e.d_red.ensure(sizeof(double) * 1100)) return 0.0f;
  double total = 0.0;
  {
    PhaseTimer pt(PH_OTHER);
    launch_total_respons(d_respons, n, e.d_red.as<double>(), e.stream);
  }
  double* h_total = e.host_scalar<double>(4);
  EMC_CHECK(hipMemcpyAsync(h_total, e.d_red.as<double>() + 1024, sizeof(double), hipMemcpyDeviceToHost, e.stream));
  sync();
  total = *h_total;
  return (real)total;
}

real cuda_total_respons(real* d_respons, real* respons, int n) {
  (void)respons;
  if (!ready()) return 0.0f;
  auto& lk = engine().like;
  // The reference passes N_images * N_slices through this int; beyond 2^31
  // the caller's value has wrapped.  If it agrees modulo 2^32 with the shape
  // of the responsibilities this buffer holds, it means that whole matrix.
  const size_t full = (size_t)lk.N * lk.R;
  if (lk.respons == d_respons && (uint32_t)n == (uint32_t)full && full > 0) return emc_hip_total_respons_n(d_respons, full);
  if (n < 0) {
    set_error("cuda_total_respons: invalid argument (n = %d; use emc_hip_total_respons_n for sizes >= 2^31)", n);
    return 0.0f;
  }
  return emc_hip_total_respons_n(d_respons, (size_t)n);
}

void cuda_update_scaling(real* d_images, real* slices, int* d_mask, real* d_respons, real* d_scaling, int N_images,
                         int N_slices, int N_2d, real* scaling) {
  if (!ready()) return;
  Engine& e = engine();
  if (!d_images || !slices || !d_mask || !d_respons || !d_scaling || N_images <= 0 || N_slices <= 0) {
    set_error("cuda_update_scaling: invalid argument");
    return;
  }
  if (!ensure_pack(d_images, d_mask, N_images, N_2d)) return;
  auto& lk = e.like;  // ensure_pack clears lk.valid if it had to repack
  const bool fused = lk.valid && lk.normalized && lk.respons == d_respons && lk.slices == slices &&
                     lk.images == d_images && lk.slices_gen == g_slices_epoch && lk.N == N_images &&
                     lk.R == N_slices && lk.N2d == N_2d;
  if (!e.d_phiden.ensure(sizeof(double) * N_images) || !e.d_invphi.ensure(sizeof(float) * N_images)) return;
  if (!fused) {
    if (!pack_slices(slices, N_slices, N_2d)) return;
    if (e.likelihood == LIKE_POISSON) {
      // Poisson: sum_r P_ri sum_p m S_rp
      launch_p_dot_a(d_respons, N_slices, N_images, e.d_A.as<double>(), e.d_phiden.as<double>(), e.stream);
    } else {
      // the reference's method: recompute sum_p m K S for every (pattern, rotation)
      launch_fill(reinterpret_cast<float*>(e.d_phiden.p), (size_t)N_images * 2, 0.0f, e.stream);
      run_estep(N_images, N_slices, 1.0f, d_scaling, nullptr, d_respons, e.d_phiden.as<double>());
    }
  }
  {
    PhaseTimer pt(PH_SCALING);
    launch_scaling_finish(N_images, e.d_B.as<double>(), e.d_phiden.as<double>(), d_scaling, e.d_invphi.as<float>(),
                          e.stream);
  }
  if (scaling)
    EMC_CHECK(hipMemcpyAsync(scaling, d_scaling, sizeof(float) * N_images, hipMemcpyDeviceToHost, e.stream));
  sync();
}

real cuda_update_slices(real* d_images, real* slices, int* d_mask, real* d_respons, real* d_scaling, int N_images,
                        int N_slices, int N_2d, sp_3matrix* model, real* d_model, real* x_coordinates,
                        real* y_coordinates, real* z_coordinates, real* d_rotations, real* weights, real* d_weight,
                        Setup setup, sp_matrix** images) {
  (void)setup;
  (void)images;
  if (!ready()) return 0.0f;
  Engine& e = engine();
  if (!d_images || !slices || !d_mask || !d_respons || !d_scaling || !model || !d_model || !d_rotations ||
      !weights || !d_weight || N_images <= 0 || N_slices <= 0) {
    set_error("cuda_update_slices: invalid argument");
    return 0.0f;
  }
  const int S = (int)model->x;
  if (model->y != model->x || model->z != model->x || (size_t)S * S != (size_t)N_2d) {
    set_error("cuda_update_slices: model %u^3 does not match N_2d=%d", model->x, N_2d);
    return 0.0f;
  }
  if (!ensure_pack(d_images, d_mask, N_images, N_2d)) return 0.0f;
  const int R = N_slices, N = N_images;
  e.R_pad = (int)round_up(R, kPadR);
  const bool fp16 = e.precision == PREC_FP16;
  const bool poisson = e.likelihood == LIKE_POISSON;
  const Proposal& prp = proposal();
  const bool prop = proposal_applies(N, R) && prp.have_p && prp.respons == d_respons && e.like.valid &&
                    e.like.proposal && e.like.normalized && e.like.respons == d_respons;
  if (!e.d_invphi.ensure(sizeof(float) * N) || !e.d_T.ensure(sizeof(double) * R) || !e.d_Tf.ensure(sizeof(float) * R) ||
      !e.d_Pscale.ensure(sizeof(float) * R) || !e.d_Sc.ensure(sizeof(float) * (size_t)e.R_pad * e.n_mp) ||
      !e.d_wts.ensure(sizeof(float) * R) || !e.d_wT.ensure(sizeof(float) * R) ||
      (poisson && (!e.d_Tphi.ensure(sizeof(float) * R) || !e.d_ones.ensure(sizeof(float) * N))))
    return 0.0f;
  EMC_CHECK(hipMemcpyAsync(e.d_wts.p, weights, sizeof(float) * R, hipMemcpyHostToDevice, e.stream));
  // 1/phi from the current scaling: the M-step divides by the updated
  // factors (emc_atomic.cu:94)
  // Gaussian: U = sum_i P K / phi; Poisson: U = sum_i P K, divided by
  // sum_i P phi instead of T
  if (poisson) launch_fill(e.d_ones.as<float>(), N, 1.0f, e.stream);
  else launch_reciprocal(d_scaling, N, e.d_invphi.as<float>(), e.stream);
  const float* opscale = poisson ? e.d_ones.as<float>() : e.d_invphi.as<float>();
  MstepArgs a{};
  a.P = d_respons;
  a.invphi = opscale;
  a.rowscale = fp16 ? e.d_Pscale.as<float>() : nullptr;
  a.Tf = poisson ? e.d_Tphi.as<float>() : e.d_Tf.as<float>();
  a.poisson = poisson;
  a.pix = e.d_pix.as<int>();
  a.slices = slices;
  a.Sc = e.d_Sc.as<float>();
  a.kscale_inv = 1.0f / e.k_scale;
  a.N = N;
  a.R = R;
  a.N_pad = e.N_pad;
  a.R_pad = e.R_pad;
  a.n_mp = e.n_mp;
  a.n_m = e.n_m;
  a.N2d = N_2d;
  a.vecP = (N % 4 == 0) && ((uintptr_t)d_respons % 16 == 0);
  // FP16 form: the M-step over the nonzero responsibilities when they are
  // few (concentrated posteriors), else the dense matrix-core M-step, or the
  // candidate M-step in a proposal iteration (d_respons holds the complete
  // posterior either way: candidates and whole-table subset scattered)
  // (the candidate M-step covers the candidates and the whole-table subset,
  // not shared rotations: with those, the dense form runs instead)
  const bool try_sparse = fp16 && e.mstep_mode != MSTEP_DENSE;
  const bool cand_ok = prop && !proposal_has_shared();
  bool sparse_m = false;
  e.p_nnz = -1;
  {
    PhaseTimer pt(PH_ROWSUM);
    if (try_sparse && (!e.d_pnnz.ensure(sizeof(int) * R) || !e.d_poff.ensure(sizeof(int) * (R + 1)))) return 0.0f;
    // the nonzero count of every rotation: in a proposal iteration from the
    // candidate slots, shared rows and subset columns, else from a scan of
    // d_respons (with T_r and the row scales of the dense FP16 form)
    const bool from_sources = prop && try_sparse && (uint64_t)R * (uint64_t)N < (1ull << 32);
    NzSources src{};
    if (from_sources) {
      if (!proposal_nz_sources(src, N, R)) return 0.0f;
      launch_nz_count(src, e.d_pnnz.as<int>(), e.stream);
    }
    const bool scan = !from_sources && (try_sparse || !cand_ok);
    if (scan)
      launch_rowsum(d_respons, R, N, opscale, e.d_T.as<double>(), e.d_Tf.as<float>(), e.d_Pscale.as<float>(), fp16,
                    e.stream, poisson ? d_scaling : nullptr, poisson ? e.d_Tphi.as<float>() : nullptr,
                    try_sparse ? e.d_pnnz.as<int>() : nullptr);
    if (try_sparse) {
      e.h_pnnz.resize(R);
      EMC_CHECK(hipMemcpyAsync(e.h_pnnz.data(), e.d_pnnz.p, sizeof(int) * R, hipMemcpyDeviceToHost, e.stream));
      sync();
      if (failed()) return 0.0f;
      long total = 0;
      for (int r = 0; r < R; r++) total += e.h_pnnz[r];
      sparse_m = e.mstep_mode == MSTEP_SPARSE || total * kSparseMstepRatio <= (long)R * N;
      if (sparse_m) {
        if (total > (long)INT32_MAX) {
          set_error("cuda_update_slices: %ld nonzero responsibilities exceed the sparse M-step's 32-bit lists",
                    total);
          return 0.0f;
        }
        e.h_poff.resize(R + 1);
        e.h_poff[0] = 0;
        for (int r = 0; r < R; r++) e.h_poff[r + 1] = e.h_poff[r] + e.h_pnnz[r];
        if (!e.d_pcol.ensure(sizeof(int) * std::max(total, 1L)) || !e.d_pval.ensure(sizeof(float) * std::max(total, 1L)))
          return 0.0f;
        EMC_CHECK(hipMemcpyAsync(e.d_poff.p, e.h_poff.data(), sizeof(int) * (R + 1), hipMemcpyHostToDevice,
                                 e.stream));
        if (from_sources) {
          // lists by a sort of the nonzeros; T_r (and Tphi_r) from them
          if (!build_nz_lists(src, total, e.d_poff.as<int>(), opscale, poisson ? d_scaling : nullptr, e.d_nz_keys,
                              e.d_nz_keys_s, e.d_nz_vals, e.d_nz_vals_s, e.d_nz_tail, e.d_nz_tmp,
                              e.d_pcol.as<int>(), e.d_pval.as<float>(), e.d_T.as<double>(), e.d_Tf.as<float>(),
                              poisson ? e.d_Tphi.as<float>() : nullptr, e.stream))
            return 0.0f;
          // the keys list every nonzero the proposal iteration wrote to d_respons
          e.resp_nz.ptr = d_respons;
          e.resp_nz.N = N;
          e.resp_nz.R = R;
          e.resp_nz.nnz = total;
          e.resp_nz.stamp = e.invalidate_stamp;
          e.resp_nz.valid = true;
        } else {
          launch_p_row_fill(d_respons, R, N, opscale, e.d_poff.as<int>(), e.d_pcol.as<int>(), e.d_pval.as<float>(),
                            e.stream);
        }
        e.p_nnz = total;
      } else if (from_sources && !cand_ok) {
        // dense M-step after all: T_r and the row scales from d_respons
        launch_rowsum(d_respons, R, N, opscale, e.d_T.as<double>(), e.d_Tf.as<float>(), e.d_Pscale.as<float>(), fp16,
                      e.stream, poisson ? d_scaling : nullptr, poisson ? e.d_Tphi.as<float>() : nullptr, nullptr);
      }
    }
    if (fp16 && !sparse_m && !cand_ok) {
      if (!e.d_P16.ensure(2 * (size_t)e.R_pad * e.N_pad) || !e.d_P16lo.ensure(2 * (size_t)e.R_pad * e.N_pad))
        return 0.0f;
      launch_convert_p(a, e.d_P16.as<_Float16>(), e.d_P16lo.as<_Float16>(), e.stream);
    }
  }
  if (cand_ok && !sparse_m) {
    // candidate-restricted M-step (+ the whole-table subset)
    if (!proposal_mstep(N, R, slices, N_2d, d_scaling)) return 0.0f;
  } else {
    PhaseTimer pt(PH_MSTEP);
    if (sparse_m) {
      if (!launch_mstep_sparse(e.h_poff, e.d_poff.as<int>(), e.d_pcol.as<int>(), e.d_pval.as<float>(),
                               e.d_K16.as<_Float16>(), e.k_exact ? nullptr : e.d_K16lo.as<_Float16>(), a, e.stream))
        return 0.0f;
    } else if (fp16)
      launch_mstep16(!e.k_exact, e.d_P16.p, e.d_P16lo.p, e.d_Kt16.p, e.d_Kt16lo.p, a, e.stream);
    else
      launch_mstep32(e.d_Kt32.as<float>(), a, e.stream);
  }
  g_slices_epoch++;
  launch_set_wT(e.d_wts.as<float>(), e.d_Tf.as<float>(), R, e.d_wT.as<float>(), e.stream);
  // deterministic: the brick form or the sorted map, whichever costs less
  // (the two give the same bits).  The brick's longest chain, at the centre
  // of the model, takes every active rotation in turn, about 1 us each; the
  // map reads all R x n_ins (x 8 trilinear) entries, about 4 ps each.
  bool det_brick = e.insert_mode == INSERT_DET_BRICK && insert_brick_fits(S, R, true);
  if (e.insert_mode == INSERT_DETERMINISTIC && insert_brick_fits(S, R, true)) {
    HostVec<float>& Th = e.h_T;
    Th.resize(R);
    EMC_CHECK(hipMemcpyAsync(Th.data(), e.d_Tf.p, sizeof(float) * R, hipMemcpyDeviceToHost, e.stream));
    sync();
    if (failed()) return 0.0f;
    size_t n_act = 0;
    for (int r = 0; r < R; r++) n_act += Th[r] > 1e-10f;
    const double entries = (double)R * e.n_ins * (e.interp == INTERP_NEAREST ? 1.0 : 8.0);
    det_brick = (double)n_act * 1.0e-6 <= entries * 4.0e-12;
  }
  if (det_brick || (e.insert_mode == INSERT_ATOMIC && insert_brick_fits(S, R, false))) {
    PhaseTimer pt(PH_INSERT);
    launch_insert_brick(e.d_Sc.as<float>(), e.n_mp, e.d_kidx.as<int>(), d_mask, e.d_wT.as<float>(), d_rotations, R,
                        S, x_coordinates, y_coordinates, z_coordinates, d_model, d_weight, e.interp, det_brick,
                        e.stream);
  } else if (e.insert_mode == INSERT_DETERMINISTIC || e.insert_mode == INSERT_SORTED ||
             e.insert_mode == INSERT_DET_BRICK) {
    auto& mk = e.map_key;
    const bool have = mk.valid && mk.rot == d_rotations && mk.x == x_coordinates && mk.y == y_coordinates &&
                      mk.z == z_coordinates && mk.mask == d_mask && mk.R == R && mk.S == S && mk.interp == e.interp &&
                      mk.stamp == e.invalidate_stamp && e.map_entries > 0;
    if (!have) {
      mk = {};
      PhaseTimer pt(PH_OTHER);
      launch_build_map(e.d_ins_pix.as<int>(), e.n_ins, d_rotations, R, S, x_coordinates, y_coordinates, z_coordinates,
                       e.interp, e, e.stream);
      if (failed()) return 0.0f;
      mk.valid = true;
      mk.rot = d_rotations;
      mk.x = x_coordinates;
      mk.y = y_coordinates;
      mk.z = z_coordinates;
      mk.mask = d_mask;
      mk.R = R;
      mk.S = S;
      mk.interp = e.interp;
      mk.stamp = e.invalidate_stamp;
    }
    PhaseTimer pt(PH_INSERT);
    launch_insert_gather(e.d_Sc.as<float>(), e.n_mp, e.d_ins_col.as<int>(), e.n_ins, e.d_wT.as<float>(),
                         e.d_map_off.as<uint32_t>(), e.d_map_ent.as<uint32_t>(),
                         e.interp == INTERP_NEAREST ? nullptr : e.d_map_wt.as<float>(), S, e.interp, d_model, d_weight,
                         e.stream);
  } else {
    PhaseTimer pt(PH_INSERT);
    launch_insert_atomic(e.d_Sc.as<float>(), e.n_mp, e.d_pix.as<int>(), e.d_ins_tiled.as<int>(), e.d_wT.as<float>(),
                         d_rotations, R, S,
                         x_coordinates, y_coordinates, z_coordinates, d_model, d_weight, e.interp, e.stream);
  }
  HostVec<float>& T = e.h_T;
  T.resize(R);
  EMC_CHECK(hipMemcpyAsync(T.data(), e.d_Tf.p, sizeof(float) * R, hipMemcpyDeviceToHost, e.stream));
  // the model to the host (emc_cuda.cu:352)
  EMC_CHECK(hipMemcpyAsync(model->data, d_model, sizeof(float) * (size_t)S * S * S, hipMemcpyDeviceToHost, e.stream));
  sync();
  e.like.valid = false;
  float overall = 0.0f;  // emc_cuda.cu:348-351
  for (int r = 0; r < R; r++) overall += T[r];
  return overall;
}

real cuda_model_max(real* model, int model_size) {
  if (!ready()) return 0.0f;
  Engine& e = engine();
  if (!model || model_size < 0) {
    set_error("cuda_model_max: invalid argument");
    return 0.0f;
  }
  if (!e.d_red.ensure(sizeof(double) * 1100)) return 0.0f;
  float* out = e.d_red.as<float>();
  launch_fill(out, 1, 0.0f, e.stream);
  launch_model_max(model, model_size, out, e.stream);
  float* m = e.host_scalar<float>(5);
  EMC_CHECK(hipMemcpyAsync(m, out, sizeof(float), hipMemcpyDeviceToHost, e.stream));
  sync();
  return *m;
}

static void* alloc_dev(size_t bytes) {
  void* p = nullptr;
  if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) {
    (void)hipGetLastError();
    set_error("out of device memory allocating %zu bytes", bytes);
    return nullptr;
  }
  return p;
}

void cuda_allocate_slices(real** slices, Setup setup, int N_slices) {
  if (!ready() || !slices) return;
  *slices = (real*)alloc_dev(sizeof(real) * (size_t)setup.side * setup.side * (size_t)N_slices);
}

void cuda_allocate_model(real** d_model, sp_3matrix* model) {
  if (!ready() || !d_model || !model) return;
  const size_t n = sp_3matrix_size(model);
  *d_model = (real*)alloc_dev(sizeof(real) * n);
  if (*d_model) EMC_CHECK(hipMemcpy(*d_model, model->data, sizeof(real) * n, hipMemcpyHostToDevice));
}

void cuda_allocate_mask(int** d_mask, sp_imatrix* mask) {
  if (!ready() || !d_mask || !mask) return;
  const size_t n = sp_imatrix_size(mask);
  *d_mask = (int*)alloc_dev(sizeof(int) * n);
  if (*d_mask) EMC_CHECK(hipMemcpy(*d_mask, mask->data, sizeof(int) * n, hipMemcpyHostToDevice));
  engine().invalidate_stamp++;
}

void cuda_reset_model(sp_3matrix* model, real* d_model) {
  if (!ready() || !model || !d_model) return;
  launch_fill(d_model, sp_3matrix_size(model), 0.0f, engine().stream);
  sync();
}

void cuda_normalize_model(sp_3matrix* model, real* d_model, real* d_weight) {
  if (!ready()) return;
  Engine& e = engine();
  if (!model || !d_model || !d_weight) {
    set_error("cuda_normalize_model: invalid argument");
    return;
  }
  if (!e.d_red.ensure(sizeof(double) * 1100)) return;
  {
    PhaseTimer pt(PH_NORMALIZE);
    launch_normalize_model(d_model, d_weight, (int)sp_3matrix_size(model), e.d_red.as<double>(), e.stream);
  }
  sync();
}

void cuda_allocate_rotations(real** d_rotations, Quaternion** rotations, int N_slices) {
  if (!ready() || !d_rotations || !rotations || N_slices < 0) return;
  std::vector<float> q((size_t)4 * N_slices);
  for (int i = 0; i < N_slices; i++) memcpy(&q[4 * (size_t)i], rotations[i]->q, sizeof(float) * 4);
  *d_rotations = (real*)alloc_dev(sizeof(real) * q.size());
  if (*d_rotations) EMC_CHECK(hipMemcpy(*d_rotations, q.data(), sizeof(real) * q.size(), hipMemcpyHostToDevice));
  engine().invalidate_stamp++;
}

void cuda_allocate_images(real** d_images, sp_matrix** images, int N_images) {
  if (!ready() || !d_images || !images || N_images <= 0) return;
  const size_t n2 = sp_matrix_size(images[0]);
  *d_images = (real*)alloc_dev(sizeof(real) * n2 * N_images);
  if (!*d_images) return;
  for (int i = 0; i < N_images; i++)
    EMC_CHECK(hipMemcpy(*d_images + n2 * i, images[i]->data, sizeof(real) * n2, hipMemcpyHostToDevice));
  engine().invalidate_stamp++;
}

void cuda_allocate_coords(real** d_x, real** d_y, real** d_z, sp_matrix* x, sp_matrix* y, sp_matrix* z) {
  if (!ready() || !d_x || !d_y || !d_z || !x || !y || !z) return;
  const size_t n = sp_matrix_size(x);
  real** dst[3] = {d_x, d_y, d_z};
  sp_matrix* src[3] = {x, y, z};
  for (int k = 0; k < 3; k++) {
    *dst[k] = (real*)alloc_dev(sizeof(real) * n);
    if (*dst[k]) EMC_CHECK(hipMemcpy(*dst[k], src[k]->data, sizeof(real) * n, hipMemcpyHostToDevice));
  }
  engine().invalidate_stamp++;
}

void cuda_allocate_real(real** x, int n) {
  if (!ready() || !x) return;
  if (n < 0) {
    set_error("cuda_allocate_real: negative size %d", n);
    return;
  }
  *x = (real*)alloc_dev((size_t)n);
}

void cuda_allocate_scaling(real** d_scaling, int N_images) {
  if (!ready() || !d_scaling || N_images < 0) return;
  *d_scaling = (real*)alloc_dev(sizeof(real) * (size_t)N_images);
  if (*d_scaling) launch_fill(*d_scaling, (size_t)N_images, 1.0f, engine().stream);
  sync();
}

// ---------------------------------------------------------------------------
// emc_hip.h
// ---------------------------------------------------------------------------
int emc_hip_last_error(char* buf, size_t n) {
  if (buf && n) {
    strncpy(buf, g_err_code ? g_err : "", n - 1);
    buf[n - 1] = 0;
  }
  return g_err_code;
}

void emc_hip_clear_error(void) {
  g_err_code = 0;
  g_err[0] = 0;
  (void)hipGetLastError();
}

void emc_hip_free(void* d_ptr) {
  if (!d_ptr) return;
  Engine& e = engine();
  if (e.stream) (void)hipStreamSynchronize(e.stream);
  (void)hipFree(d_ptr);
  e.invalidate_stamp++;
  e.like.valid = false;
}

void emc_hip_synchronize(void) {
  if (engine().stream) sync();
}

void emc_hip_shutdown(void) {
  if (!g_engine) return;
  if (g_engine->stream) (void)hipStreamSynchronize(g_engine->stream);
  g_engine->release_all();
  for (hipEvent_t ev : g_engine->ev_pool) (void)hipEventDestroy(ev);
  if (g_engine->stream) (void)hipStreamDestroy(g_engine->stream);
  delete g_engine;
  g_engine = nullptr;
}

void emc_hip_allocate_bytes(void** d_ptr, size_t bytes) {
  if (!ready() || !d_ptr) return;
  *d_ptr = alloc_dev(bytes);
}

int emc_hip_host_alloc(void** h_ptr, size_t bytes) {
  if (!ready() || !h_ptr) return 1;
  *h_ptr = nullptr;
  if (hipHostMalloc(h_ptr, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    *h_ptr = nullptr;
    set_error("out of page-locked host memory allocating %zu bytes", bytes);
    return 1;
  }
  return 0;
}

void emc_hip_host_free(void* h_ptr) {
  if (!h_ptr) return;
  Engine& e = engine();
  if (e.stream) (void)hipStreamSynchronize(e.stream);
  (void)hipHostFree(h_ptr);
}

int emc_hip_memcpy_to_host(void* dst, const void* d_src, size_t bytes) {
  if (!ready()) return 1;
  EMC_CHECK(hipMemcpyAsync(dst, d_src, bytes, hipMemcpyDeviceToHost, engine().stream));
  sync();
  return failed() ? 1 : 0;
}

int emc_hip_memcpy_to_device(void* d_dst, const void* src, size_t bytes) {
  if (!ready()) return 1;
  EMC_CHECK(hipMemcpyAsync(d_dst, src, bytes, hipMemcpyHostToDevice, engine().stream));
  sync();
  Engine& e = engine();
  e.invalidate_stamp++;
  e.like.valid = false;
  return failed() ? 1 : 0;
}

int emc_hip_set_option(const char* key, const char* value) {
  Engine& e = engine();
  if (!key || !value) return 1;
  auto is = [&](const char* a, const char* b) { return strcmp(a, b) == 0; };
  if (is(key, "precision")) {
    if (is(value, "fp32")) e.precision = PREC_FP32;
    else if (is(value, "fp16")) e.precision = PREC_FP16;
    else return 1;
  } else if (is(key, "insert")) {
    if (is(value, "atomic")) e.insert_mode = INSERT_ATOMIC;
    else if (is(value, "global_atomic")) e.insert_mode = INSERT_GLOBAL_ATOMIC;
    else if (is(value, "deterministic")) e.insert_mode = INSERT_DETERMINISTIC;
    else if (is(value, "sorted_scatter")) e.insert_mode = INSERT_SORTED;
    else if (is(value, "deterministic_brick")) e.insert_mode = INSERT_DET_BRICK;
    else return 1;
  } else if (is(key, "interp")) {
    if (is(value, "nearest")) e.interp = INTERP_NEAREST;
    else if (is(value, "trilinear")) e.interp = INTERP_TRILINEAR;
    else return 1;
  } else if (is(key, "timing")) {
    e.timing = value[0] == '1';
  } else if (is(key, "verbose")) {
    e.verbose = value[0] == '1';
  } else if (is(key, "likelihood")) {
    if (is(value, "gaussian")) e.likelihood = LIKE_GAUSSIAN;
    else if (is(value, "poisson")) e.likelihood = LIKE_POISSON;
    else return 1;
  } else if (is(key, "expand_pack")) {
    if (is(value, "auto")) e.expand_tiled = false;
    else if (is(value, "tiled")) e.expand_tiled = true;
    else return 1;
  } else if (is(key, "expand_pixels")) {
    if (is(value, "tiles")) e.expand_rows = false;
    else if (is(value, "rows")) e.expand_rows = true;
    else return 1;
  } else if (is(key, "expand_order")) {
    if (is(value, "normal")) e.expand_sorted = true;
    else if (is(value, "table")) e.expand_sorted = false;
    else return 1;
  } else if (is(key, "expand_rotations")) {
    char* end = nullptr;
    const long v = strtol(value, &end, 10);
    if (!end || *end != 0 || v < 1 || v > 64) return 1;
    e.expand_rotations = (int)v;
  } else if (is(key, "refine")) {
    if (is(value, "on")) e.refine = true;
    else if (is(value, "off")) e.refine = false;
    else return 1;
  } else if (is(key, "refine_tau")) {
    char* end = nullptr;
    const double v = strtod(value, &end);
    if (!end || *end != 0 || !(v > 0.0) || v > 1.0) return 1;
    e.refine_tau = (float)v;
  } else if (is(key, "gemm_tile")) {
    if (is(value, "256")) e.gemm_shape = 0;
    else if (is(value, "128")) e.gemm_shape = 1;
    else if (is(value, "256x128")) e.gemm_shape = 2;
    else if (is(value, "256w8")) e.gemm_shape = 3;
    else return 1;
  } else if (is(key, "gemm_tail")) {
    if (is(value, "split")) e.tail_split = true;
    else if (is(value, "whole")) e.tail_split = false;
    else return 1;
  } else if (is(key, "respons_fill")) {
    if (is(value, "delta")) e.respons_delta = true;
    else if (is(value, "full")) e.respons_delta = false;
    else return 1;
  } else if (is(key, "sparse")) {
    if (is(value, "on")) e.sparse = true;
    else if (is(value, "off")) e.sparse = false;
    else return 1;
  } else if (is(key, "mstep")) {
    if (is(value, "auto")) e.mstep_mode = MSTEP_AUTO;
    else if (is(value, "dense")) e.mstep_mode = MSTEP_DENSE;
    else if (is(value, "sparse")) e.mstep_mode = MSTEP_SPARSE;
    else return 1;
  } else if (is(key, "proposal")) {
    if (is(value, "on")) proposal().active = true;
    else if (is(value, "off")) proposal().active = false;
    else return 1;
  } else if (is(key, "audit_fraction") || is(key, "min_mass") || is(key, "audit_max_fail")) {
    char* end = nullptr;
    const double v = strtod(value, &end);
    if (!end || *end != 0 || !(v >= 0.0 && v <= 1.0)) return 1;
    Proposal& p = proposal();
    (is(key, "min_mass") ? p.min_mass : is(key, "audit_max_fail") ? p.max_fail : p.audit_frac) = (float)v;
  } else if (is(key, "proposal_shared")) {
    char* end = nullptr;
    const long v = strtol(value, &end, 10);
    if (!end || *end != 0 || v < 0 || v > (1 << 20)) return 1;
    proposal().max_shared = (int)v;
  } else if (is(key, "proposal_group_min")) {
    char* end = nullptr;
    const long v = strtol(value, &end, 10);
    if (!end || *end != 0 || v < 0 || v > (1L << 30)) return 1;
    proposal().group_min = (int)v;
  } else if (is(key, "proposal_core_div")) {
    char* end = nullptr;
    const long v = strtol(value, &end, 10);
    if (!end || *end != 0 || v < 1 || v > 64) return 1;
    proposal().core_div = (int)v;
  } else if (is(key, "proposal_track")) {
    if (is(value, "on")) proposal().track = true;
    else if (is(value, "off")) proposal().track = false;
    else return 1;
  } else if (is(key, "proposal_gate")) {
    if (is(value, "predicted")) proposal().gate_predicted = true;
    else if (is(value, "audit")) proposal().gate_predicted = false;
    else return 1;
  } else {
    return 1;
  }
  e.like.valid = false;
  return 0;
}

void emc_hip_invalidate(void) {
  Engine& e = engine();
  e.invalidate_stamp++;
  e.like.valid = false;
  e.slice_pack = {};
}

static const char* kPhaseNames[PH_COUNT] = {"expand", "pack", "likelihood", "softmax", "scaling",
                                            "mstep_prep", "mstep", "insert", "normalize", "other", "audit",
                                            "shared", "groups", "audit_gemm", "refine"};

int emc_hip_phase_times(double* ms, long long* launches, int n) {
  Engine& e = engine();
  e.resolve_phase_times();
  int k = 0;
  for (; k < n && k < PH_COUNT; k++) {
    if (ms) ms[k] = e.phase_ms[k];
    if (launches) launches[k] = e.phase_launches[k];
  }
  return k;
}

void emc_hip_reset_phase_times(void) {
  Engine& e = engine();
  e.resolve_phase_times();
  for (int k = 0; k < PH_COUNT; k++) {
    e.phase_ms[k] = 0;
    e.phase_launches[k] = 0;
  }
}

const char* emc_hip_phase_name(int k) { return k >= 0 && k < PH_COUNT ? kPhaseNames[k] : nullptr; }

int emc_hip_images_from_csr(real* d_images, const int* row_off, const int* pix, const unsigned short* counts,
                            int N_images, int N_2d) {
  if (!ready()) return 1;
  Engine& e = engine();
  if (!d_images || !row_off || N_images <= 0 || N_2d <= 0 || row_off[0] != 0) {
    set_error("emc_hip_images_from_csr: invalid argument");
    return 1;
  }
  const size_t nnz = (size_t)row_off[N_images];
  for (int i = 0; i < N_images; i++)
    if (row_off[i + 1] < row_off[i]) {
      set_error("emc_hip_images_from_csr: row offsets decrease at pattern %d", i);
      return 1;
    }
  for (size_t j = 0; j < nnz; j++)
    if (pix[j] < 0 || pix[j] >= N_2d) {
      set_error("emc_hip_images_from_csr: pixel index %d out of range", pix[j]);
      return 1;
    }
  DBuf d_off, d_pix, d_cnt;
  if (!d_off.ensure(sizeof(int) * (N_images + 1)) || !d_pix.ensure(sizeof(int) * std::max<size_t>(nnz, 1)) ||
      !d_cnt.ensure(sizeof(unsigned short) * std::max<size_t>(nnz, 1)))
    return 1;
  EMC_CHECK(hipMemcpyAsync(d_off.p, row_off, sizeof(int) * (N_images + 1), hipMemcpyHostToDevice, e.stream));
  if (nnz) {
    EMC_CHECK(hipMemcpyAsync(d_pix.p, pix, sizeof(int) * nnz, hipMemcpyHostToDevice, e.stream));
    EMC_CHECK(hipMemcpyAsync(d_cnt.p, counts, sizeof(unsigned short) * nnz, hipMemcpyHostToDevice, e.stream));
  }
  launch_images_from_csr(d_off.as<int>(), d_pix.as<int>(), d_cnt.as<unsigned short>(), N_images, N_2d, d_images,
                         e.stream);
  sync();
  e.invalidate_stamp++;  // the patterns at d_images changed
  e.like.valid = false;
  return failed() ? 1 : 0;
}

int emc_hip_pack_info(int* out, int n) {
  Engine& e = engine();
  const int v[4] = {e.n_m, e.n_mp, e.k_exact ? 1 : 0, e.precision == PREC_FP16 ? (e.k_exact ? 2 : 3) : 1};
  int k = 0;
  for (; k < n && k < 4; k++) out[k] = v[k];
  return k;
}

long long emc_hip_mstep_nnz(void) { return engine().p_nnz; }

long long emc_hip_refine_count(void) { return engine().ref_entries; }

}  // extern "C"
