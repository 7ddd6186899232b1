// Internal state and kernel launchers of libemc_hip (not part of the C-ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <new>
#include <vector>

#include "emc.h"
#include "emc_hip.h"

namespace emc {

// ---- errors -------------------------------------------------------------
void set_error(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
bool failed();  // sticky error present

#define EMC_CHECK(call)                                                                       \
  do {                                                                                        \
    hipError_t e_ = (call);                                                                   \
    if (e_ != hipSuccess) ::emc::set_error("%s:%d: %s: %s", __FILE__, __LINE__, #call,        \
                                           hipGetErrorString(e_));                            \
  } while (0)
#define EMC_LAUNCHED() EMC_CHECK(hipGetLastError())

// ---- page-locked host memory for the library's own transfers -------------
// A device->host copy into pageable memory is staged by the runtime copy by
// copy, the host waiting for each; into page-locked memory the copies queue
// on the stream and the host waits once.  HostVec is a std::vector in
// page-locked memory (ordinary memory if page-locking fails).
template <class T>
struct PinnedAllocator {
  using value_type = T;
  static constexpr size_t kHead = 64;  // in front of each block: 1 page-locked, 0 malloc
  PinnedAllocator() = default;
  template <class U>
  PinnedAllocator(const PinnedAllocator<U>&) {}
  T* allocate(size_t n) {
    const size_t bytes = kHead + n * sizeof(T);
    char* p = nullptr;
    if (hipHostMalloc((void**)&p, bytes, hipHostMallocDefault) == hipSuccess && p) {
      *reinterpret_cast<int*>(p) = 1;
    } else {
      (void)hipGetLastError();
      p = static_cast<char*>(std::malloc(bytes));
      if (!p) throw std::bad_alloc();
      *reinterpret_cast<int*>(p) = 0;
    }
    return reinterpret_cast<T*>(p + kHead);
  }
  void deallocate(T* q, size_t) {
    if (!q) return;
    char* p = reinterpret_cast<char*>(q) - kHead;
    if (*reinterpret_cast<int*>(p) == 1)
      (void)hipHostFree(p);
    else
      std::free(p);
  }
  template <class U>
  bool operator==(const PinnedAllocator<U>&) const { return true; }
  template <class U>
  bool operator!=(const PinnedAllocator<U>&) const { return false; }
};
template <class T>
using HostVec = std::vector<T, PinnedAllocator<T>>;

// ---- device buffers owned by the engine ----------------------------------
struct DBuf {
  void* p = nullptr;
  size_t bytes = 0;
  // Grow to at least `b` bytes (contents not preserved).  False on failure.
  bool ensure(size_t b);
  void release();
  template <class T>
  T* as() const { return static_cast<T*>(p); }
};

// ---- phases for timing -----------------------------------------------------
enum Phase {
  PH_EXPAND = 0, PH_PACK, PH_LIKE, PH_SOFTMAX, PH_SCALING, PH_ROWSUM, PH_MSTEP,
  PH_INSERT, PH_NORMALIZE, PH_OTHER, PH_AUDIT, PH_SHARED, PH_GROUPS, PH_AUDIT_GEMM, PH_REFINE, PH_COUNT
};

enum Precision { PREC_FP32 = 0, PREC_FP16 = 1 };
// atomic: brick-owned LDS accumulation (per-pixel global atomics where the
// model side is not a multiple of the brick); global_atomic: per-pixel
// global atomics; deterministic: the sorted-scatter gather
enum InsertMode {
  INSERT_ATOMIC = 0,
  INSERT_DETERMINISTIC = 1,
  INSERT_GLOBAL_ATOMIC = 2,
  INSERT_SORTED = 3,
  INSERT_DET_BRICK = 4
};
// auto: the nonzero-list M-step when at most 1 in kSparseMstepRatio
// responsibilities is nonzero, else the dense matrix-core M-step
enum MstepMode { MSTEP_AUTO = 0, MSTEP_DENSE = 1, MSTEP_SPARSE = 2 };
constexpr long kSparseMstepRatio = 200;
enum Interp { INTERP_NEAREST = 0, INTERP_TRILINEAR = 1 };
// Gaussian: the reference's L = -sum m (S - K/phi)^2 / (2 n_m sigma^2).
// Poisson: L = sum m (K log(phi max(S, kSlogFloor)) - phi S); the floor bounds
// the range of the FP16 log-slice operand.
enum Likelihood { LIKE_GAUSSIAN = 0, LIKE_POISSON = 1 };
constexpr float kSlogFloor = 1e-6f;

// Tile sizes shared by the GEMM kernels and the padding of packed operands.
constexpr int kTileM = 128;    // FP32 kernels: 128 x 128
constexpr int kTileN = 128;
constexpr int kTile16 = 256;   // FP16 kernels: 256 x 256
constexpr int kPadK = 256;     // pixel padding: multiple of every BK and pixel tile
constexpr int kPadN = 256;     // pattern padding: multiple of every pattern tile
constexpr int kPadR = 256;     // rotation padding: multiple of every rotation tile

inline size_t round_up(size_t x, size_t m) { return (x + m - 1) / m * m; }

struct Engine {
  int device = -1;
  int n_cu = 256;          // compute units of the device
  DBuf d_tail_part;        // FP32 partial tiles of the split-k tail of the FP16 likelihood GEMM
  hipStream_t stream = nullptr;

  // options
  int precision = PREC_FP32;
  int insert_mode = INSERT_ATOMIC;
  int interp = INTERP_NEAREST;
  int likelihood = LIKE_GAUSSIAN;
  bool sparse = false;  // Poisson E-step from CSR counts (see k_sparse.hip)
  int mstep_mode = MSTEP_AUTO;  // FP16 form: M-step over the nonzero responsibilities
  bool respons_delta = true;    // proposal iterations rewrite only the changed responsibilities
  bool tail_split = true;       // FP16 likelihood GEMM: last partial round of tiles split along k
  int gemm_shape = 0;           // FP16 likelihood GEMM tiles: 0 256 x 256 (16 waves), 1 128 x 128 (two per unit),
                                // 2 256 x 128, 3 256 x 256 (8 waves)
  bool expand_tiled = false;    // fused expand + pack in rotation blocks x pixel tiles even where the slice fits LDS
  bool expand_sorted = true;    // expand takes the rotations in the order of their slice normals (k_expand.hip)
  bool expand_rows = true;      // fused expand + pack: pixels in rows (true) or 8 x 8 tiles per wave
  int expand_rotations = 8;     // rotations per workgroup of that form
  bool refine = true;           // responsibilities >= refine_tau recomputed from FP64 likelihoods (k_refine.hip)
  float refine_tau = 1e-9f;
  bool timing = false;
  bool verbose = false;

  // generation counters: bumped whenever the library writes a buffer kind
  uint64_t gen = 1;
  uint64_t next_gen() { return ++gen; }

  // ---- packed patterns (depends on images, mask, N, N2d, precision) ----
  struct {
    const float* images = nullptr;
    const int* mask = nullptr;
    int N = 0, N2d = 0, precision = -1, likelihood = -1, sparse = -1;
    uint64_t stamp = 0;  // invalidation stamp at build time
  } pack_key;
  uint64_t pix_gen = 0;  // bumped whenever the compact pixel list is rebuilt
  // slice operands packed by the expand itself (FP16 form): valid for this
  // slices buffer at this write epoch, pixel list and likelihood
  struct {
    const float* slices = nullptr;
    uint64_t epoch = 0, pix_gen = 0;
    int R = 0, likelihood = -1;
  } slice_pack;
  int n_m = 0;      // masked pixels (mask != 0)
  int n_mp = 0;     // padded to kPadK
  int n_ins = 0;    // pixels with mask == 1 (inserted)
  int N_pad = 0;    // patterns padded to the GEMM tile
  bool k_exact = true;   // FP16 operand holds the counts exactly
  float k_scale = 1.0f;  // FP16 operand = K * k_scale (power of two)
  DBuf d_pix;       // int [n_mp]: pixel index of compact column, -1 padding
  DBuf d_B;         // double [N]: sum_p m K^2 (Poisson: sum_p m K)
  DBuf d_K32;       // float [N_pad][n_mp]
  DBuf d_Kt32;      // float [n_mp][N_pad]
  DBuf d_K16, d_K16lo;    // half [N_pad][n_mp]
  DBuf d_Kt16, d_Kt16lo;  // half [n_mp][N_pad]
  DBuf d_flags;     // scratch flags for the pack kernels
  // CSR copy of the counts (sparse mode): row offsets [N+1], compact pixel
  // index and count of every nonzero
  DBuf d_csr_off, d_csr_col, d_csr_cnt;
  size_t csr_nnz = 0;
  // every pattern's nonzero masked counts for the refinement: offsets
  // [N+1], pixel index, count (float); built for cnt_images at cnt_pix_gen
  DBuf d_cnt_off, d_cnt_px, d_cnt_val;
  const float* cnt_images = nullptr;
  int cnt_N = 0;
  uint64_t cnt_pix_gen = 0, cnt_stamp = 0;
  // refinement scratch: entries, their count, per-pattern sums, sum P log P change
  DBuf d_ref_ent, d_ref_count, d_ref_pat, d_ref_tot, d_ref_slot, d_ref_tmp;
  double ref_entries = 0;  // entries the last refinement recomputed

  // ---- packed slices (per likelihood call) ----
  int R_pad = 0;
  DBuf d_S32;           // float [R_pad][n_mp]
  DBuf d_S16, d_S16lo;  // half [R_pad][n_mp]
  DBuf d_Sscale;        // float [R]: 1/scale_r, where the FP16 operand is S * scale_r (power of two)
  DBuf d_A;             // double [R]: sum_p m S^2 (Poisson: sum_p m S)
  DBuf d_LSt;           // half2 [n_mp][R_pad]: (hi, lo) transposed (sparse mode)
  DBuf d_kidx;          // int [N2d]: packed index of each masked-in pixel (inverse of d_pix), -1 elsewhere
  DBuf d_redo;          // int [R_pad]: rows the provisional-scale pack must redo
  DBuf d_xpart_max, d_xpart_A;  // [R_pad][pixel tiles]: per-tile row maximum and A_r partial of that pack
  DBuf d_rot_order;     // int [R]: expand's processing order of the rotations, for the table in order_key
  struct {
    const float* z = nullptr;
    const int* mask = nullptr;
    int S = 0;
    uint64_t stamp = 0;
  } zrange_key;         // d_zrange holds the z range of these coordinates and mask
  struct {
    const float* rot = nullptr;
    int R = 0;
    uint64_t stamp = 0;
  } order_key;
  int sscale_R = -1;    // rotations whose row scales d_Sscale holds (-1: none)
  uint64_t s16_gen = 0;  // bumped whenever d_S16 / d_S16lo are written
  uint64_t lst_gen = 0;  // s16_gen of the slices d_LSt holds (built when the CSR E-step needs it)

  // ---- likelihood -> softmax -> scaling fusion state ----
  struct {
    bool valid = false;
    const float* respons = nullptr;  // buffer holding L (then P)
    const float* slices = nullptr;
    const float* images = nullptr;
    uint64_t slices_gen = 0;
    int N = 0, R = 0, N2d = 0;
    double sigma2x = 0.0;  // 2 n_m sigma^2
    bool normalized = false;  // softmax ran on this buffer after the likelihood
    bool have_total = false;
    double total = 0.0;       // sum P log P from the softmax pass
    bool proposal = false;    // L came from the candidate-restricted E-step
  } like;
  DBuf d_phi_used;   // float [N]: scaling used by the likelihood
  DBuf d_smax;       // float  [chunks][N] partial maxima
  DBuf d_ssum;       // double [chunks][N] partial sums
  DBuf d_sacc;       // double [4][chunks][N] partial P-weighted sums
  DBuf d_phiden;     // double [N]: sum_r P_ri C_ir
  DBuf d_red;        // double scratch for reductions
  DBuf d_invphi;     // float [N]: 1/scaling for the M-step

  // ---- M-step ----
  DBuf d_T;          // double [R]
  DBuf d_Tf;         // float [R]
  DBuf d_Tphi;       // float [R]: sum_i P_ri phi_i (Poisson M-step divisor)
  DBuf d_ones;       // float [N]: 1 (Poisson: the M-step operand is P itself)
  DBuf d_Pscale;     // float [R]: FP16 M-step operand row scale
  DBuf d_P16, d_P16lo;  // half [R_pad][N_pad]: FP16 M-step operand (P/phi * row scale) hi/lo
  DBuf d_pnnz;       // int [R]: nonzero responsibilities of each rotation
  DBuf d_poff;       // int [R+1]: their offsets in the lists below
  DBuf d_pcol, d_pval;  // int / float [nnz]: pattern, P * operand scale (patterns ascending)
  long p_nnz = -1;      // nonzeros of the last M-step (-1: dense form)
  HostVec<int> h_pnnz;       // read back from the device
  std::vector<int> h_poff;   // uploaded (staged by the runtime: the vector may change right after)
  DBuf d_nz_keys, d_nz_keys_s, d_nz_vals, d_nz_vals_s, d_nz_tail, d_nz_tmp;  // nonzero lists from sources
  // What a responsibility buffer holds, as far as this library wrote it:
  // when valid, every nonzero of `ptr` ([R][N]) is at one of the nnz
  // sorted keys r * N + i in d_nz_keys_s (the last M-step's lists, built
  // from the posterior the library wrote there), zeros elsewhere.  The next
  // proposal iteration then zeroes those entries instead of the whole matrix.
  struct {
    const float* ptr = nullptr;
    int N = 0, R = 0;
    long nnz = 0;
    uint64_t stamp = 0;
    bool valid = false;
  } resp_nz;
  // long rows of the list M-step: segments (first, end nonzero), rows,
  // their first segment, partial sums [segments][n_mp]
  std::vector<int2> h_sp_seg;
  std::vector<int> h_sp_rows, h_sp_rowseg, h_sp_light, h_sp_zero;
  DBuf d_sp_seg, d_sp_rows, d_sp_rowseg, d_sp_part, d_sp_light, d_sp_zero;
  DBuf d_mask4;          // int [N2d / 4]: bit q set when pixel 4v + q is masked in (compact pixel list pix_gen)
  uint64_t mask4_gen = 0;
  DBuf d_Sc;         // float [R_pad][n_mp]: updated compact slices for insertion
  DBuf d_model_brick;  // float [S^3]: the model in 4 x 4 x 4 bricks for the expand's gathers
  DBuf d_wts;        // float [R]: rotation weights (uploaded per call)
  DBuf d_wT;         // float [R]: weights[r] * T_r
  DBuf d_act;        // int [R + 1]: rotations with T_r > 1e-10, then their count (brick insertion)
  DBuf d_zrange;     // float [2]: range of the detector's z over the inserted pixels

  // ---- insertion map for the deterministic mode ----
  struct {
    const float* rot = nullptr;
    const float *x = nullptr, *y = nullptr, *z = nullptr;
    const int* mask = nullptr;
    int R = 0, S = 0, interp = -1;
    uint64_t stamp = 0;
    bool valid = false;
  } map_key;
  DBuf d_ins_pix;    // int [n_ins]: inserted pixels in x-major order
  DBuf d_ins_col;    // int [n_ins]: their compact column
  DBuf d_ins_tiled;  // int [n_ins]: compact columns of the inserted pixels in 8 x 8 tiles
  DBuf d_map_off;    // uint32 [S^3 + 1]
  DBuf d_map_ent;    // uint32 [entries]: entry id e = r * n_ins + t
  DBuf d_map_wt;     // float [entries]: trilinear weight (trilinear only)
  DBuf d_map_heavy;  // uint32 [S^3 + 1]: voxels with >= 64 entries (the count last)
  size_t map_entries = 0;
  uint32_t map_heavy = 0;

  uint64_t invalidate_stamp = 1;

  // ---- timing: an event pair per timed phase, recorded on the stream and
  // resolved only when the times are read (no synchronisation inside an
  // iteration) ----
  double phase_ms[PH_COUNT] = {0};
  long long phase_launches[PH_COUNT] = {0};
  struct PhaseEvents {
    hipEvent_t a, b;
    int phase;
  };
  std::vector<hipEvent_t> ev_pool;       // created events, reused after each resolve
  size_t ev_used = 0;                    // pool events recorded since the last resolve
  std::vector<PhaseEvents> ev_pending;   // recorded pairs not yet added to phase_ms
  hipEvent_t next_event();
  void resolve_phase_times();

  // page-locked slots for device->host reads of single values (slot k used
  // by one call site; each is read after the stream is synchronised)
  HostVec<double> h_scalars = HostVec<double>(16);
  HostVec<float> h_T;  // per-rotation totals read back by the M-step
  template <class T>
  T* host_scalar(int k) {
    return reinterpret_cast<T*>(h_scalars.data() + k);
  }

  bool init();
  void release_all();
};

Engine& engine();          // initialised on first use (device = current)
bool ready();              // engine initialised and no sticky error

// Phase timer: records events around a region when timing is on.
struct PhaseTimer {
  int phase;
  bool on;
  hipEvent_t start = nullptr;
  explicit PhaseTimer(int ph);
  ~PhaseTimer();
};

// ---- kernel launchers (defined in the k_*.hip files) ----------------------
// expand
// (bricked: model is the 4 x 4 x 4-brick copy made by launch_brick_model)
void launch_expand(const float* model, int S, const float* rot, int R, const float* x,
                   const float* y, const float* z, float* slices, int interp, hipStream_t st, bool bricked = false);
void launch_brick_model(const float* model, int S, float* out, hipStream_t st);
// pack patterns: compact columns, B_i, FP32 and FP16 copies (+ transposes)
void launch_pack_mask(const int* mask, int N2d, int* pix, int n_mp, hipStream_t st);
// B = sum_p m K^2, or sum_p m K when poisson
void launch_pack_images_f32(const float* images, const int* pix, int N, int N2d, int n_mp,
                            int N_pad, float* K32, float* Kt32, double* B, hipStream_t st,
                            bool poisson = false);
void launch_pack_images_f16(const float* images, const int* pix, int N, int N2d, int n_mp,
                            int N_pad, float kscale, bool split, _Float16* K16, _Float16* K16lo,
                            _Float16* Kt16, _Float16* Kt16lo, double* B, hipStream_t st,
                            bool poisson = false);
void launch_images_stats(const float* images, const int* pix, int N, int N2d, int n_mp,
                         unsigned int* flags, hipStream_t st);
// pack slices: compact FP32 or FP16 hi/lo, per-rotation scale, A_r; with
// poisson the operand is log(max(S, kSlogFloor)) and A_r = sum_p m S
void launch_pack_slices_f32(const float* slices, const int* pix, int R, int N2d, int n_mp,
                            int R_pad, float* S32, double* A, hipStream_t st, bool poisson = false);
bool expand_pack_fits(int S, int n_mp);
// kidx: packed index of every pixel (-1: masked out); n_m masked pixels,
// packed entries n_m .. n_mp - 1 are padding
void launch_expand_pack_f16(const float* model, int S, const float* rot, int R, int R_pad, const float* x,
                            const float* y, const float* z, float* slices, const int* pix, const int* kidx, int n_m,
                            int n_mp, _Float16* S16, _Float16* S16lo, float* scale, double* A, int interp,
                            bool poisson, hipStream_t st, bool bricked = false);
void launch_pack_slices_f16(const float* slices, const int* pix, int R, int N2d, int n_mp,
                            int R_pad, _Float16* S16, _Float16* S16lo, float* scale, double* A,
                            hipStream_t st, bool poisson = false);

// GEMMs (k_gemm.hip)
struct LikeEpilogue {
  float* L;               // [R][N] output, or nullptr
  const float* P;         // [R][N] responsibilities (PC mode), or nullptr
  double* D;              // [N] sum_r P_ri C_ir accumulator (PC mode)
  const double* A;        // [R]
  const double* B;        // [N]
  const float* phi;       // [N]
  const float* cscale;    // [R] C = acc * cscale[r] (FP16 path), or nullptr
  float kscale_inv;       // 1 / k_scale (FP16 path)
  double inv_2nms2;       // 1 / (2 n_m sigma^2)
  int N, R;
  int poisson;            // L = C + B_i log phi_i - phi_i A_r (C over log S)
  int vecL;               // L rows 16-byte aligned (N % 4 == 0, aligned base)
};
void launch_estep(int precision, bool ksplit, const void* K, const void* Klo, const void* S,
                  const void* Slo, int N_pad, int R_pad, int n_mp, const LikeEpilogue& ep,
                  hipStream_t st);
struct MstepArgs {
  const float* P;         // [R][N]
  const float* invphi;    // [N]
  const float* rowscale;  // [R] (FP16 path)
  const float* Tf;        // [R] divisor of U
  const int* pix;         // [n_mp]
  float* slices;          // [R][N2d] full output (masked columns written)
  float* Sc;              // [R_pad][n_mp] compact output
  float kscale_inv;
  int N, R, N_pad, R_pad, n_mp, n_m, N2d;
  int vecP;               // P rows 16-byte aligned (N % 4 == 0, aligned base)
  int raw = 0;            // write the unnormalised sums U to Sc only
  int poisson = 0;        // S' = Tf > 1e-30 ? U / Tf : 0 (else the reference's rule)
};
void launch_mstep32(const float* Kt, const MstepArgs& a, hipStream_t st);
// FP16 hi/lo forms (k_gemm16.hip)
void launch_estep16(bool ksplit, const void* K, const void* Klo, const void* S, const void* Slo, int N_pad,
                    int R_pad, int n_mp, const LikeEpilogue& ep, hipStream_t st);
void launch_convert_p(const MstepArgs& a, _Float16* hi, _Float16* lo, hipStream_t st);
void launch_mstep16(bool ksplit, const void* Phi, const void* Plo, const void* Kt, const void* Ktlo,
                    const MstepArgs& a, hipStream_t st);

// softmax / totals / scaling (k_softmax.hip)
int softmax_chunks(int N, int R);
void launch_softmax(float* L, int R, int N, int chunks, float* smax, double* ssum, double* sacc,
                    const double* A, bool fuse_phi, hipStream_t st);
// the two passes separately: chunk partials, then P with nmerge >= chunks
// partials merged (smax / ssum / sacc hold [nmerge][N])
void launch_softmax_partial(const float* L, int R, int N, int chunks, float* smax, double* ssum, hipStream_t st);
void launch_softmax_finish(float* L, int R, int N, int chunks, int nmerge, const float* smax, const double* ssum,
                           double* sacc, const double* A, bool fuse_phi, hipStream_t st);
void launch_softmax_reduce(int N, int chunks, const double* sacc, bool fuse_phi,
                           const double* B, const float* phi_used, double sigma2x,
                           double* phiden, double* red, hipStream_t st, bool poisson = false,
                           const unsigned char* skip = nullptr);
void launch_total_respons(const float* P, size_t n, double* red, hipStream_t st);
void launch_scaling_finish(int N, const double* B, const double* phiden, float* scaling,
                           float* invphi, hipStream_t st);
void launch_reciprocal(const float* s, int N, float* inv, hipStream_t st);
// out[0] = sum of v[0..n) in a fixed order (one workgroup)
void launch_sum_f64(const double* v, int n, double* out, hipStream_t st);
// model helpers (k_model.hip)
void launch_rowsum(const float* P, int R, int N, const float* invphi, double* T, float* Tf,
                   float* rowscale, bool want_scale, hipStream_t st, const float* phi = nullptr,
                   float* Tphi = nullptr, int* nnz = nullptr);
// den[i] = sum_r P[r][i] A[r] (Poisson scaling denominator)
void launch_p_dot_a(const float* P, int R, int N, const double* A, double* den, hipStream_t st);
// ins_k: compact indices of the inserted pixels (mask == 1) in 8 x 8 tiles
void launch_insert_atomic(const float* Sc, int n_mp, const int* pix, const int* ins_k,
                          const float* wT, const float* rot, int R, int S, const float* x,
                          const float* y, const float* z, float* model, float* weight,
                          int interp, hipStream_t st);
void launch_build_map(const int* ins_pix, int n_ins, const float* rot, int R, int S,
                      const float* x, const float* y, const float* z, int interp,
                      Engine& eng, hipStream_t st);
void launch_insert_gather(const float* Sc, int n_mp, const int* ins_col, int n_ins,
                          const float* wT, const uint32_t* off, const uint32_t* ent,
                          const float* ent_wt, int S, int interp, float* model, float* weight,
                          hipStream_t st);
void launch_set_wT(const float* wts, const float* Tf, int R, float* wT, hipStream_t st);
// brick-owned insertion (k_model.hip), LDS atomics or the deterministic
// sequential order; the model side must be a multiple of the brick edge
// (8 atomic, 4 deterministic)
bool insert_brick_fits(int S, int R, bool deterministic);
void launch_insert_brick(const float* Sc, int n_mp, const int* kidx, const int* mask, const float* wT,
                         const float* rot, int R, int S, const float* x, const float* y, const float* z,
                         float* model, float* weight, int interp, bool deterministic, hipStream_t st);
// Refinement of the significant responsibilities (k_refine.hip).  Sources:
// [rows][ld] blocks of responsibilities (row -> rotation row_rot[row] or
// row, column -> pattern col_pat[col] or col, patterns with skip set left
// out; own: the refined value is also written back there) and the
// candidate slots Lc [N][topk][slots] of the patterns outside insub.
struct RefineRows {
  float* P = nullptr;
  int rows = 0, ld = 0, cols = 0;
  const int* row_rot = nullptr;
  const int* col_pat = nullptr;
  const unsigned char* skip = nullptr;
  bool own = false;
};
struct RefineInputs {
  float* respons = nullptr;  // [R][N] responsibility matrix (every refined entry written)
  int N = 0, R = 0, N2d = 0;
  const float* slices = nullptr;  // [R][N2d] FP32 slices the likelihood used
  const float* images = nullptr;  // [N][N2d] counts
  RefineRows rows[3];
  int nrows = 0;
  float* Lc = nullptr;
  const int *cand = nullptr, *cell_off = nullptr, *cell_rot = nullptr;
  const unsigned char* insub = nullptr;
  int topk = 0, slots = 0;
  double* phiden = nullptr;  // [N] scaling denominators to correct, or nullptr
  bool poisson = false;
  double sigma2x = 0.0;      // 2 n_m sigma^2 (Gaussian)
  float tau = 1e-9f;
};
// Recompute the entries >= in.tau; *dtotal receives the change of sum P log P.
bool refine_posterior(const RefineInputs& in, double* dtotal);

// red: kNormMaxBlocks partial sums, then the factor (kNormMaxBlocks + 2
// doubles); the partial-sum grid never exceeds kNormMaxBlocks blocks.
constexpr int kNormMaxBlocks = 1024;
// Device address of the normalisation scratch of the allocation-free
// entry point on `device` (noalloc.hip), resolved on first use; nullptr
// and an error set on failure.
double* noalloc_scratch(int device);
void launch_normalize_model(float* model, const float* weight, int n, double* red,
                            hipStream_t st);
void launch_model_max(const float* model, int n, float* out, hipStream_t st);
void launch_fill(float* p, size_t n, float v, hipStream_t st);
// CSR counts and the sparse Poisson E-step (k_sparse.hip)
void launch_csr_row_nnz(const float* images, const int* pix, int N, int N2d, int n_mp, int* nnz, hipStream_t st);
void launch_csr_fill(const float* images, const int* pix, int N, int N2d, int n_mp, const int* off, int* col,
                     unsigned short* cnt, hipStream_t st);
void launch_transpose_hilo(const _Float16* hi, const _Float16* lo, int R_pad, int n_mp, void* out, hipStream_t st);
void launch_estep_csr(const int* off, const int* col, const unsigned short* cnt, const void* LSt, int R_pad,
                      const LikeEpilogue& ep, hipStream_t st);
// M-step over the nonzero responsibilities (k_sparse.hip): lists of each
// rotation's nonzero P (patterns ascending), then S' from them
void launch_p_row_fill(const float* P, int R, int N, const float* opscale, const int* off, int* col, float* val,
                       hipStream_t st);
bool launch_mstep_sparse(const std::vector<int>& h_off, const int* off, const int* col, const float* val,
                         const _Float16* K, const _Float16* Klo, const MstepArgs& a, hipStream_t st);
void launch_images_from_csr(const int* off, const int* pix, const unsigned short* cnt, int N, int N2d, float* images,
                            hipStream_t st);

}  // namespace emc
