// Launch arguments of the candidate-restricted iteration's kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "engine.hpp"

namespace emc {

constexpr int kSlot = 32;       // rotation slots per coarse cell (largest cell must fit)
constexpr int kCandRows = 128;  // entries per E-step workgroup
constexpr int kSegChunks = 24;  // 32-entry chunks per M-step segment
constexpr int kMaxTopk = 64;    // candidate cells per pattern at most
constexpr int kMaxCore = 64;    // core cells of a pattern group at most
constexpr int kGroupTile = 256; // patterns and core rotations per group-block workgroup

struct CandEstepArgs {
  const _Float16 *K, *Klo;          // [N_pad][n_mp] packed counts (lo only when split)
  const _Float16 *S, *Slo;          // [R_pad][n_mp] packed slices hi / lo
  const _Float16* zero;             // >= n_mp zero halves
  const int2* work;                 // (cell, first entry) per workgroup
  int n_work;
  const int *cell_off, *cell_rot;   // fine rotations of each cell (CSR)
  const int *ent_off, *ent_pat, *ent_j;  // entries of each cell (CSR)
  const double *A, *B;              // sum_p m S^2 [R], sum_p m K^2 [N]
  const float* phi;                 // [N] scaling used by the likelihood
  const float* sscale;              // [R] 1 / FP16 slice row scale
  float kscale_inv;
  double inv_2nms2;
  int n_mp, topk;
  int poisson;                      // Poisson epilogue (C over log S)
  float* Lc;                        // [N][topk][kSlot]
};

struct GroupEstepArgs {
  const _Float16 *K, *Klo, *S, *Slo, *zero;
  const int4* work;                 // (first pattern in gpat, patterns, first core rotation, rotations)
  int n_work;
  const int* gpat;                  // patterns sorted by group
  const int* crot;                  // core rotations of every group
  const unsigned char *ccell, *cslot;  // core index of each one's cell, its slot in the cell
  const signed char* rank;          // [N][kMaxCore] candidate rank of each core cell (-1: none)
  const double *A, *B;
  const float* phi;
  const float* sscale;
  float kscale_inv;
  double inv_2nms2;
  int n_mp, topk;
  int poisson;
  float* Lc;
};

struct CandSoftmaxArgs {
  float* Lc;                        // L in, P out
  const int* cand;                  // [N][topk]
  const int *cell_off, *cell_rot;
  const unsigned char* insub;       // [N] pattern scored on the whole table
  const double *A, *B;
  const float* phi;
  double sigma2x;                   // 2 n_m sigma^2
  int N, topk;
  int poisson;                      // scaling denominator sum_r P_ri A_r
  float* respons;                   // dense [R][N] (zeroed) or nullptr
  double *phiden, *plogp;           // [N]
  // with the shared rotations scored beside the candidates, the candidates
  // are one (max, sum) partial at index `slot` of [nmerge][N]:
  // mode 0: normalise over the candidates alone (phiden, plogp written);
  // mode 1: write the candidates' partial into smax / ssum;
  // mode 2: normalise with all merged partials, fused sums into sacc
  //         ([4][nmerge][N]: P log P, P, P A, P L) at `slot`
  int mode;
  float* smax;
  double *ssum, *sacc;
  int nmerge, slot;
};

struct CandPrepArgs {
  const float* Pc;                  // [N][topk][kSlot]
  const int *ent_off, *ent_pat, *ent_j, *cell_off, *cell_rot;
  const long* chunk_off;            // [C+1] first 32-entry chunk of each cell
  const int* chunk_cell;            // [chunks] cell of each chunk
  double *part_T, *part_Tphi;       // [chunks][kSlot] partial sums
  float* part_max;                  // [chunks][kSlot]
  const float* invphi;              // operand P * invphi (ones for Poisson)
  const float* phi;                 // Poisson: also Tphi_r = sum P phi, else nullptr
  const double* Tsub;               // [R] subset row sums or nullptr
  const float* Tphisub;             // [R] subset sum P phi (Poisson) or nullptr
  double* T;
  float *Tf, *rowscale;
  float* Tphi;                      // [R] (Poisson)
  _Float16* panels;                 // [chunks][2][kSlot][32]
  int C, topk;
};

struct CandMstepArgs {
  const _Float16 *K, *Klo;
  const _Float16* zero;
  const _Float16* panels;
  const long* chunk_off;
  // segments: cell, first chunk, chunk count, partial-sum slot (-1: the
  // cell's only segment, which finishes S' itself)
  const int* seg_cell;
  const long* seg_chunk0;
  const int *seg_nch, *seg_part;
  int n_seg;
  float* part;                      // [parts][kSlot][n_mp]
  const int *mcell, *mcell_part;    // cells split into segments, their parts (CSR)
  int n_mcell;
  const int *ent_off, *ent_pat, *cell_off, *cell_rot, *pix;
  const float *rowscale, *Tf;        // Tf: divisor of U (T_r, or Tphi_r for Poisson)
  float* Sc;                        // [R_pad][n_mp]: subset U in (if has_sub), S' out
  float* slices;                    // [R][N2d]
  float kscale_inv;
  int has_sub, C, n_ptiles, n_mp, n_m, N2d;
  int poisson;
};

void launch_cand_estep(bool ksplit, const CandEstepArgs& a, hipStream_t st);
void launch_cand_softmax(const CandSoftmaxArgs& a, hipStream_t st);
void launch_cand_prep(const CandPrepArgs& a, long n_chunks, hipStream_t st);
void launch_cand_mstep(bool ksplit, const CandMstepArgs& a, hipStream_t st);
void launch_gather_rows(const _Float16* K, const _Float16* Klo, const int* sub, int n_sub, int n_sub_pad, int n_mp,
                        _Float16* Ks, _Float16* Kslo, hipStream_t st);
void launch_gather_rows_t(const _Float16* K, const int* sub, int n_sub, int n_sub_pad, int n_mp, _Float16* Kt,
                          hipStream_t st);
void launch_gather_f64(const double* v, const int* sub, int n, double* out, hipStream_t st);
void launch_gather_f32(const float* v, const int* sub, int n, float* out, hipStream_t st);
// captured mass: candidate cells' rotations plus the R_sh shared rotations
void launch_sub_finish(const float* Psub, int n_sub, const int* sub, const double* phiden_sub, double* phiden,
                       const int* cand, int topk, const int* cell_off, const int* cell_rot, const int* sh_rot,
                       int R_sh, float* mass, hipStream_t st);
void launch_shared_scatter(const float* Psh, int R_sh, int N, const int* sh_rot, const unsigned char* insub,
                           float* respons, hipStream_t st);
void launch_merge_top(int N, int ktop, const unsigned char* insub, int* top, float* topm, const int* top_s,
                      const float* topm_s, const int* sh_cells, hipStream_t st);
// Cells per partial of the whole-table top cells (one wave per 64 patterns
// and kTopGroup cells; small groups give enough waves to fill the device).
constexpr int kTopGroup = 16;
// the ktop (<= 4) cells holding the most posterior mass of every pattern
// (Pc == nullptr: the subset columns only); part_c / part_m:
// n_sub * ceil(C / kTopGroup) * 4 scratch entries
void launch_topcells(const float* Pc, int N, int topk, const unsigned char* insub, const int* cand,
                     const float* Psub, int n_sub, const int* sub, int C, const int* cell_off, const int* cell_rot,
                     int ktop, int* top, float* topm, int* part_c, float* part_m, hipStream_t st);
// the subset's posterior Psub [R][n_sub] into its columns of respons
// (nonzero_only: only its nonzero entries, into a matrix already zero there)
void launch_sub_scatter(const float* Psub, int R, int n_sub, const int* sub, int N, float* respons, bool nonzero_only,
                        hipStream_t st);
// respons = 0 except the subset columns (colmap[i] >= 0: Psub column); N % 4 == 0, 16-byte aligned rows
void launch_fill_subset(float* respons, int R, int N, const int* colmap, const float* Psub, int n_sub,
                        hipStream_t st);

// candidate lists on the GPU (k_lists.hip)
// mark: ceil(C / 4) * 4 bytes
void launch_mark_support(const int* top, const float* topm, int N, int ktop, float share_mass, int C,
                         unsigned char* mark, hipStream_t st);
void launch_recentre_cand(const int* top, const float* topm, int N, int ktop, const int* nbr, int nnbr,
                          const int* net, const unsigned char* shared, int C, int topk, int* cand, int* moved,
                          hipStream_t st);
// count[C] = entries naming no valid cell (count: C + 1 ints); ncand[i]
void launch_cand_hist(const int* cand, int N, int topk, int C, const int* cell_off, int* count, int* ncand,
                      hipStream_t st);
// count[c], c < C, of sorted keys [n]
void launch_counts_sorted(const int* keys, long n, int C, int* count, hipStream_t st);
bool sort_entries(const int* cand, int N, int topk, int C, DBuf& keys_sorted, DBuf& vals_sorted, DBuf& temp,
                  int* ent_pat, int* ent_j, hipStream_t st);

// pattern groups of the candidate E-step (k_groups.hip)
bool build_group_keys(const int* cand, int N, int topk, int C, int* key, int* count, int* gpat, DBuf& key_sorted,
                      DBuf& temp, hipStream_t st);
// hist: C * C ints of scratch; core: [C][kMaxCore]; *trunc: qualifying core
// cells left out by the kMaxCore cap, over all groups
void launch_group_core(const int* cand, const int* key, const int* count, int N, int topk, int C, int min_size,
                       int core_div, const int* cell_off, int* hist, int* core, int* ncore, int* nrot, int* trunc,
                       hipStream_t st);
void launch_group_rots(const int* core, const int* ncore, const int* cofs, const int* cell_off, const int* cell_rot,
                       int C, int* crot, unsigned char* ccell, unsigned char* cslot, hipStream_t st);
void launch_group_rank(const int* cand, const int* key, int N, int topk, int C, const int* core, const int* ncore,
                       signed char* rank, int* rkey, hipStream_t st);
void launch_group_estep(bool ksplit, const GroupEstepArgs& a, hipStream_t st);

// nonzero responsibilities per rotation from the candidate slots, shared
// rows and subset columns (k_lists.hip); R * N < 2^32
struct NzSources {
  const float* Pc;                 // [N][topk][kSlot]
  const int *cand, *cell_off, *cell_rot;
  const unsigned char* insub;      // [N]
  int N, topk;
  const float* Psh;                // [R_sh][N]
  const int* sh_rot;
  int R_sh;
  const float* Psub;               // [R][n_sub]
  const int* sub;
  int R, n_sub;
};
void launch_nz_count(const NzSources& a, int* rownnz, hipStream_t st);
// off: exclusive prefix of the counts ([R+1], device); col / val: [nnz]
bool build_nz_lists(const NzSources& a, long nnz, const int* off, const float* opscale, const float* phi,
                    DBuf& keys, DBuf& keys_sorted, DBuf& vals, DBuf& vals_sorted, DBuf& tail, DBuf& temp, int* col,
                    float* val, double* T, float* Tf, float* Tphi, hipStream_t st);
// respons[keys[t]] = 0 for the n sorted keys r * N + i
void launch_zero_keys(const unsigned* keys, long n, float* respons, hipStream_t st);

}  // namespace emc
