// State of the learned rotation proposal and the candidate-restricted
// iteration.
#pragma once

#include <cstdint>
#include <vector>

#include "cnn.hpp"
#include "engine.hpp"

namespace emc {

struct Proposal {
  Cnn cnn;
  bool configured = false;
  bool predicted = false;  // candidates present
  bool active = false;     // option "proposal" on
  int C = 0;               // coarse cells
  int R = 0;               // fine rotations
  int N = 0;               // patterns of the candidate lists
  int topk = 32;
  int batch = 256;         // network batch
  float audit_frac = 0.02f;
  float min_mass = 0.999f;
  // fraction of the audited patterns below min_mass above which the whole
  // iteration runs on the whole table (1 = never)
  float max_fail = 0.005f;
  bool gate_predicted = true;  // full search when the network's top-k mass < min_mass
  bool track = true;       // re-centre the candidates on the last posterior
  uint64_t seed = 0;
  int iteration = 0;       // candidate-restricted E-steps run

  // coarse cells: fine rotations of each cell (CSR), nearest cells of each
  std::vector<int> h_cell_off, h_cell_rot, h_nbr;  // set-up only
  static constexpr int kNbr = 12;
  static constexpr int kTop = 4;     // posterior cells per pattern used to re-centre
  DBuf d_net_cand;                   // the network's candidates [N][topk]
  DBuf d_nbr;                        // [C][kNbr] nearest cells (-1: none)
  DBuf d_mark, d_shared;             // [C] support marks, shared cells (bytes)
  DBuf d_count, d_ncand, d_moved;    // entries per cell [C+1], rotations per pattern [N], changed lists
  DBuf d_keys_s, d_vals_s, d_sort_tmp;  // cell-sorted entries and the sort's scratch
  HostVec<int> h_count, h_colmap;  // entries per cell; subset index of each pattern (-1: none)
  DBuf d_colmap;
  HostVec<unsigned char> h_mark;
  DBuf d_top, d_topm;                // [N][kTop] cells, masses of the last posterior
  // where the last posterior is: none, Lc / Lsub (candidate iteration), or
  // already reduced into d_top / d_topm (whole-table iteration of post_N patterns)
  enum { POST_NONE = 0, POST_CAND = 1, POST_TOP = 2 };
  int post = POST_NONE;
  int post_N = 0;
  // shared cells: the cells holding a pattern's best cell (or >= kShareMass
  // of some pattern's posterior) in the last posterior, when there are at
  // most max_shared of them.  Every pattern is scored on all their rotations
  // by the dense kernel, and the candidate cells exclude them.
  int max_shared = 512;
  static constexpr float kShareMass = 0.01f;
  HostVec<int> h_sh_cells, h_sh_off, h_sh_rot;  // cells, CSR of their rotations
  HostVec<unsigned char> h_shared;              // [C]
  int n_shc = 0, R_sh = 0, R_sh_pad = 0;
  DBuf d_sh_cells, d_sh_off, d_sh_rot, d_Ssh, d_Sshlo, d_Ash, d_sscale_sh, d_Lsh, d_top_s, d_topm_s;
  void clear_shared() { h_sh_cells.clear(); h_sh_off.assign(1, 0); h_sh_rot.clear(); n_shc = R_sh = R_sh_pad = 0; }
  DBuf d_iota;                       // 0..max(N, R_sh)-1
  DBuf d_stage;                      // staging of the batched list uploads
  DBuf d_part_c, d_part_m;           // top-cell partials of the subset columns
  DBuf d_cell_off, d_cell_rot;
  // per pattern: candidate rotation count, network top-k mass, search mode
  // (the candidate cells [N][topk] live in d_cand)
  HostVec<int> h_ncand;
  HostVec<float> h_pmass;
  HostVec<unsigned char> h_mode;  // 1 = whole table
  DBuf d_cand, d_pmass;
  // entries per cell (CSR, patterns ascending) and E-step work items
  HostVec<int> h_ent_off;
  HostVec<long> h_chunk_off;
  DBuf d_ent_off, d_ent_pat, d_ent_j, d_work, d_chunk_off, d_chunk_cell;
  int n_work = 0;
  long n_entries = 0, n_chunks = 0;
  // M-step segments (cells with many entries are split) and their partials
  DBuf d_seg_cell, d_seg_chunk0, d_seg_nch, d_seg_part, d_mcell, d_mcell_part, d_part;
  DBuf d_part_T, d_part_Tphi, d_part_max;
  int n_seg = 0, n_mcell = 0, n_parts = 0;
  // pattern groups of the candidate E-step: patterns sharing their first
  // candidate cell, whose core cells (candidates of at least half of them)
  // are scored as dense blocks; the per-cell E-step takes the other entries
  // (core cells beyond kMaxCore included: st_core_trunc counts them)
  int group_min = 64;                // smallest group scored as blocks (0: no groups)
  int core_div = 4;                  // core cell: held by >= 1 / core_div of its group's patterns
  DBuf d_gkey, d_gcount, d_gpat, d_gkey_s, d_ghist, d_gcore, d_gncore, d_gnrot, d_gcofs, d_crot, d_ccell, d_cslot,
      d_gtrunc, d_rank, d_rkey, d_gwork, d_rcount, d_rncand;
  HostVec<int> h_gcount, h_gncore, h_gnrot, h_gcofs, h_rcount, h_rent_off, h_rn;
  HostVec<float> h_mass;
  HostVec<int4> h_gwork;
  int n_gwork = 0;
  double st_group_pairs = 0;         // (pattern, rotation) pairs of the group blocks
  double st_resid_pairs = 0;         // candidate pairs left to the per-cell kernel
  double st_core_trunc = 0;          // qualifying group-core cells beyond kMaxCore (scored per cell)
  int h_gtrunc = 0;
  double st_cleared = 0;             // responsibility entries cleared by the last write (N * R unless delta)
  DBuf d_rent_off, d_rent_pat, d_rent_j, d_rwork, d_rkeys_s, d_rvals_s;
  int n_rwork = 0;
  DBuf d_targets;  // float [N][C] training targets (coarse-cell posterior mass)
  DBuf d_perm;     // int [N] shuffled pattern order of a training epoch
  DBuf d_Lc;      // float [N][topk][32]: L, then P
  DBuf d_panels;  // half [chunks][2][32][32] M-step operand
  DBuf d_crowscale;
  DBuf d_plogp;   // double [N]
  DBuf d_zero;    // zero halves (gather target of padding rows)
  // whole-table subset of this iteration (audited + full search)
  HostVec<int> h_sub;
  HostVec<unsigned char> h_insub, h_audit;
  int n_sub = 0, n_sub_pad = 0;
  DBuf d_sub, d_insub, d_Ksub, d_Ksublo, d_Ktsub, d_Ktsublo, d_Bsub, d_phisub, d_invphisub, d_Lsub, d_phiden_sub,
      d_Psub16, d_Psub16lo, d_mass, d_tot;
  // state shared between the calls of one iteration
  const float* respons = nullptr;  // buffer of the current proposal responsibilities
  bool have_p = false;             // softmax done, M-step may run on the candidates
  double sub_total = 0;            // sum P log P over the whole-table subset
  // statistics of the last iteration
  double st_full = 0, st_audited = 0, st_fail = 0, st_min_mass = 1, st_mean_mass = 1, st_pairs = 0,
         st_gated = 0, st_train_loss = 0, st_train_ms = 0, st_predict_ms = 0, st_cand_pairs = 0,
         st_moved = 0, st_dense = 0, st_shared = 0, st_recentre_ms = 0;

  void release();
};

Proposal& proposal();

// Candidate-restricted steps, called by the emc.h entry points when the
// proposal is active for (N, R).  Packed patterns and slices are current.
bool proposal_applies(int N, int R);
// Scores the whole-table subset first; when more than max_fail of the
// audited patterns fall below min_mass, *dense is set and the caller runs
// this iteration on the whole table instead of the candidates.
bool proposal_estep(int N, int R, float sigma, const float* phi_used, bool* dense);
bool proposal_softmax(float* respons, int N, int R, double sigma2x, double* total);
bool proposal_mstep(int N, int R, float* slices, int N2d, const float* scaling);
// Whole-table responsibilities [R][N] of an iteration run without the
// candidates (a fallback iteration, or a dense one before prediction):
// reduced to each pattern's top cells for re-centring, when tracking.
bool proposal_tracks(int N, int R);
bool proposal_has_shared();
// where the last proposal softmax left each part of the posterior
struct NzSources;
bool proposal_nz_sources(NzSources& src, int N, int R);
bool proposal_dense_posterior(const float* respons, int N, int R);

}  // namespace emc
