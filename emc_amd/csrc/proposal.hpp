// State of the learned rotation proposal.
#pragma once

#include <cstdint>
#include <vector>

#include "cnn.hpp"
#include "engine.hpp"

namespace emc {

struct Proposal {
  Cnn cnn;
  bool configured = false;
  bool predicted = false;
  int C = 0;           // coarse cells
  int R = 0;           // fine rotations
  int N = 0;           // patterns (set by predict)
  int topk = 32;
  int slot = 32;       // per-cell rotation slots (>= largest cell, multiple of 16)
  float audit_frac = 0.02f;
  float min_mass = 0.999f;
  uint64_t seed = 0;
  int iteration = 0;   // candidate-restricted iterations run

  // coarse cells: fine rotations of each cell (CSR), cell of each fine rotation
  std::vector<int> h_cell_off, h_cell_rot;
  DBuf d_cell_off, d_cell_rot, d_rot_cell, d_rot_slot;
  // per pattern: top-k cells (row-major [N][topk]) and search mode (1 = full)
  std::vector<int> h_cand, h_mode;
  DBuf d_cand;
  // work lists of the candidate-restricted contractions, rebuilt when the
  // candidates or the modes change
  bool lists_valid = false;
  std::vector<int> h_ent_off;        // [C+1] entries per cell
  DBuf d_ent_off, d_ent_pat, d_ent_j;  // entry -> pattern, position in its top-k
  DBuf d_work;                       // int2 (cell, first entry) per E-step workgroup
  int n_work = 0;
  long n_entries = 0;
  // candidate log-likelihoods -> responsibilities [N][topk][slot]
  DBuf d_Lc;
  // patterns scored on the whole table this iteration (audit + full search)
  std::vector<int> h_sub;
  int n_sub = 0, n_sub_pad = 0;
  std::vector<unsigned char> h_audit;  // pattern audited this iteration
  DBuf d_sub, d_sub_is_audit, d_Ksub, d_Ktsub, d_Bsub, d_phisub, d_Lsub, d_Psub16, d_Psub16lo, d_invphisub;
  DBuf d_mass;          // captured mass of audited patterns
  DBuf d_psum;          // per-pattern sum P log P (candidate patterns)
  DBuf d_zero_row;      // n_mp zero halves (gather target for padding rows)
  DBuf d_U;             // float [R_pad][n_mp] accumulators of the subset M-step
  bool have_estep = false;
  double last_total = 0.0;
  // statistics of the last iteration
  double st_full = 0, st_audited = 0, st_fail = 0, st_min_mass = 1, st_mean_mass = 1, st_pairs = 0,
         st_train_loss = 0;
};

Proposal& proposal();

}  // namespace emc
