/*
 * emc_hip.h -- additions to the EMC C-ABI that the reference does not have.
 *
 * None of these has a counterpart in FilipeMaia/emc's emc.h; they add what a
 * production caller needs around the reference entry points (error codes,
 * freeing, synchronisation, option selection, phase timing) without changing
 * the signatures of those entry points.
 *
 * Threading: the library keeps one HIP stream on the current device and is
 * not thread safe.  All entry points of emc.h and this header enqueue work on
 * that stream.
 */
#ifndef EMC_HIP_H
#define EMC_HIP_H

#include <stddef.h>

#include "emc.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Status of the last failed call: returns 0 if no error has occurred since
 * the last emc_hip_clear_error(), else a non-zero code, and copies the
 * message into buf (n bytes, may be NULL).  Errors are sticky: after one,
 * entry points return without doing work (and print the message once to
 * stderr) until emc_hip_clear_error() is called.  With the environment
 * variable EMC_HIP_ABORT_ON_ERROR=1 the process aborts at the first error. */
int emc_hip_last_error(char *buf, size_t n);
void emc_hip_clear_error(void);

/* Free a device buffer returned by any cuda_allocate_* function. */
void emc_hip_free(void *d_ptr);
/* Wait for all work enqueued by the library. */
void emc_hip_synchronize(void);
/* Release every internal buffer and the stream (optional, at exit). */
void emc_hip_shutdown(void);

/* sum over n entries of P log P: the size_t form of cuda_total_respons, for
 * N_images * N_slices >= 2^31 (cuda_total_respons also accepts the wrapped
 * int the reference's driver passes when it matches the buffer's shape). */
real emc_hip_total_respons_n(const real *d_respons, size_t n);

/* cuda_normalize_model (emc.h; reference emc_cuda.cu:414-433) that
 * allocates nothing, for callers that cannot call hipMalloc (code under
 * stream capture, tight serving loops).  No counterpart in the reference.
 * Same kernels and result, bit for bit; d_model and d_weight hold
 * model_size voxels.  Its scratch (1,024 FP64 partial sums and the factor,
 * 8 KB) is a static device array of the library, one per GPU, held until
 * the process ends: any model_size from 0 to INT_MAX fits; a negative size
 * or a null pointer is rejected (-1, error set).  Calls on one GPU share
 * that scratch, so calls that overlap (other host threads, or streams other
 * than the library's) race.  Returns 0 on success. */
int emc_hip_normalize_model_noalloc(real *d_model, const real *d_weight, int model_size);

/* Per-pattern best rotation (argmax_r P, lowest r on ties) into best[N]
 * and the responsibility of every rotation averaged over the patterns into
 * average[R] (either may be NULL), from d_respons [R][N]: what the
 * reference's viewer reads as best_quaternion_%04d.data and
 * average_resp_%04d.h5 (viewer/slice_module.py:42,
 * viewer/rotations_module.py:43-47).  No counterpart in the reference's
 * emc.h.  Returns 0 on success. */
int emc_hip_respons_summary(const real *d_respons, int N_slices, int N_images, int *best, float *average);

/* Device buffer of `bytes` bytes (the size_t form of cuda_allocate_real). */
void emc_hip_allocate_bytes(void **d_ptr, size_t bytes);
/* Page-locked host buffer (no counterpart in the reference): host arrays
 * the library fills every iteration, such as the model array
 * cuda_update_slices copies back (emc_cuda.cu:352), are copied at the
 * link's rate instead of through the runtime's staging of pageable memory.
 * Returns 0 on success; free with emc_hip_host_free. */
int emc_hip_host_alloc(void **h_ptr, size_t bytes);
void emc_hip_host_free(void *h_ptr);
/* Copies between host and device on the library stream; they return after
 * the copy is complete.  Return 0 on success. */
int emc_hip_memcpy_to_host(void *dst, const void *d_src, size_t bytes);
int emc_hip_memcpy_to_device(void *d_dst, const void *src, size_t bytes);

/* Options (key, value), all strings; returns 0 if accepted.
 *   "precision"  "fp32" (default): the reference's method, FP32 arithmetic
 *                 on FP32 matrix cores;
 *                "fp16": the pattern x slice and M-step contractions on FP16
 *                 matrix cores with hi/lo operand splitting and FP32
 *                 accumulation, FP64 normalisations.
 *   "insert"     "atomic" (default): each workgroup owns an 8^3 brick of
 *                 the model (and a share of the rotations with T_r > 1e-10
 *                 when there are too few bricks to fill the GPU) and adds
 *                 into LDS every contribution that lands in it, found from
 *                 the brick's box mapped back onto the detector, then adds
 *                 the brick to the model: plain stores when it holds all
 *                 the rotations, one global atomic per voxel otherwise (LDS
 *                 float atomics, so the order of the additions is not
 *                 fixed); per-pixel global atomics where the model side is
 *                 not a multiple of 8;
 *                "global_atomic": per-pixel global float atomics;
 *                "deterministic": every voxel's contributions added in the
 *                 order of the reference's sequential insertion (rotations
 *                 ascending, pixels x-major), so bit-reproducible run to run
 *                 and equal to that insertion: by "deterministic_brick"
 *                 or "sorted_scatter", whichever the number of rotations
 *                 with T_r > 1e-10 makes cheaper (the brick form needs a
 *                 side that is a multiple of 4) -- the two give the same
 *                 bits;
 *                "deterministic_brick": 4^3 bricks, one wave each, taking
 *                 the active rotations in order; colliding pixels add one
 *                 at a time in pixel order (work follows the active
 *                 rotations; side a multiple of 4 and fewer than 2^24
 *                 rotations, else "sorted_scatter" is used);
 *                "sorted_scatter": the same sums from a voxel-sorted map of
 *                 every (rotation, pixel) entry, built once per rotation
 *                 table (memory: 4 bytes per entry, 36 trilinear; work
 *                 fixed by the table).
 *   "interp"     "nearest" (default, the reference's rule) or "trilinear".
 *   "timing"     "1" records per-phase GPU time (an event pair per phase on
 *                 the library's stream, read back when the times are
 *                 queried: no synchronisation inside an entry point), "0"
 *                 off.
 *   "verbose"    "1" prints the reference's per-call diagnostics.
 *   "likelihood" "gaussian" (default, the reference's); "poisson":
 *                 L = sum_p m (K log(phi max(S, 1e-6)) - phi S), scaling
 *                 phi = sum K / sum_r P sum S, S' = sum_i P K / sum_i P phi.
 *   "sparse"     "on" (Poisson, precision "fp16"): the whole-table E-step
 *                 reads the counts as CSR (int32 pixel, uint16 count) and
 *                 contracts them with the log-slice operand; "off" (default).
 *   "mstep"      FP16 form only.  "auto" (default): when at most 1 in 200
 *                 responsibilities is nonzero, the M-step sums each rotation's
 *                 nonzero P_ri (lists built from d_respons, P kept in FP32)
 *                 instead of the dense contraction; "dense", "sparse": always
 *                 that form.  The dropped terms are exact zeros.
 *   "expand_pack"  "auto" (default): the fused expand + FP16 pack keeps a
 *                 slice of at most 64 KB in LDS, larger slices run in blocks
 *                 of 8 rotations x 64 x 64 pixel tiles with provisional row
 *                 scales; "tiled": the latter at every size.
 *   "expand_pixels"  "rows" (default): in the fused expand + pack each
 *                 wave takes 64 consecutive pixels of a row (full 256-byte
 *                 slice writes); "tiles": an 8 x 8 pixel tile (fewer model
 *                 cache lines per wave; same results, measured 8% slower at
 *                 128^3).
 *   "expand_order"  "normal" (default): the expand kernels take the
 *                 rotations in the order of a Hilbert curve over the
 *                 direction of their slice normals, so that the slices in
 *                 flight gather from the same model region; "table": in
 *                 table order.  Results are the same either way (each
 *                 rotation's row is its own); the order is computed once
 *                 per rotation buffer (address and count).
 *   "expand_rotations"  rotations per workgroup of the tiled expand + pack
 *                 (default "8"; 1 .. 64).
 *   "refine"     "on" (default): after the softmax, every responsibility of
 *                 at least "refine_tau" (default "1e-9") is recomputed from
 *                 its likelihood summed in FP64 over the pattern's nonzero
 *                 counts (the matrix cores' FP32 sums over the pixels and
 *                 the FP32 storage of L are off by up to ~5e-3 at 15,000
 *                 pixels, a relative error of ~5e-3 in every P); each
 *                 pattern's refined entries are renormalised among
 *                 themselves, the others keep their values.  A pattern with
 *                 more than 1,024 entries at the threshold refines those
 *                 >= 1e-6 instead, or those >= 1e-3.  The threshold
 *                 reaches down to 1e-9 because voxels whose whole weight
 *                 comes from small responsibilities take their value from
 *                 those alone (M / W): at 1e-6 such voxels of a 256^3
 *                 model still differ from the FP64 iteration by ~2e-4 of
 *                 the model's maximum, at 1e-9 by ~3e-5; "off".
 *   "gemm_tile"  "256" (default): 256 x 256 tiles of the FP16 likelihood
 *                 GEMM, one 16-wave workgroup per compute unit, three LDS
 *                 stages; "256w8": the same tiles with 8 waves; "128":
 *                 128 x 128 tiles, two workgroups per compute unit;
 *                 "256x128": 256 x 128 tiles with four LDS stages (same
 *                 results, bit for bit).
 *   "gemm_tail"  "split" (default): the FP16 likelihood GEMM's tiles that
 *                 do not fill a whole round of the compute units (all of
 *                 them when there are fewer tiles than units) are split along
 *                 the pixel sum into as many ranges as finish them soonest
 *                 (partial sums added in a fixed order: results are
 *                 reproducible, and equal the unsplit ones to FP32
 *                 rounding); "whole": every tile in one workgroup.
 *   "respons_fill"  proposal mode.  "delta" (default): when d_respons still
 *                 holds the previous proposal iteration's responsibilities,
 *                 as this library wrote them and its M-step listed their
 *                 nonzeros, an iteration zeroes only those entries before
 *                 writing its own, instead of all N_images * N_slices.  A
 *                 caller that writes d_respons by other means than this
 *                 library must call emc_hip_invalidate() (or choose "full")
 *                 before the next iteration.  "full": every iteration writes
 *                 the whole matrix.  The contents are the same either way.
 *   "proposal"   "on": once candidates exist for this pattern count (see
 *                 the proposal workflow below) and precision is "fp16", the
 *                 E-step scores each pattern only against the fine rotations
 *                 of its candidate coarse cells; "off" (default).
 *   "audit_fraction"  fraction of patterns (default "0.02") scored on the
 *                 whole table every proposal iteration, chosen at random.
 *   "min_mass"   captured posterior mass (default "0.999") below which a
 *                 pattern returns to whole-table search.
 *   "audit_max_fail"  fraction of the audited patterns (default "0.005")
 *                 allowed below min_mass: the subset is scored before the
 *                 candidates, and when more of it falls below, the iteration
 *                 runs on the whole table instead ("1": never).
 *   "proposal_track" "on" (default): before every proposal iteration each
 *                 pattern's candidates are re-centred on its last posterior
 *                 (the cells holding most of its mass and their nearest
 *                 cells, then the network's cells; with topk <= 64);
 *                 "off": the network's cells throughout.
 *   "proposal_shared" largest number of shared cells (default "512"; "0":
 *                 none).  With tracking, the cells holding some pattern's
 *                 best cell (or >= 1% of its posterior) in the last iteration
 *                 are shared when there are at most this many: every pattern
 *                 is scored on all their rotations (dense kernel) besides its
 *                 topk candidate cells, which then exclude them.
 *   "proposal_group_min" smallest group of patterns (sharing their first
 *                 candidate cell) whose common candidate cells are scored as
 *                 one dense block (default "64"; "0": every candidate cell
 *                 by the per-cell kernel).  Takes effect when the candidate
 *                 lists are next built.
 *   "proposal_core_div"  a candidate cell is in its group's dense block when
 *                 at least 1 / this many of the group's patterns hold it
 *                 (default "4"; 1 .. 64).  Takes effect when the candidate
 *                 lists are next built.
 *   "proposal_gate"  "predicted" (default): a pattern also starts in
 *                 whole-table search when the network's top-k cells hold less
 *                 than min_mass of its predicted distribution; "audit": only
 *                 whole-table scoring moves patterns to full search.
 */
int emc_hip_set_option(const char *key, const char *value);

/* Invalidate every cached derivative of caller buffers (packed patterns,
 * mask index, insertion map, the record of which responsibilities are
 * nonzero).  Needed only if the caller writes images, the mask, rotations,
 * coordinates or responsibilities by means other than this library. */
void emc_hip_invalidate(void);

/* Per-phase GPU time accumulated while "timing" is on, in milliseconds, and
 * the number of kernel launches per phase.  Phases:
 *   0 expand, 1 pack (slice operands), 2 likelihood GEMM, 3 softmax,
 *   4 scaling, 5 M-step row sums, 6 M-step GEMM, 7 insert, 8 normalise,
 *   9 other, 10 audit (proposal mode: everything done for the patterns
 *   scored on the whole table), 11 shared (proposal mode: the likelihood of
 *   every pattern on the shared cells' rotations), 12 groups (proposal mode:
 *   the dense blocks of pattern groups; phase 2 is then the per-cell kernel
 *   for the other candidate entries), 13 audit_gemm (proposal mode: the
 *   whole-table likelihood of the subset; its other work is phase 10), 14
 *   refine (the FP64 recomputation of the significant responsibilities).
 * Returns the number of phases written (<= n). */
int emc_hip_phase_times(double *ms, long long *launches, int n);
void emc_hip_reset_phase_times(void);
/* Name of phase k, or NULL. */
const char *emc_hip_phase_name(int k);

/* ---- Rotation-proposal network (no counterpart in the reference) ----
 * A small CNN maps a pattern (log1p of its counts on the detector grid,
 * unmasked pixels 0) to a distribution over coarse rotation cells: four
 * 3x3 / stride-2 convolutions (16, 32, 64, 64 channels, ReLU), dense 256 +
 * ReLU, dense n_cells, softmax.  BF16 operands, FP32 accumulation, FP32
 * master weights, Adam.  These calls expose it for testing; the proposal
 * workflow below drives it. */
int emc_hip_cnn_init(int side, int n_cells, int batch, unsigned long long seed);
long emc_hip_cnn_param_count(void);
/* Copy the parameters and/or the last gradients (either pointer may be NULL). */
int emc_hip_cnn_get(float *params, float *grads);
int emc_hip_cnn_set(const float *params);
/* Logits of patterns idx[0..B) of d_images, written [n_cells][B]. */
int emc_hip_cnn_forward(const real *d_images, const int *d_mask, int N_2d, const int *idx, int B,
                        float *logits);
/* One step of softmax cross-entropy training against targets [n_cells][B]
 * (rows sum to 1): forward, backward, Adam with learning rate lr (0 keeps
 * the weights, the gradients are still computed).  *loss = mean loss. */
int emc_hip_cnn_step(const real *d_images, const int *d_mask, int N_2d, const int *idx, int B,
                     const float *targets, float lr, double *loss);

/* ---- Proposal workflow (no counterpart in the reference) ----
 * Coarse cells: the antipodal half of the sampling-n_coarse table; every
 * fine rotation belongs to the cell of the nearest coarse rotation (at most
 * 32 per cell).  Typical use, with the emc.h calls unchanged:
 *   emc_hip_proposal_init(q, R, 4, side, 32, 256, seed);
 *   iterations 1-5 (dense), each followed by
 *     emc_hip_proposal_train(..., d_respons, ...)   (P summed per cell as targets)
 *   emc_hip_proposal_predict(...);  emc_hip_set_option("proposal", "on");
 *   further iterations score topk cells per pattern (+ the audit subset). */
/* Configure for the fine table q_fine [n_rot][4] (host); (re)initialises the
 * network for side x side patterns and the given batch. */
int emc_hip_proposal_init(const real *q_fine, int n_rot, int n_coarse, int side, int topk, int batch,
                          unsigned long long seed);
/* Cells as CSR: cell_off [C+1], cell_rot [n_rot] (either may be NULL);
 * returns C, or -1 if not configured. */
int emc_hip_proposal_cells(int *cell_off, int *cell_rot);
/* `epochs` passes of Adam (learning rate lr) over all patterns against the
 * per-cell sums of the responsibilities d_respons [N_slices][N_images];
 * *loss = mean cross-entropy of the last pass. */
int emc_hip_proposal_train(const real *d_images, const int *d_mask, int N_2d, const real *d_respons, int N_images,
                           int N_slices, int epochs, float lr, double *loss);
/* Candidate cells of every pattern from the network (its topk largest
 * outputs); resets the search modes. */
int emc_hip_proposal_predict(const real *d_images, const int *d_mask, int N_2d, int N_images);
/* Candidate cells given by the caller, cand [N_images][topk] (distinct per
 * pattern); resets the search modes. */
int emc_hip_proposal_set_candidates(const int *cand, int N_images, int topk);
/* Current candidates [N][topk], search modes (1 = whole table) and the
 * network's top-k mass (1 for caller-given candidates); NULL skips. */
int emc_hip_proposal_get(int *cand, int *mode, float *mass, int N_images);
/* Shared cells of the current proposal iteration (cells every pattern is
 * scored on, besides its candidate cells): their count; the first max_cells
 * written to cells (NULL skips). */
int emc_hip_proposal_shared(int *cells, int max_cells);
/* Statistics of the last proposal iteration: out[0] iteration, [1] patterns
 * in whole-table search, [2] audited patterns, [3] audited below min_mass,
 * [4] smallest and [5] mean captured mass of the audited patterns, [6]
 * pattern x rotation pairs scored, [7] patterns gated to full search at
 * prediction, [8] last training loss, [9] last training time (ms), [10] last
 * prediction time (ms), [11] patterns scored on the whole table, [12]
 * candidate pairs of all patterns (the candidate E-step's work), [13]
 * patterns whose candidates were re-centred before the iteration, [14] 1 if
 * the audit sent the iteration to the whole table, [15] shared cells, [16]
 * host time of the re-centring and list building (ms), [17] (pattern,
 * rotation) pairs of the group blocks, [18] candidate pairs left to the
 * per-cell kernel, [19] shared rotations, [20] responsibility entries the
 * last iteration cleared before writing its own (N_images * N_slices, or the
 * previous iteration's nonzeros with "respons_fill" "delta"), [21] cells
 * that qualified for a group's core beyond its cap of 64 (scored by the
 * per-cell kernel instead).  Returns the number of values
 * written (<= n). */
int emc_hip_proposal_stats(double *out, int n);

/* Patterns given as CSR (no counterpart in the reference): pattern i holds
 * counts[j] photons at pixel pix[j] (p = y * side + x) for row_off[i] <= j <
 * row_off[i+1]; every other pixel 0.  Fills the dense [N_images][N_2d] buffer
 * d_images the emc.h calls take (so a caller uploads only the nonzeros).
 * Returns 0 on success. */
int emc_hip_images_from_csr(real *d_images, const int *row_off, const int *pix, const unsigned short *counts,
                            int N_images, int N_2d);

/* Information on the packed pattern operand of the last E-step:
 * out[0] = masked pixel count n_m, out[1] = padded count, out[2] = 1 if the
 * FP16 operand holds the counts exactly, out[3] = GEMM passes per product. */
int emc_hip_pack_info(int *out, int n);

/* Nonzero responsibilities the last dense-form M-step summed over (no
 * counterpart in the reference): >= 0 when it ran over the per-rotation lists
 * of nonzero P (option "mstep"), -1 when it ran as the dense contraction. */
long long emc_hip_mstep_nnz(void);

/* Responsibilities the last refinement recomputed in FP64 (option
 * "refine"; no counterpart in the reference); 0 when it is off. */
long long emc_hip_refine_count(void);

#ifdef __cplusplus
}
#endif

#endif /* EMC_HIP_H */
