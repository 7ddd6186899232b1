/*
 * CPU oracle for the EMC expand/maximize/compress iteration.
 *
 * TEST INFRASTRUCTURE ONLY.  Only the tests, the smoke check and the CPU
 * baseline leg of the benchmark may load this library, and only as the
 * checker.  The product (libemc_hip.so, the emc driver) never links it.
 *
 * Two restatements of the reference, FilipeMaia/emc:
 *
 *   r32_*  FP32, faithful to the reference's kernels: same float expressions,
 *          same operand order, same strided per-thread partial sums and
 *          block tree reductions (256 threads) as emc_cuda.cu / emc_atomic.cu,
 *          compiled without FMA contraction.  Where the reference leaves the
 *          order open (thrust::reduce, float atomics) a fixed sequential
 *          order is used and said so at the function.
 *
 *   o64_*  FP64 oracle of the same equations (all arithmetic in double; the
 *          nearest-voxel index is computed exactly as the reference does in
 *          float so that both oracles see the same geometry).  Also holds
 *          the options the reference lacks: trilinear interpolation and the
 *          Poisson likelihood.
 *
 * Host maths shared by both: the rotation table (emc.c:43-363), the detector
 * coordinates (emc.c:365-393).
 *
 * Layouts follow the reference ABI: slices[r*N2d+p], images[i*N2d+p],
 * respons[r*N+i], model[z*S*S+y*S+x], pixel p = y*S+x.
 */
#ifndef EMC_ORACLE_H
#define EMC_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- host maths ---- */
int o_n_to_samples(int n);
/* Full table, 20(n+5n^3) entries, in the reference's order and float bits. */
int o_rotation_list(int n, float *q, float *w);
/* Antipodally deduplicated half: first occurrence of each {q,-q} pair kept,
 * weight = w(q) + w(-q).  Returns the count (10(n+5n^3)). */
int o_rotation_list_half(int n, float *q, float *w);
/* Per-pixel coordinates, layout [p = y*side + x]. */
void o_coordinates(int side, float wavelength, float pixel_size, float detector_distance,
                   float *x, float *y, float *z);
/* Nearest voxel index along each axis, float maths as emc_cuda.cu:65-88. */
void o_voxel_index(const float *q, float xc, float yc, float zc, int S, int *vx, int *vy, int *vz);

/* ---- FP32 restatement of the reference path ---- */
void r32_expand(const float *model, int S, const float *q, int R,
                const float *x, const float *y, const float *z, float *slices);
void r32_likelihood(const float *slices, const float *images, const int *mask, float sigma,
                    const float *scaling, float *respons, int N2d, int N, int R);
void r32_softmax(float *respons, int R, int N);
double r32_total_respons(const float *respons, size_t n);
void r32_update_scaling(const float *images, const float *slices, const int *mask,
                        const float *respons, float *scaling, int N, int R, int N2d);
void r32_mstep(const float *images, float *slices, const int *mask, const float *respons,
               const float *scaling, int N, int R, int N2d, float *T);
void r32_insert_all(float *model, float *weight, int S, const float *slices, const int *mask,
                    const float *T, const float *q, const float *weights, int R, const float *x,
                    const float *y, const float *z);
void r32_insert_all_par(float *model, float *weight, int S, const float *slices, const int *mask,
                        const float *T, const float *q, const float *weights, int R, const float *x,
                        const float *y, const float *z);
/* M-step (slices overwritten at masked pixels, T[r] written), then insert.
 * Returns sum_r T_r.  Insert order: r, then x outer, y inner, sequential. */
float r32_update_slices(const float *images, float *slices, const int *mask, const float *respons,
                        const float *scaling, int N, int R, int N2d, int S, float *model,
                        float *weight, const float *x, const float *y, const float *z,
                        const float *q, const float *weights, float *T);
void r32_normalize_model(float *model, const float *weight, int n);
float r32_model_max(const float *model, int n);
void r32_insert_slice(float *model, float *weight, int S, const float *slice, const int *mask,
                      float w, const float *q, const float *x, const float *y, const float *z);

/* ---- FP64 oracle ---- */
enum { O_NEAREST = 0, O_TRILINEAR = 1 };
/* O_GAUSSIAN: the reference's L = -sum m (S - K/phi)^2 / (2 n_m sigma^2);
 * phi = sum m K^2 / sum_r P sum m K S; S' = sum_i P K/phi / T (T > 1e-10).
 * O_POISSON (no counterpart in the reference's GPU path):
 * L = sum m [K (log phi + log max(S, 1e-6)) - phi S];
 * phi = sum m K / sum_r P sum m S; S' = sum_i P K / sum_i P phi (> 1e-30, else 0).
 * Insertion weights are w_r T_r in both. */
enum { O_GAUSSIAN = 0, O_POISSON = 1 };
void o64_expand(const double *model, int S, const float *q, int R, const float *x,
                const float *y, const float *z, int interp, double *slices);
void o64_likelihood(const double *slices, const float *images, const int *mask, double sigma,
                    const double *scaling, int N2d, int N, int R, int kind, double *L);
void o64_softmax(const double *L, int R, int N, double *P);
double o64_total_respons(const double *P, size_t n);
void o64_update_scaling(const float *images, const double *slices, const int *mask,
                        const double *P, int N, int R, int N2d, int kind, double *scaling);
void o64_mstep(const float *images, const int *mask, const double *P, const double *scaling,
               int N, int R, int N2d, int kind, const double *slices_in, double *slices, double *T);
void o64_insert(const double *slices, const int *mask, const double *T, const float *weights,
                int R, int N2d, int S, const float *q, const float *x, const float *y,
                const float *z, int interp, double *model, double *weight);
void o64_normalize_model(double *model, const double *weight, int n);
/* One FP64 iteration (nearest voxel) for large problems: masked pixels only,
 * nonzero counts as lists, zero responsibilities skipped.  allowed [R][N] or
 * NULL restricts each pattern's softmax.  Writes L and P [R][N], the new
 * scaling [N] and the normalised model [S^3], and when given the
 * accumulators before normalisation (raw_out, weight_out [S^3]); returns
 * sum P log P. */
double o64_iteration_large(const double *model_in, const float *images, const int *mask,
                           const double *scaling_in, int N, int R, int S, const float *q,
                           const float *weights, const float *x, const float *y, const float *z,
                           double sigma, int kind, const unsigned char *allowed, double *L,
                           double *P, double *scaling_out, double *model_out, double *raw_out,
                           double *weight_out);

/* ---- whole iteration (driver order emc.c:989-1084), FP32 restatement ----
 * model/weight in-out; scaling in-out; respons scratch [R][N]; slices scratch
 * [R][N2d].  Returns sum P log P.  nthreads <= 0 means all host cores. */
double r32_iteration(float *model, float *weight, float *slices, const float *images,
                     const int *mask, float *scaling, float *respons, int N, int R, int S,
                     const float *x, const float *y, const float *z, const float *q,
                     const float *weights, float sigma, int nthreads);
int o_max_threads(void);
/* Threads of the OpenMP loops from now on (n > 0). */
void o_set_threads(int n);

#ifdef __cplusplus
}
#endif

#endif
