/*
 * emc.h -- the EMC expand/maximize/compress C-ABI.
 *
 * These are the GPU entry points of FilipeMaia/emc (emc.h:20-56 in the
 * reference), with identical names, argument order, argument types and
 * return types, so that a C driver written against the reference header links
 * against libemc_hip.so unchanged.  Every entry point is implemented with
 * hand-written HIP kernels for gfx950; nothing here goes through a CUDA
 * compatibility layer.
 *
 * Conventions shared by all entry points (they follow the reference):
 *   - device pointers are plain `real *` / `int *` obtained from the
 *     cuda_allocate_* functions below (hipMalloc underneath);
 *   - slices and responsibilities are laid out slice-major:
 *       slices[r * N_2d + p],  respons[r * N_images + i];
 *   - images are pattern-major: images[i * N_2d + p];
 *   - rotations are quaternions q[r*4 + 0..3], q0 the scalar part;
 *   - every function returns only after the results it hands to the host are
 *     complete (host arrays filled, return values computed).
 *
 * Additions that the reference does not have (error reporting, freeing,
 * option selection, the precision form and the learned proposal) live in
 * emc_hip.h; nothing in this file changes behaviour relative to the
 * reference apart from the points listed at each function.
 */
#ifndef EMC_H
#define EMC_H

#include "spimage_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* emc.h:8-10 -- q[0] is the scalar part. */
typedef struct {
  real q[4];
} Quaternion;

/* emc.h:12-18 */
typedef struct {
  int side;
  real wavelength;
  real pixel_size;
  int detector_size;
  real detector_distance;
} Setup;

/* Host (CPU) nearest-neighbour insertion of one slice (emc.h:20-22,
 * defined at emc.c:491-539).  Accumulates model += w*slice and weight += w at
 * the voxel nearest to each pixel whose mask value is 1. */
void insert_slice(sp_3matrix *model, sp_3matrix *weight, sp_matrix *slice,
                  sp_imatrix *mask, real w, Quaternion *rot, sp_matrix *x_coordinates,
                  sp_matrix *y_coordinates, sp_matrix *z_coordinates);

/* M-step + compress (emc.h:24-29, emc_cuda.cu:292-362, emc_atomic.cu:76-133).
 * For every rotation r: T_r = sum_i P_ri; for masked pixels
 * slices[r][p] = sum_i images[i][p] * P_ri / scaling[i], divided by T_r when
 * T_r > 1e-10; then, for T_r > 1e-10, inserts the slice into d_model and
 * d_weight with weight weights[r] * T_r (nearest neighbour, bounds 0<=v<side).
 * Copies the (not yet normalised) model to model->data and returns sum_r T_r.
 * `weights` is a host array of N_slices rotation weights. */
real cuda_update_slices(real *d_images, real *slices, int *d_mask,
                        real *d_respons, real *d_scaling, int N_images, int N_slices, int N_2d,
                        sp_3matrix *model, real *d_model,
                        real *x_coordinates, real *y_coordinates,
                        real *z_coordinates, real *d_rotations, real *weights,
                        real *d_weight, Setup setup, sp_matrix **images);

/* Scaling update (emc.h:32-34, emc_cuda.cu:195-264):
 * scaling[i] = sum_p m_p K_ip^2 / sum_r P_ri sum_p m_p K_ip S_rp, written to
 * d_scaling and copied to the host array `scaling` (N_images values). */
void cuda_update_scaling(real *d_images, real *slices, int *d_mask,
                         real *d_respons, real *d_scaling, int N_images, int N_slices, int N_2d,
                         real *scaling);

/* Log-likelihood (emc.h:35-37, emc_cuda.cu:113-193):
 * d_respons[r][i] = -sum_p m_p (S_rp - K_ip/scaling_i)^2 / (2 * n_m * sigma^2)
 * with n_m the number of pixels whose mask is non-zero.  The raw values are
 * copied to the host array `respons` (N_slices*N_images values).  Difference
 * from the reference: `respons` may be NULL, which skips that copy (it is
 * N*R*4 bytes, 10 GB at 50,000 patterns x 50,100 rotations). */
void cuda_calculate_responsabilities(real *slices, real *images, int *d_mask,
                                     real sigma, real *d_scaling, real *d_respons,
                                     int N_2d, int N_images, int N_slices, real *respons);

/* Expand (emc.h:39-41, emc_cuda.cu:51-111,267-290): for every rotation r and
 * pixel p, d_slices[r][p] = d_model[v] with v the voxel nearest to the
 * rotated pixel coordinate (0 < v < side on every axis), else 0. */
void cuda_get_slices(sp_3matrix *model, real *d_model, real *d_slices, real *d_rot,
                     real *d_x_coordinates,
                     real *d_y_coordinates, real *d_z_coordinates, int N_slices);

/* Device buffer for N_slices slices of side*side pixels (emc_cuda.cu:370-372). */
void cuda_allocate_slices(real **slices, Setup setup, int N_slices);
/* max(0, max_v model[v]) (emc_cuda.cu:364-368). */
real cuda_model_max(real *model, int model_size);
/* Device copy of a volume (emc_cuda.cu:374-377). */
void cuda_allocate_model(real **d_model, sp_3matrix *model);
/* Device copy of the detector mask (emc_cuda.cu:379-382). */
void cuda_allocate_mask(int **d_mask, sp_imatrix *mask);
/* model[:] = 0 (emc_cuda.cu:410-412). */
void cuda_reset_model(sp_3matrix *model, real *d_model);
/* model = weight > 0 ? model / weight : 0, then model /= mean(model)
 * (emc_cuda.cu:414-433). */
void cuda_normalize_model(sp_3matrix *model, real *d_model, real *d_weight);
/* Device copy of N_slices quaternions, 4 reals each (emc_cuda.cu:384-390). */
void cuda_allocate_rotations(real **d_rotations, Quaternion **rotations, int N_slices);
/* Device copy of N_images images, pattern-major (emc_cuda.cu:392-398). */
void cuda_allocate_images(real **d_images, sp_matrix **images, int N_images);
/* Device copies of the three pixel-coordinate maps (emc_cuda.cu:400-408). */
void cuda_allocate_coords(real **d_x, real **d_y, real **d_z, sp_matrix *x,
                          sp_matrix *y, sp_matrix *z);
/* Uninitialised device buffer of n BYTES (emc_cuda.cu:435-437).  As in the
 * reference the count is an int; emc_hip_allocate_bytes in emc_hip.h takes a
 * size_t for buffers of 2 GiB and more. */
void cuda_allocate_real(real **x, int n);
/* Device scaling array filled with 1 (emc_cuda.cu:439-443). */
void cuda_allocate_scaling(real **scaling, int N_images);
/* Per-pattern softmax over rotations, in place (emc_cuda.cu:445-482):
 * m_i = max(-1e10, max_r L_ri); P_ri = exp(L_ri - m_i) where L_ri - m_i > -1e10
 * (else 0), then P_ri /= sum_r P_ri.  The reference kernel has a race between
 * reading the block maximum and clearing the reduction buffer
 * (emc_cuda.cu:457-461); this implementation computes the true maximum. */
void cuda_normalize_responsabilities(real *d_respons, int N_slices, int N_images);
/* sum over n entries of P log P, terms with P <= 0 count 0 (emc_cuda.cu:484-505). */
real cuda_total_respons(real *d_respons, real *respons, int n);

#ifdef __cplusplus
}
#endif

#endif /* EMC_H */
