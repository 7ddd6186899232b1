/*
 * emc_host.h -- host-side maths of the EMC driver (C, no GPU).
 *
 * These produce the inputs of the GPU path: the rotation table, the detector
 * coordinates, the detector mask and, for benchmarking and testing, synthetic
 * diffraction data.  They correspond to the host functions of emc.c in
 * FilipeMaia/emc; each declaration cites the part it replaces.
 */
#ifndef EMC_HOST_H
#define EMC_HOST_H

#include <stddef.h>
#include <stdint.h>

#include "emc.h"

#ifdef __cplusplus
extern "C" {
#endif

/* emc.c:43: number of quaternions of the sampling-n table, 20(n + 5n^3). */
int emc_n_to_samples(int n);

/* emc.c:49-363 (generate_rotation_list): the 600-cell vertices plus the
 * interior points of its subdivided edges, faces and cells, quaternions
 * normalised, weights proportional to 1/|q|^3 of the un-normalised point and
 * normalised to sum 1.  Same order and float values as the reference.
 * half = 0: the full table of 20(n+5n^3) entries (it holds q and -q);
 * half = 1: one entry per rotation, 10(n+5n^3) entries: the first entry of
 *           each {q, -q} pair in table order, weight w(q) + w(-q).
 * q receives 4 floats per entry, w one.  Returns the number of entries, or
 * -1 on failure (n < 1 or an internal consistency check failing). */
int emc_rotation_table(int n, int half, float *q, float *w);

/* Coarse cell of each fine rotation (no counterpart in the reference): the
 * coarse entry c maximising |q_fine[r] . q_coarse[c]| (the smallest rotation
 * angle between the two; q and -q are the same rotation), lowest c on ties.
 * Returns 0, or -1 on invalid arguments. */
int emc_assign_cells(const float *q_fine, int n_fine, const float *q_coarse, int n_coarse, int *cell);

/* Same as generate_rotation_list (emc.c:49) with the reference's signature:
 * allocates the Quaternion* array and the weights with malloc. */
int generate_rotation_list(const int n, Quaternion ***return_list, real **return_weights);

/* emc.c:365-393 (calculate_coordinates): per-pixel Fourier-space coordinates
 * in voxel units: x = px - side/2 + 0.5, y likewise, z = tan(theta/2) * r_px.
 * Output layout [p = y*side + x]. */
void emc_coordinates(Setup setup, float *x, float *y, float *z);
/* The reference signature, filling three side x side sp_matrix. */
void calculate_coordinates(Setup setup, sp_matrix *x_coordinates,
                           sp_matrix *y_coordinates, sp_matrix *z_coordinates);

/* Detector mask: 1 = pixel used, 0 = excluded.  Excludes pixels closer than
 * `beamstop` pixels to the detector centre (side/2 - 0.5, side/2 - 0.5), a
 * horizontal gap of `gap_h` rows through the centre and a cross-shaped gap of
 * `gap_cross` rows and columns.  circle != 0 also excludes pixels farther
 * than side/2 from the centre (emc.c:856-864).  Returns the pixel count. */
int emc_make_mask(int side, float beamstop, int gap_h, int gap_cross, int circle, int *mask);

/* Synthetic single-particle diffraction data. */
typedef struct {
  int side;            /* S: model side and detector side in pixels */
  int n_blobs;         /* Gaussian blobs in the real-space object */
  float blob_radius;   /* blob centres uniform in a ball of this radius (voxels) */
  float sigma_min, sigma_max; /* blob widths, uniform (voxels) */
  int n_patterns;      /* N */
  float phi_sigma;     /* pattern scale ~ lognormal(0, phi_sigma), normalised to mean 1 */
  float photons;       /* mean photon count per pattern over the masked pixels */
  uint64_t seed;
} emc_synth_params;

/* True intensity |F(k)|^2 of the blob object on the model grid: voxel v holds
 * the value at frequency k = v - side/2 - 0.5 (the centre convention of the
 * reference's nearest-voxel rule), normalised to mean 1.  OpenMP-parallel,
 * results independent of the thread count. */
int emc_synth_intensity(const emc_synth_params *p, float *volume);

/* Patterns: orientation q_i uniform (Shoemake, as emc.c:27-38), scale phi_i,
 * expected counts lambda_ip = phi_i * c * I(side/2 + 0.5 + R(q_i) c_p) by
 * trilinear interpolation of `volume` (zero outside), Poisson sampled.  c is
 * set so that the mean over patterns of sum_p m_p lambda_ip is `photons`.
 * Unmasked pixels get 0.  images: N x side^2 floats (integer counts);
 * q_true: N x 4 (may be NULL); phi_true: N (may be NULL).  Returns 0. */
int emc_synth_patterns(const emc_synth_params *p, const float *volume, const float *x,
                       const float *y, const float *z, const int *mask, float *images,
                       float *q_true, float *phi_true);

/* Uniform(0,1) initial model (the reference's model_input == 0, emc.c:893-898),
 * from a counter-based generator so it is reproducible from `seed`. */
void emc_synth_random_model(int side, uint64_t seed, float *model);

/* A uniformly random unit quaternion from three uniforms (emc.c:27-38). */
void emc_quaternion_from_uniforms(double u1, double u2, double u3, float q[4]);

/* Counter-based uniform in [0,1): a hash of (seed, stream, a, b). */
double emc_uniform(uint64_t seed, uint64_t stream, uint64_t a, uint64_t b);

#ifdef __cplusplus
}
#endif

#endif /* EMC_HOST_H */
