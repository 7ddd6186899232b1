/*
 * Detector geometry and the host-side slice insertion of the EMC C-ABI.
 */
#include <math.h>
#include <stdlib.h>

#include "emc_host.h"

/* emc.c:365-393.  Pixel (x, y) sits at (x - side/2 + 0.5, y - side/2 + 0.5)
 * in voxel units; the Ewald-sphere depth is z = r * (1 - cos t) / sin t with
 * t = atan2(r * pixel_size, distance), i.e. r * tan(t/2).  The reference mixes
 * float storage with double library calls; the same conversions are kept so
 * the coordinates match it bit for bit. */
void emc_coordinates(Setup setup, float *xo, float *yo, float *zo) {
  const int side = setup.side;
  const int half = side / 2;
  for (int y = 0; y < side; y++)
    for (int x = 0; x < side; x++) {
      const double dx = (float)(x - half) + 0.5;
      const double dy = (float)(y - half) + 0.5;
      const float r = (float)sqrt(dx * dx + dy * dy);
      const float t = (float)atan2(r * setup.pixel_size, setup.detector_distance);
      const float kr = (float)(sin(t) / setup.wavelength);
      const float kz = (float)((1.0 - cos(t)) / setup.wavelength);
      const size_t p = (size_t)y * side + x;
      xo[p] = (float)dx;
      yo[p] = (float)dy;
      zo[p] = kz / kr * r;
    }
}

void calculate_coordinates(Setup setup, sp_matrix *x_coordinates, sp_matrix *y_coordinates,
                           sp_matrix *z_coordinates) {
  /* sp_matrix (row = x, col = y) is column major, i.e. data[y*side + x]:
   * the same linear layout emc_coordinates writes. */
  emc_coordinates(setup, x_coordinates->data, y_coordinates->data, z_coordinates->data);
}

int emc_make_mask(int side, float beamstop, int gap_h, int gap_cross, int circle, int *mask) {
  const double c = side / 2.0 - 0.5;
  int count = 0;
  for (int y = 0; y < side; y++)
    for (int x = 0; x < side; x++) {
      const double dx = x - c, dy = y - c;
      const double r2 = dx * dx + dy * dy;
      int m = 1;
      if (r2 < (double)beamstop * beamstop) m = 0;
      if (gap_h > 0 && y >= side / 2 - gap_h / 2 && y < side / 2 - gap_h / 2 + gap_h) m = 0;
      if (gap_cross > 0) {
        const int lo = side / 2 - gap_cross / 2, hi = lo + gap_cross;
        if ((y >= lo && y < hi) || (x >= lo && x < hi)) m = 0;
      }
      /* emc.c:856-864 */
      if (circle && sqrt(pow((double)x - side / 2.0 + 0.5, 2) + pow((double)y - side / 2.0 + 0.5, 2)) > side / 2.0)
        m = 0;
      mask[(size_t)y * side + x] = m;
      count += m;
    }
  return count;
}

/* emc.c:491-539, the host twin of the GPU insertion.  The reference computes
 * the rotated coordinate in double (2.0 literals) and rounds with round();
 * the voxel bounds are 0 <= v < side and only pixels whose mask is 1 are
 * inserted.  Used by the driver for the random-orientation initial model. */
void insert_slice(sp_3matrix *model, sp_3matrix *weight, sp_matrix *slice, sp_imatrix *mask,
                  real w, Quaternion *rot, sp_matrix *x_coordinates, sp_matrix *y_coordinates,
                  sp_matrix *z_coordinates) {
  const int rows = (int)sp_matrix_rows(slice), cols = (int)sp_matrix_cols(slice);
  const int mx = (int)sp_3matrix_x(model), my = (int)sp_3matrix_y(model), mz = (int)sp_3matrix_z(model);
  const float *q = rot->q;
  /* rotation matrix, evaluated as the reference writes it: the diagonal
   * entries are float expressions (and their product with the float
   * coordinate stays float), the off-diagonal ones are promoted to double by
   * the 2.0 literal */
  const float a00 = q[0] * q[0] + q[1] * q[1] - q[2] * q[2] - q[3] * q[3];
  const double a01 = 2.0 * q[1] * q[2] - 2.0 * q[0] * q[3];
  const double a02 = 2.0 * q[1] * q[3] + 2.0 * q[0] * q[2];
  const double a10 = 2.0 * q[1] * q[2] + 2.0 * q[0] * q[3];
  const float a11 = q[0] * q[0] - q[1] * q[1] + q[2] * q[2] - q[3] * q[3];
  const double a12 = 2.0 * q[2] * q[3] - 2.0 * q[0] * q[1];
  const double a20 = 2.0 * q[1] * q[3] - 2.0 * q[0] * q[2];
  const double a21 = 2.0 * q[2] * q[3] + 2.0 * q[0] * q[1];
  const float a22 = q[0] * q[0] - q[1] * q[1] - q[2] * q[2] + q[3] * q[3];
  for (int x = 0; x < rows; x++)
    for (int y = 0; y < cols; y++) {
      if (sp_imatrix_get(mask, x, y) != 1) continue;
      const float cx = sp_matrix_get(x_coordinates, x, y);
      const float cy = sp_matrix_get(y_coordinates, x, y);
      const float cz = sp_matrix_get(z_coordinates, x, y);
      /* new_x etc. are stored in float variables in the reference */
      const float nx = (float)((double)(a00 * cx) + a01 * cy + a02 * cz);
      const float ny = (float)(a10 * cx + (double)(a11 * cy) + a12 * cz);
      const float nz = (float)(a20 * cx + a21 * cy + (double)(a22 * cz));
      const int vx = (int)round((float)mx / 2.0 + 0.5 + nx);
      const int vy = (int)round((float)my / 2.0 + 0.5 + ny);
      const int vz = (int)round((float)mz / 2.0 + 0.5 + nz);
      if (vx < 0 || vx >= mx || vy < 0 || vy >= my || vz < 0 || vz >= mz) continue;
      sp_3matrix_set(model, vx, vy, vz, sp_3matrix_get(model, vx, vy, vz) + w * sp_matrix_get(slice, x, y));
      sp_3matrix_set(weight, vx, vy, vz, sp_3matrix_get(weight, vx, vy, vz) + w);
    }
}

