/*
 * Minimal stand-in for the libspimage container types that the EMC C-ABI
 * passes across the boundary.  libspimage itself is not a dependency of this
 * project; only the memory layout of its three matrix structs matters here.
 *
 *   sp_matrix   2-D real matrix, column major: element (row, col) lives at
 *               data[col * rows + row].  For an S x S detector image "row" is
 *               the pixel x index and "col" the pixel y index, so the linear
 *               pixel index is p = y * S + x (emc_cuda.cu:67 indexes
 *               x_coordinates[y*x_max+x] the same way).
 *   sp_imatrix  same layout with int elements (detector mask).
 *   sp_3matrix  3-D real volume, element (x, y, z) at data[z*x*y + y*x + x]
 *               (emc_cuda.cu:92 indexes the model as
 *               round_z*model_x*model_y + round_y*model_x + round_x).
 *
 * `real` is float, as in the reference build (kernels take float* for real*,
 * emc_cuda.cu:114,130).
 */
#ifndef EMC_SPIMAGE_TYPES_H
#define EMC_SPIMAGE_TYPES_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef float real;

typedef struct {
  unsigned int rows;
  unsigned int cols;
  real *data;
} sp_matrix;

typedef struct {
  unsigned int rows;
  unsigned int cols;
  int *data;
} sp_imatrix;

typedef struct {
  unsigned int x;
  unsigned int y;
  unsigned int z;
  real *data;
} sp_3matrix;

static inline unsigned int sp_matrix_rows(const sp_matrix *m) { return m->rows; }
static inline unsigned int sp_matrix_cols(const sp_matrix *m) { return m->cols; }
static inline size_t sp_matrix_size(const sp_matrix *m) { return (size_t)m->rows * m->cols; }
static inline real sp_matrix_get(const sp_matrix *m, unsigned int row, unsigned int col) {
  return m->data[(size_t)col * m->rows + row];
}
static inline void sp_matrix_set(sp_matrix *m, unsigned int row, unsigned int col, real v) {
  m->data[(size_t)col * m->rows + row] = v;
}

static inline size_t sp_imatrix_size(const sp_imatrix *m) { return (size_t)m->rows * m->cols; }
static inline int sp_imatrix_get(const sp_imatrix *m, unsigned int row, unsigned int col) {
  return m->data[(size_t)col * m->rows + row];
}
static inline void sp_imatrix_set(sp_imatrix *m, unsigned int row, unsigned int col, int v) {
  m->data[(size_t)col * m->rows + row] = v;
}

static inline unsigned int sp_3matrix_x(const sp_3matrix *m) { return m->x; }
static inline unsigned int sp_3matrix_y(const sp_3matrix *m) { return m->y; }
static inline unsigned int sp_3matrix_z(const sp_3matrix *m) { return m->z; }
static inline size_t sp_3matrix_size(const sp_3matrix *m) { return (size_t)m->x * m->y * m->z; }
static inline real sp_3matrix_get(const sp_3matrix *m, unsigned int x, unsigned int y, unsigned int z) {
  return m->data[(size_t)z * m->x * m->y + (size_t)y * m->x + x];
}
static inline void sp_3matrix_set(sp_3matrix *m, unsigned int x, unsigned int y, unsigned int z, real v) {
  m->data[(size_t)z * m->x * m->y + (size_t)y * m->x + x] = v;
}

#ifdef __cplusplus
}
#endif

#endif /* EMC_SPIMAGE_TYPES_H */
