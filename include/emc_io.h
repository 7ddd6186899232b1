/*
 * emc_io.h -- the driver's on-disk formats (C, no GPU, no libhdf5).
 *
 * The reference reads its patterns and mask and writes its models through
 * libspimage's HDF5 files (emc.c:777-864, 917-950, 1072-1082) and the later
 * viewer reads further HDF5 outputs with h5py (the viewer/ scripts).  Neither
 * libspimage nor libhdf5 is available here, so this is a small HDF5 reader
 * and writer of its own for the subset those files use: a version-0
 * superblock, one root group (symbol table: v1 B-tree, local heap, symbol
 * nodes), datasets with version-1 object headers, contiguous storage, rank
 * 0-3, little-endian IEEE float32/float64 and integers.  Files written
 * with libhdf5's default (earliest) file format and contiguous datasets in
 * the root group are read as well; anything else (chunked or compressed
 * data, superblock version >= 2, groups below the root) is refused with an
 * error.
 */
#ifndef EMC_IO_H
#define EMC_IO_H

#ifdef __cplusplus
extern "C" {
#endif

enum { EMC_H5_F32 = 0, EMC_H5_I32 = 1, EMC_H5_F64 = 2 };

/* One dataset to write: dims in C order (last varies fastest), rank 0..3
 * (rank 0: a scalar, one element); data holds prod(dims) elements of type. */
typedef struct {
  const char *name;
  int type;
  int rank;
  long long dims[3];
  const void *data;
} emc_h5_dataset;

/* Write the datasets (distinct names, at most 64) into a new file at path,
 * replacing any file there.  Returns 0, or -1 (message in emc_io_error). */
int emc_h5_write(const char *path, const emc_h5_dataset *sets, int n);

/* Read dataset `name` of the root group, converted to `type`
 * (EMC_H5_F32/I32/F64), into buf of cap elements.  *rank and dims[0..rank)
 * receive its shape (either may be NULL).  Returns the element count, or -1
 * (not found, unsupported layout or type, cap too small: emc_io_error). */
long long emc_h5_read(const char *path, const char *name, int type, void *buf, long long cap, int *rank,
                      long long dims[3]);

/* Message of the last failed call of this header. */
const char *emc_io_error(void);

/* An image or volume in the layout of libspimage's HDF5 files (the files
 * of emc.c:804,840,917,944,1082): datasets "real" (float32, dims (nz,) ny, nx:
 * x fastest, as the reference's model[z][y][x] and image[y][x]), "mask"
 * (int32, same dims), "num_dimensions", "image_center", "phased",
 * "shifted", "scaled", "detector_distance", "lambda", "pixel_size",
 * "version".  nz = 1 writes a 2-D image.  mask may be NULL (all ones).
 * libspimage itself is absent here, so agreement with its files rests on
 * this description of them (dataset names, types and axis order). */
int emc_write_spimage(const char *path, const float *real, const int *mask, int nx, int ny, int nz);
/* Read such a file: |"real" + i "imag"| (the modulus the reference takes
 * with sp_cabs) into real, the mask into mask (may be NULL); the dataset
 * must hold nx*ny*nz elements.  Returns 0 or -1. */
int emc_read_spimage(const char *path, float *real, int *mask, int nx, int ny, int nz);
/* Shape of the "real" dataset of such a file (nz = 1 for an image).
 * Returns 0 or -1. */
int emc_spimage_shape(const char *path, int *nx, int *ny, int *nz);

/* The reference driver's pattern ingest (emc.c:777-832, rescale_intensity
 * 0): image i is read from the file named by printf(pattern, i) (the
 * reference's "Data/image_%.4d.h5") and sampled on the side x side detector
 * grid with stride `stride` about its centre (emc.c:777: read_stride 4):
 * images[i][y*side + x] = |img(sx, sy)| at sx = (int)(stride * ((x - side/2)
 * + 0.5) + nx/2 - 0.5) (integer x - side/2 and nx/2), likewise sy.  Fails
 * if a sampled pixel falls outside an image.  Returns 0 or -1. */
int emc_read_patterns(const char *pattern, int n_images, int side, int stride, float *images);
/* The mask (emc.c:834-864): with a path (mask_input 1), 0 where the file's
 * |value| at the strided position is 0, else 1; with path NULL
 * (mask_input 0) all 1; then every pixel farther than side/2 from
 * (side/2 - 0.5, side/2 - 0.5) is set to 0.  Returns 0 or -1. */
int emc_read_mask(const char *path, int side, int stride, int *mask);

#ifdef __cplusplus
}
#endif

#endif /* EMC_IO_H */
