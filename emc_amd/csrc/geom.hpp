// Rotation of detector coordinates into the model grid, shared by slicing
// (expand) and insertion so both use one voxel index per (rotation, pixel).
#pragma once

#include <hip/hip_runtime.h>

namespace emc {

// Rotation matrix of q (q0 scalar) written term by term as emc_cuda.cu:65-85
// (and emc_atomic.cu:32-52) evaluate it in float.  FMA contraction is off so
// the value is the same as the oracle's no-FMA float evaluation.
__device__ __forceinline__ void rot_matrix(const float* __restrict__ q, float e[9]) {
#pragma clang fp contract(off)
  const float a = q[0], b = q[1], c = q[2], d = q[3];
  e[0] = a * a + b * b - c * c - d * d;
  e[1] = 2.0f * b * c - 2.0f * a * d;
  e[2] = 2.0f * b * d + 2.0f * a * c;
  e[3] = 2.0f * b * c + 2.0f * a * d;
  e[4] = a * a - b * b + c * c - d * d;
  e[5] = 2.0f * c * d - 2.0f * a * b;
  e[6] = 2.0f * b * d - 2.0f * a * c;
  e[7] = 2.0f * c * d + 2.0f * a * b;
  e[8] = a * a - b * b - c * c + d * d;
}

// Continuous voxel coordinate (S/2 + 0.5) + R c per axis; rounding it gives
// the reference's nearest voxel (emc_cuda.cu:86-88).
__device__ __forceinline__ void voxel_coord(const float e[9], float half, float cx, float cy, float cz,
                                            float& ux, float& uy, float& uz) {
#pragma clang fp contract(off)
  const float nx = e[0] * cx + e[1] * cy + e[2] * cz;
  const float ny = e[3] * cx + e[4] * cy + e[5] * cz;
  const float nz = e[6] * cx + e[7] * cy + e[8] * cz;
  ux = half + nx;
  uy = half + ny;
  uz = half + nz;
}

// Trilinear footprint: corner c (bit 0 x, bit 1 y, bit 2 z) of the cell
// containing u, its weight, and whether it lies in 0 <= v < S.
__device__ __forceinline__ bool tri_corner(int c, float ux, float uy, float uz, int S, int& vox, float& w) {
  const float fx0 = floorf(ux), fy0 = floorf(uy), fz0 = floorf(uz);
  const int ix = (int)fx0 + (c & 1), iy = (int)fy0 + ((c >> 1) & 1), iz = (int)fz0 + ((c >> 2) & 1);
  const float fx = ux - fx0, fy = uy - fy0, fz = uz - fz0;
  w = ((c & 1) ? fx : 1.0f - fx) * (((c >> 1) & 1) ? fy : 1.0f - fy) * (((c >> 2) & 1) ? fz : 1.0f - fz);
  if (ix < 0 || ix >= S || iy < 0 || iy >= S || iz < 0 || iz >= S) return false;
  vox = (iz * S + iy) * S + ix;
  return true;
}

// Model copy for the expand's gathers, in 4 x 4 x 4 voxel bricks (256 B,
// two cache lines of 4 x 4 x 2): a slice plane crosses few bricks, where the
// linear x-fastest layout puts nearly every voxel of a tilted plane in its
// own cache line.  S % 4 == 0.
__device__ __forceinline__ size_t brick_index(int x, int y, int z, int S) {
  const int nb = S >> 2;
  return ((((size_t)(z >> 2) * nb + (y >> 2)) * nb + (x >> 2)) << 6) + (((z & 3) * 4 + (y & 3)) * 4 + (x & 3));
}

// Model value at voxel (x, y, z), inside the grid, from the linear or the
// bricked layout.
template <bool BRICK>
__device__ __forceinline__ float model_at(const float* __restrict__ m, int x, int y, int z, int S) {
  return BRICK ? m[brick_index(x, y, z, S)] : m[((size_t)z * S + y) * S + x];
}

// Trilinear corner c as voxel coordinates (see tri_corner).
__device__ __forceinline__ bool tri_corner_xyz(int c, float ux, float uy, float uz, int S, int& ix, int& iy, int& iz,
                                               float& w) {
  const float fx0 = floorf(ux), fy0 = floorf(uy), fz0 = floorf(uz);
  ix = (int)fx0 + (c & 1);
  iy = (int)fy0 + ((c >> 1) & 1);
  iz = (int)fz0 + ((c >> 2) & 1);
  const float fx = ux - fx0, fy = uy - fy0, fz = uz - fz0;
  w = ((c & 1) ? fx : 1.0f - fx) * (((c >> 1) & 1) ? fy : 1.0f - fy) * (((c >> 2) & 1) ? fz : 1.0f - fz);
  return !(ix < 0 || ix >= S || iy < 0 || iy >= S || iz < 0 || iz >= S);
}

}  // namespace emc
