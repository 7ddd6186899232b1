// Pieces shared by the FP32 and FP16 matrix-core contractions: workgroup ->
// tile mapping, the likelihood and M-step epilogues, the responsibility
// loader of the M-step.
#pragma once

#include <hip/hip_runtime.h>

#include "engine.hpp"

namespace emc {

typedef float floatx4_t __attribute__((ext_vector_type(4)));

// Workgroup -> output tile.  Consecutive workgroup ids are dispatched round
// robin over the 8 XCDs, each with its own L2; remap so that every XCD works
// on a contiguous range of tiles, and walk that range in groups of 8 tile
// rows so neighbouring workgroups share operand panels in L2.
// Logical tile of block blockIdx.x among nt blocks: XCD x takes the
// consecutive logical tiles from x * nt / 8 on.
__device__ __forceinline__ int logical_of_block(int nt) {
  const int bid = blockIdx.x;
  const int xcd = bid & 7, slot = bid >> 3;
  const int q = nt >> 3, rem = nt & 7;
  return xcd * q + min(xcd, rem) + slot;
}

// (tm, tn) of a logical tile: groups of 8 row tiles walked column by column.
__device__ __forceinline__ void tile_of_logical(int logical, int tiles_m, int tiles_n, int& tm, int& tn) {
  constexpr int G = 8;
  const int per_group = G * tiles_n;
  const int group = logical / per_group;
  const int first_m = group * G;
  const int gsize = min(tiles_m - first_m, G);
  const int in_group = logical - group * per_group;
  tm = first_m + in_group % gsize;
  tn = in_group / gsize;
}

__device__ __forceinline__ void tile_of_block(int tiles_m, int tiles_n, int& tm, int& tn) {
  tile_of_logical(logical_of_block(tiles_m * tiles_n), tiles_m, tiles_n, tm, tn);
}

// Likelihood epilogue for one lane block: rows i0..i0+3 (patterns), column r
// (rotation).  C = acc * cscale[r] / kscale, then in FP64
// L = -(A_r - 2 C/phi_i + B_i/phi_i^2) / (2 n_m sigma^2) (Gaussian) or
// L = C + B_i log phi_i - phi_i A_r (Poisson, C over log S), stored to L[r][i].
__device__ __forceinline__ void like_store(const LikeEpilogue& ep, int i0, int r, floatx4_t acc) {
  if (r >= ep.R) return;
  const double cs = (double)(ep.cscale ? ep.cscale[r] : 1.0f) * (double)ep.kscale_inv;
  const double a = ep.A[r];
  float out[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int i = i0 + j;
    out[j] = 0.0f;
    if (i < ep.N) {
      const double c = (double)acc[j] * cs;
      const double phi = (double)ep.phi[i];
      out[j] = ep.poisson ? (float)(c + ep.B[i] * log(phi) - phi * a)
                          : (float)(-(a - 2.0 * c / phi + ep.B[i] / (phi * phi)) * ep.inv_2nms2);
    }
  }
  float* dst = ep.L + (size_t)r * ep.N + i0;
  if (ep.vecL && i0 + 3 < ep.N) {
    *reinterpret_cast<floatx4_t*>(dst) = floatx4_t{out[0], out[1], out[2], out[3]};
  } else {
#pragma unroll
    for (int j = 0; j < 4; j++)
      if (i0 + j < ep.N) dst[j] = out[j];
  }
}

// sum_r P_ri C_ir epilogue: reduce the lane's values over its columns
// (rotations) across the 16 lanes of a group, one atomic per row.
__device__ __forceinline__ void like_pc_flush(const LikeEpilogue& ep, int i0, double v[4]) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    double s = v[j];
#pragma unroll
    for (int o = 8; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((lane & 15) == 0 && i0 + j < ep.N) atomicAdd(&ep.D[i0 + j], s);
  }
}

// Row r of the responsibilities, 4 patterns from i: P_ri * invphi_i * rs,
// zero outside the matrix.
__device__ __forceinline__ floatx4_t load_p4(const MstepArgs& a, int r, int i, float rs) {
  floatx4_t v = {0, 0, 0, 0};
  if (r >= a.R) return v;
  const float* row = a.P + (size_t)r * a.N;
  if (a.vecP) {
    if (i < a.N) {
      v = *reinterpret_cast<const floatx4_t*>(row + i);
      const floatx4_t ip = *reinterpret_cast<const floatx4_t*>(a.invphi + i);
      v = v * ip * rs;
    }
  } else {
#pragma unroll
    for (int c = 0; c < 4; c++)
      if (i + c < a.N) v[c] = row[i + c] * a.invphi[i + c] * rs;
  }
  return v;
}

// M-step epilogue: rows r0..r0+3 (rotations), column k (packed pixel).
// U = acc * extra / rowscale[r]; S' = T_r > 1e-10 ? U / T_r : U written to
// the full slice and to the packed copy used by insertion (raw: U itself to
// the packed copy only).
__device__ __forceinline__ void mstep_store(const MstepArgs& a, int r0, int k, floatx4_t acc, float extra) {
  if (k >= a.n_m) return;
  const int p = a.pix[k];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int r = r0 + j;
    if (r >= a.R) continue;
    float u = acc[j] * extra;
    if (a.rowscale) u = u / a.rowscale[r];
    if (a.raw) {
      a.Sc[(size_t)r * a.n_mp + k] = u;
      continue;
    }
    const float T = a.Tf[r];
    const float s = a.poisson ? (T > 1e-30f ? u / T : 0.0f) : (T > 1e-10f ? u / T : u);
    a.slices[(size_t)r * a.N2d + p] = s;
    a.Sc[(size_t)r * a.n_mp + k] = s;
  }
}

}  // namespace emc
