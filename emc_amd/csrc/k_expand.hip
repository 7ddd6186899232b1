// Expand (slicing the model at every rotation) and operand packing kernels.
//
// Slicing follows emc_cuda.cu:51-111: the pixel coordinate is rotated by the
// quaternion's rotation matrix, written exactly as the reference writes it in
// float, and the voxel index along each axis is roundf(S/2 + 0.5 + c).  FMA
// contraction is disabled in these functions so the index matches the
// oracle's float evaluation bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <utility>
#include <vector>

#include "engine.hpp"
#include "gemm_common.hpp"
#include "geom.hpp"
#include "reduce.hpp"

namespace emc {

// Pixel p handled by lane slot t of a workgroup: each wave takes an 8 x 8
// pixel tile, so the voxels it gathers lie on a small patch of the rotated
// plane (few cache lines per wave) while its slice writes stay 32-byte rows.
// Needs S % 8 == 0; otherwise pixels in row order.
__device__ __forceinline__ int expand_pixel(int t, int S) {
  if (S & 7) return t;
  const int tile = t >> 6, l = t & 63, tiles_x = S >> 3;
  return ((tile / tiles_x) * 8 + (l >> 3)) * S + (tile % tiles_x) * 8 + (l & 7);
}

// Slice packing runs 1,024 threads per workgroup (rotation): the fused
// expand + pack keeps the slice in 64 KB of LDS, and two such workgroups
// still fill a CU with waves for the model gathers; the separate pack uses
// the same count so that both sum A_r in the same order.
constexpr int kPackThreads = 1024;
constexpr int kPackPer = 16;

// Bricked copy of the model for the gathers (brick_index).
__global__ __launch_bounds__(256) void k_brick_model(const float* __restrict__ model, int S, float* __restrict__ out) {
  const size_t v = (size_t)blockIdx.x * 256 + threadIdx.x;
  const size_t n = (size_t)S * S * S;
  if (v >= n) return;
  const int x = (int)(v % S), y = (int)((v / S) % S), z = (int)(v / ((size_t)S * S));
  out[brick_index(x, y, z, S)] = model[v];
}

void launch_brick_model(const float* model, int S, float* out, hipStream_t st) {
  const size_t n = (size_t)S * S * S;
  hipLaunchKernelGGL(k_brick_model, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, model, S, out);
  EMC_LAUNCHED();
}

// ---------------------------------------------------------------------------
// Processing order of the rotations.  The voxels a slice gathers lie on the
// rotated detector surface, whose normal is R e_z = (e[2], e[5], e[8]):
// rotations with nearby normals (whatever their in-plane angle) gather from
// nearly the same voxels.  The expand kernels therefore take the rotations
// in the order of a Hilbert curve over the normal's direction (the
// octahedral map of the hemisphere n_z >= 0; n and -n give the same plane
// through the centre), so that the workgroups in flight on one XCD share
// their model bricks in that XCD's L2.  Each rotation's output stays at its
// own row: the order changes nothing but the schedule.
// ---------------------------------------------------------------------------
static uint32_t hilbert_index(uint32_t n, uint32_t x, uint32_t y) {
  uint32_t d = 0;
  for (uint32_t s = n / 2; s > 0; s /= 2) {
    const uint32_t rx = (x & s) ? 1u : 0u, ry = (y & s) ? 1u : 0u;
    d += s * s * ((3u * rx) ^ ry);
    if (ry == 0) {  // rotate the quadrant so the curve stays continuous
      if (rx == 1) {
        x = n - 1 - x;
        y = n - 1 - y;
      }
      const uint32_t t = x;
      x = y;
      y = t;
    }
  }
  return d;
}

static const int* rotation_order(const float* rot, int R, hipStream_t st) {
  Engine& e = engine();
  if (!e.expand_sorted || R <= 1) return nullptr;
  auto& k = e.order_key;
  if (k.rot == rot && k.R == R && k.stamp == e.invalidate_stamp) return e.d_rot_order.as<int>();
  std::vector<float> q((size_t)R * 4);
  EMC_CHECK(hipMemcpyAsync(q.data(), rot, sizeof(float) * 4 * (size_t)R, hipMemcpyDeviceToHost, st));
  EMC_CHECK(hipStreamSynchronize(st));
  if (failed()) return nullptr;
  constexpr uint32_t kGrid = 1024;
  std::vector<std::pair<uint32_t, int>> key((size_t)R);
  for (int r = 0; r < R; r++) {
    const double a = q[4 * (size_t)r], b = q[4 * (size_t)r + 1], c = q[4 * (size_t)r + 2], d = q[4 * (size_t)r + 3];
    double nx = 2.0 * b * d + 2.0 * a * c, ny = 2.0 * c * d - 2.0 * a * b, nz = a * a - b * b - c * c + d * d;
    if (nz < 0.0) {
      nx = -nx;
      ny = -ny;
      nz = -nz;
    }
    const double l1 = std::fabs(nx) + std::fabs(ny) + nz;
    const double u = l1 > 0.0 ? nx / l1 : 0.0, v = l1 > 0.0 ? ny / l1 : 0.0;  // |u| + |v| <= 1
    // the diamond |u| + |v| <= 1 turned into the square [-1, 1]^2
    const double su = u + v, sv = u - v;
    const auto cell = [](double w) {
      const long c = (long)std::floor((w + 1.0) * 0.5 * kGrid);
      return (uint32_t)std::min<long>(kGrid - 1, std::max<long>(0, c));
    };
    key[(size_t)r] = {hilbert_index(kGrid, cell(su), cell(sv)), r};
  }
  std::sort(key.begin(), key.end());
  std::vector<int> order((size_t)R);
  for (int r = 0; r < R; r++) order[(size_t)r] = key[(size_t)r].second;
  if (!e.d_rot_order.ensure(sizeof(int) * (size_t)R)) return nullptr;
  EMC_CHECK(hipMemcpyAsync(e.d_rot_order.p, order.data(), sizeof(int) * (size_t)R, hipMemcpyHostToDevice, st));
  EMC_CHECK(hipStreamSynchronize(st));
  if (failed()) return nullptr;
  k.rot = rot;
  k.R = R;
  k.stamp = e.invalidate_stamp;
  return e.d_rot_order.as<int>();
}

// Rotation of logical index lg (order == nullptr: the table order; indices
// at or beyond R, the padding rows, map to themselves).
__device__ __forceinline__ int rotation_of(const int* __restrict__ order, int lg, int R) {
  return order && lg < R ? order[lg] : lg;
}

// One workgroup per rotation; model reads are gathers served by L2 (the
// pixel loops are unrolled so that several gathers per lane are in flight).  Each
// XCD takes a contiguous range of rotations (logical_of_block): neighbouring
// table entries are nearby orientations whose slices gather overlapping
// model bricks, so that XCD's L2 serves them.
template <bool BRICK>
__global__ __launch_bounds__(256) void k_expand_nearest(const float* __restrict__ model, int S,
                                                        const float* __restrict__ rot,
                                                        const float* __restrict__ x,
                                                        const float* __restrict__ y,
                                                        const float* __restrict__ z,
                                                        float* __restrict__ slices,
                                                        const int* __restrict__ order) {
  const int r = rotation_of(order, logical_of_block(gridDim.x), gridDim.x);
  float e[9];
  rot_matrix(&rot[4 * (size_t)r], e);
  const float half = (float)S / 2.0f + 0.5f;
  const int N2d = S * S;
  float* out = slices + (size_t)r * N2d;
#pragma unroll 4
  for (int t = threadIdx.x; t < N2d; t += blockDim.x) {
    const int p = expand_pixel(t, S);
    float ux, uy, uz;
    voxel_coord(e, half, x[p], y[p], z[p], ux, uy, uz);
    const int vx = (int)roundf(ux), vy = (int)roundf(uy), vz = (int)roundf(uz);
    float v = 0.0f;
    if (vx > 0 && vx < S && vy > 0 && vy < S && vz > 0 && vz < S) v = model_at<BRICK>(model, vx, vy, vz, S);
    out[p] = v;
  }
}

// Trilinear slicing (not in the reference): the 8 voxels around the
// continuous coordinate, those inside 0 <= v < S, weighted by the usual
// products of fractional distances.
template <bool BRICK>
__global__ __launch_bounds__(256) void k_expand_trilinear(const float* __restrict__ model, int S,
                                                          const float* __restrict__ rot,
                                                          const float* __restrict__ x,
                                                          const float* __restrict__ y,
                                                          const float* __restrict__ z,
                                                          float* __restrict__ slices,
                                                          const int* __restrict__ order) {
  const int r = rotation_of(order, logical_of_block(gridDim.x), gridDim.x);
  float e[9];
  rot_matrix(&rot[4 * (size_t)r], e);
  const float half = (float)S / 2.0f + 0.5f;
  const int N2d = S * S;
  float* out = slices + (size_t)r * N2d;
#pragma unroll 4
  for (int t = threadIdx.x; t < N2d; t += blockDim.x) {
    const int p = expand_pixel(t, S);
    float ux, uy, uz;
    voxel_coord(e, half, x[p], y[p], z[p], ux, uy, uz);
    float acc = 0.0f;
#pragma unroll
    for (int c = 0; c < 8; c++) {
      int ix, iy, iz;
      float w;
      if (tri_corner_xyz(c, ux, uy, uz, S, ix, iy, iz, w)) acc += w * model_at<BRICK>(model, ix, iy, iz, S);
    }
    out[p] = acc;
  }
}

void launch_expand(const float* model, int S, const float* rot, int R, const float* x,
                   const float* y, const float* z, float* slices, int interp, hipStream_t st, bool bricked) {
  if (R <= 0) return;
  const int* ord = rotation_order(rot, R, st);
  if (failed()) return;
  if (interp == INTERP_TRILINEAR) {
    if (bricked)
      hipLaunchKernelGGL(k_expand_trilinear<true>, dim3(R), dim3(256), 0, st, model, S, rot, x, y, z, slices, ord);
    else
      hipLaunchKernelGGL(k_expand_trilinear<false>, dim3(R), dim3(256), 0, st, model, S, rot, x, y, z, slices, ord);
  } else {
    if (bricked)
      hipLaunchKernelGGL(k_expand_nearest<true>, dim3(R), dim3(256), 0, st, model, S, rot, x, y, z, slices, ord);
    else
      hipLaunchKernelGGL(k_expand_nearest<false>, dim3(R), dim3(256), 0, st, model, S, rot, x, y, z, slices, ord);
  }
  EMC_LAUNCHED();
}

// ---------------------------------------------------------------------------
// Pattern packing: gather the masked pixels of every pattern into a dense,
// padded row (the GEMM operand), and sum_p m K^2 in double.
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_images_stats(const float* __restrict__ images,
                                                      const int* __restrict__ pix, int N, int N2d,
                                                      int n_mp, unsigned int* flags) {
  const int i = blockIdx.x;
  unsigned int inexact = 0, maxbits = 0;
  for (int k = threadIdx.x; k < n_mp; k += blockDim.x) {
    const int p = pix[k];
    if (p < 0) continue;
    const float v = images[(size_t)i * N2d + p];
    const float a = fabsf(v);
    if (!(v == rintf(v)) || !(a <= 2048.0f)) inexact = 1;
    if (a == a) maxbits = max(maxbits, __float_as_uint(a));
  }
  inexact = block_max_u32(inexact);
  maxbits = block_max_u32(maxbits);
  if (threadIdx.x == 0) {
    if (inexact) atomicOr(&flags[0], 1u);
    atomicMax(&flags[1], maxbits);
  }
}

void launch_images_stats(const float* images, const int* pix, int N, int N2d, int n_mp,
                         unsigned int* flags, hipStream_t st) {
  hipLaunchKernelGGL(k_images_stats, dim3(N), dim3(256), 0, st, images, pix, N, N2d, n_mp, flags);
  EMC_LAUNCHED();
}

__global__ __launch_bounds__(256) void k_pack_images_f32(const float* __restrict__ images,
                                                         const int* __restrict__ pix, int N, int N2d,
                                                         int n_mp, float* __restrict__ K32,
                                                         double* __restrict__ B, int poisson) {
  const int i = blockIdx.x;  // 0 .. N_pad-1
  double acc = 0.0;
  for (int k = threadIdx.x; k < n_mp; k += blockDim.x) {
    const int p = pix[k];
    float v = 0.0f;
    if (i < N && p >= 0) v = images[(size_t)i * N2d + p];
    K32[(size_t)i * n_mp + k] = v;
    acc += poisson ? (double)v : (double)v * v;
  }
  acc = block_sum_f64(acc);
  if (threadIdx.x == 0 && i < N) B[i] = acc;
}

__global__ __launch_bounds__(256) void k_pack_images_f16(const float* __restrict__ images,
                                                         const int* __restrict__ pix, int N, int N2d,
                                                         int n_mp, float kscale, int split,
                                                         _Float16* __restrict__ K16,
                                                         _Float16* __restrict__ K16lo,
                                                         double* __restrict__ B, int poisson) {
  const int i = blockIdx.x;
  double acc = 0.0;
  for (int k = threadIdx.x; k < n_mp; k += blockDim.x) {
    const int p = pix[k];
    float v = 0.0f;
    if (i < N && p >= 0) v = images[(size_t)i * N2d + p];
    const float s = v * kscale;
    const _Float16 hi = (_Float16)s;
    K16[(size_t)i * n_mp + k] = hi;
    if (split) K16lo[(size_t)i * n_mp + k] = (_Float16)(s - (float)hi);
    acc += poisson ? (double)v : (double)v * v;
  }
  acc = block_sum_f64(acc);
  if (threadIdx.x == 0 && i < N) B[i] = acc;
}

// out[c][r] = in[r][c] for an rows x cols matrix, 64x64 tiles through LDS.
template <class T>
__global__ __launch_bounds__(256) void k_transpose(const T* __restrict__ in, T* __restrict__ out,
                                                   int rows, int cols) {
  __shared__ T tile[64][65];
  const int r0 = blockIdx.y * 64, c0 = blockIdx.x * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;  // 64 x 4
  for (int j = ty; j < 64; j += 4) {
    const int r = r0 + j, c = c0 + tx;
    tile[j][tx] = (r < rows && c < cols) ? in[(size_t)r * cols + c] : T(0);
  }
  __syncthreads();
  for (int j = ty; j < 64; j += 4) {
    const int c = c0 + j, r = r0 + tx;
    if (c < cols && r < rows) out[(size_t)c * rows + r] = tile[tx][j];
  }
}

template <class T>
static void transpose(const T* in, T* out, int rows, int cols, hipStream_t st) {
  dim3 grid((cols + 63) / 64, (rows + 63) / 64);
  hipLaunchKernelGGL(k_transpose<T>, grid, dim3(256), 0, st, in, out, rows, cols);
  EMC_LAUNCHED();
}

void launch_pack_images_f32(const float* images, const int* pix, int N, int N2d, int n_mp,
                            int N_pad, float* K32, float* Kt32, double* B, hipStream_t st, bool poisson) {
  hipLaunchKernelGGL(k_pack_images_f32, dim3(N_pad), dim3(256), 0, st, images, pix, N, N2d, n_mp, K32, B,
                     (int)poisson);
  EMC_LAUNCHED();
  transpose<float>(K32, Kt32, N_pad, n_mp, st);
}

void launch_pack_images_f16(const float* images, const int* pix, int N, int N2d, int n_mp,
                            int N_pad, float kscale, bool split, _Float16* K16, _Float16* K16lo,
                            _Float16* Kt16, _Float16* Kt16lo, double* B, hipStream_t st, bool poisson) {
  hipLaunchKernelGGL(k_pack_images_f16, dim3(N_pad), dim3(256), 0, st, images, pix, N, N2d, n_mp,
                     kscale, (int)split, K16, K16lo, B, (int)poisson);
  EMC_LAUNCHED();
  transpose<_Float16>(K16, Kt16, N_pad, n_mp, st);
  if (split) transpose<_Float16>(K16lo, Kt16lo, N_pad, n_mp, st);
}

// ---------------------------------------------------------------------------
// Slice packing: compact masked columns, A_r = sum_p m S^2 in double.  For
// the Poisson likelihood the operand is log(max(S, kSlogFloor)) and
// A_r = sum_p m S.
// ---------------------------------------------------------------------------

__device__ __forceinline__ float slice_operand(float v, int poisson) {
  return poisson ? logf(fmaxf(v, kSlogFloor)) : v;
}

__global__ __launch_bounds__(256) void k_pack_slices_f32(const float* __restrict__ slices,
                                                         const int* __restrict__ pix, int R, int N2d,
                                                         int n_mp, float* __restrict__ S32,
                                                         double* __restrict__ A, int poisson) {
  const int r = blockIdx.x;
  double acc = 0.0;
  for (int k = threadIdx.x; k < n_mp; k += blockDim.x) {
    const int p = pix[k];
    float v = 0.0f, o = 0.0f;
    if (r < R && p >= 0) {
      v = slices[(size_t)r * N2d + p];
      o = slice_operand(v, poisson);
    }
    S32[(size_t)r * n_mp + k] = o;
    acc += poisson ? (double)v : (double)v * v;
  }
  acc = block_sum_f64(acc);
  if (threadIdx.x == 0 && r < R) A[r] = acc;
}

// Per-rotation power-of-two scale so that max |S| * scale <= 2^14, then the
// FP16 pair hi = fp16(S*scale), lo = fp16(S*scale - hi).  Launched with the
// fused kernel's thread count: A_r is summed in the same order.
__global__ __launch_bounds__(kPackThreads) void k_pack_slices_f16(const float* __restrict__ slices,
                                                         const int* __restrict__ pix, int R, int N2d,
                                                         int n_mp, _Float16* __restrict__ S16,
                                                         _Float16* __restrict__ S16lo,
                                                         float* __restrict__ scale,
                                                         double* __restrict__ A, int poisson) {
  const int r = blockIdx.x;
  const bool live = r < R;
  float mx = 0.0f;
  if (live)
    for (int k = threadIdx.x; k < n_mp; k += blockDim.x) {
      const int p = pix[k];
      if (p >= 0) mx = fmaxf(mx, fabsf(slice_operand(slices[(size_t)r * N2d + p], poisson)));
    }
  mx = block_max_f32(mx);
  float sc = 1.0f;
  if (mx > 0.0f && mx < 3.0e38f) {
    int e;
    frexpf(mx, &e);  // mx < 2^e
    sc = ldexpf(1.0f, max(-120, min(120, 14 - e)));
  }
  double acc = 0.0;
  for (int k = threadIdx.x; k < n_mp; k += blockDim.x) {
    const int p = pix[k];
    float v = 0.0f, o = 0.0f;
    if (live && p >= 0) {
      v = slices[(size_t)r * N2d + p];
      o = slice_operand(v, poisson);
    }
    const float s = o * sc;
    const _Float16 hi = (_Float16)s;
    S16[(size_t)r * n_mp + k] = hi;
    S16lo[(size_t)r * n_mp + k] = (_Float16)(s - (float)hi);
    acc += poisson ? (double)v : (double)v * v;
  }
  acc = block_sum_f64(acc);
  if (threadIdx.x == 0 && live) {
    A[r] = acc;
    scale[r] = 1.0f / sc;  // exact: sc is a power of two
  }
}

// Expand and FP16 packing in one pass: one workgroup per rotation computes
// its slice (written to slices, the caller's output, and kept in LDS), then
// packs the masked pixels from LDS: operand (log for Poisson) once into
// registers, the row maximum, the hi/lo pair and A_r.  Needs N2d floats of
// LDS and n_mp <= kPackThreads * kPackPer.

template <bool TRI, bool BRICK>
__global__ __launch_bounds__(kPackThreads) void k_expand_pack_f16(const float* __restrict__ model, int S,
                                                         const float* __restrict__ rot,
                                                         const float* __restrict__ x, const float* __restrict__ y,
                                                         const float* __restrict__ z, float* __restrict__ slices,
                                                         const int* __restrict__ pix, int R, int n_mp,
                                                         _Float16* __restrict__ S16, _Float16* __restrict__ S16lo,
                                                         float* __restrict__ scale, double* __restrict__ A,
                                                         int poisson, const int* __restrict__ order, int rows) {
  extern __shared__ float sv[];
  const int r = rotation_of(order, logical_of_block(gridDim.x), R);
  const bool live = r < R;
  const int N2d = S * S;
  if (live) {
    float e[9];
    rot_matrix(&rot[4 * (size_t)r], e);
    const float half = (float)S / 2.0f + 0.5f;
    float* out = slices + (size_t)r * N2d;
  #pragma unroll 4
  for (int t = threadIdx.x; t < N2d; t += blockDim.x) {
      const int p = rows ? t : expand_pixel(t, S);
      float ux, uy, uz;
      voxel_coord(e, half, x[p], y[p], z[p], ux, uy, uz);
      float v = 0.0f;
      if (TRI) {
#pragma unroll
        for (int c = 0; c < 8; c++) {
          int ix, iy, iz;
          float w;
          if (tri_corner_xyz(c, ux, uy, uz, S, ix, iy, iz, w)) v += w * model_at<BRICK>(model, ix, iy, iz, S);
        }
      } else {
        const int vx = (int)roundf(ux), vy = (int)roundf(uy), vz = (int)roundf(uz);
        if (vx > 0 && vx < S && vy > 0 && vy < S && vz > 0 && vz < S) v = model_at<BRICK>(model, vx, vy, vz, S);
      }
      out[p] = v;
      sv[p] = v;
    }
  }
  __syncthreads();
  float o[kPackPer];
  float mx = 0.0f;
  double acc = 0.0;
#pragma unroll
  for (int q = 0; q < kPackPer; q++) {
    const int k = threadIdx.x + q * kPackThreads;
    o[q] = 0.0f;
    if (live && k < n_mp) {
      const int p = pix[k];
      if (p >= 0) {
        const float v = sv[p];
        o[q] = slice_operand(v, poisson);
        mx = fmaxf(mx, fabsf(o[q]));
        acc += poisson ? (double)v : (double)v * v;
      }
    }
  }
  mx = block_max_f32(mx);
  float sc = 1.0f;
  if (mx > 0.0f && mx < 3.0e38f) {
    int ex;
    frexpf(mx, &ex);  // mx < 2^ex
    sc = ldexpf(1.0f, max(-120, min(120, 14 - ex)));
  }
#pragma unroll
  for (int q = 0; q < kPackPer; q++) {
    const int k = threadIdx.x + q * kPackThreads;
    if (k < n_mp) {
      const float sx = o[q] * sc;
      const _Float16 hi = (_Float16)sx;
      S16[(size_t)r * n_mp + k] = hi;
      S16lo[(size_t)r * n_mp + k] = (_Float16)(sx - (float)hi);
    }
  }
  acc = block_sum_f64(acc);
  if (threadIdx.x == 0 && live) {
    A[r] = acc;
    scale[r] = 1.0f / sc;  // exact: sc is a power of two
  }
}

// Slices too large for LDS: the row scale is not known until the whole
// slice is computed, so the pair is written with a provisional scale (the
// rotation's scale from the previous pack, prov_inv = 1 / scale, 0 when
// there is none) and the row is flagged when its own scale differs;
// k_pack_fix then repacks the flagged rows from the FP32 slice.  The pairs
// are the separate pack's either way (sc is a power of two).
//
// A workgroup takes xr (option expand_rotations, 8) consecutive rotations on
// one kXT x kXT pixel tile:
// nearby orientations on a small patch gather from one small region of the
// model, which stays in cache across the block's rotations (a whole slice
// per workgroup, with hundreds of slices in flight per XCD, streamed the
// model from memory for every slice).  Per (rotation, tile) it leaves the
// maximum |operand| and the partial A_r; k_pack_tile_finish reduces them in
// tile order (A_r is reproducible), sets the scale and the redo flag and
// zeroes the padding entries.
constexpr int kXT = 64;

template <bool TRI, bool BRICK>
__global__ __launch_bounds__(256) void k_expand_pack_tile(const float* __restrict__ model, int S,
                                                          const float* __restrict__ rot,
                                                          const float* __restrict__ x, const float* __restrict__ y,
                                                          const float* __restrict__ z, float* __restrict__ slices,
                                                          const int* __restrict__ kidx, int R, int R_pad, int n_mp,
                                                          const float* __restrict__ prov_inv,
                                                          _Float16* __restrict__ S16, _Float16* __restrict__ S16lo,
                                                          float* __restrict__ part_max, double* __restrict__ part_A,
                                                          int poisson, int xr, const int* __restrict__ order) {
  const int ntx = (S + kXT - 1) / kXT, ntiles = ntx * ntx;
  const int tile = blockIdx.y, tx0 = (tile % ntx) * kXT, ty0 = (tile / ntx) * kXT;
  const int N2d = S * S;
  const float half = (float)S / 2.0f + 0.5f;
  for (int rr = 0; rr < xr; rr++) {
    const int lg = blockIdx.x * xr + rr;
    if (lg >= R_pad) break;  // uniform over the workgroup
    const int r = rotation_of(order, lg, R);
    const bool live = r < R;
    _Float16* hrow = S16 + (size_t)r * n_mp;
    _Float16* lrow = S16lo + (size_t)r * n_mp;
    float e[9];
    float sc0 = 1.0f;
    if (live) {
      rot_matrix(&rot[4 * (size_t)r], e);
      const float pv = prov_inv ? prov_inv[r] : 0.0f;
      if (pv > 0.0f) sc0 = 1.0f / pv;  // exact: pv is a power of two
    }
    float mx = 0.0f;
    double acc = 0.0;
#pragma unroll 4
    for (int t = threadIdx.x; t < kXT * kXT; t += 256) {
      const int px = tx0 + (t & (kXT - 1)), py = ty0 + t / kXT;
      if (px >= S || py >= S) continue;
      const int p = py * S + px;
      const int k = kidx[p];
      if (!live) {
        if (k >= 0) {
          hrow[k] = (_Float16)0.0f;
          lrow[k] = (_Float16)0.0f;
        }
        continue;
      }
      float ux, uy, uz;
      voxel_coord(e, half, x[p], y[p], z[p], ux, uy, uz);
      float v = 0.0f;
      if (TRI) {
#pragma unroll
        for (int c = 0; c < 8; c++) {
          int ix, iy, iz;
          float w;
          if (tri_corner_xyz(c, ux, uy, uz, S, ix, iy, iz, w)) v += w * model_at<BRICK>(model, ix, iy, iz, S);
        }
      } else {
        const int vx = (int)roundf(ux), vy = (int)roundf(uy), vz = (int)roundf(uz);
        if (vx > 0 && vx < S && vy > 0 && vy < S && vz > 0 && vz < S) v = model_at<BRICK>(model, vx, vy, vz, S);
      }
      slices[(size_t)r * N2d + p] = v;
      if (k >= 0) {
        const float o = slice_operand(v, poisson);
        mx = fmaxf(mx, fabsf(o));
        acc += poisson ? (double)v : (double)v * v;
        const float sx = o * sc0;
        const _Float16 hi = (_Float16)sx;
        hrow[k] = hi;
        lrow[k] = (_Float16)(sx - (float)hi);
      }
    }
    mx = block_max_f32(mx);
    acc = block_sum_f64(acc);
    if (threadIdx.x == 0 && live) {
      part_max[(size_t)r * ntiles + tile] = mx;
      part_A[(size_t)r * ntiles + tile] = acc;
    }
  }
}

// One workgroup per rotation: the tiles' maxima and A partials in tile order,
// the row scale, the redo flag against the provisional scale, and zero
// padding entries n_m .. n_mp - 1.
__global__ __launch_bounds__(64) void k_pack_tile_finish(const float* __restrict__ part_max,
                                                         const double* __restrict__ part_A, int ntiles, int R,
                                                         int n_m, int n_mp, float* __restrict__ scale,
                                                         double* __restrict__ A, int* __restrict__ redo,
                                                         int have_prov, _Float16* __restrict__ S16,
                                                         _Float16* __restrict__ S16lo) {
  const int r = blockIdx.x;
  for (int k = n_m + threadIdx.x; k < n_mp; k += 64) {
    S16[(size_t)r * n_mp + k] = (_Float16)0.0f;
    S16lo[(size_t)r * n_mp + k] = (_Float16)0.0f;
  }
  if (r >= R || threadIdx.x != 0) return;
  float mx = 0.0f;
  double acc = 0.0;
  for (int t = 0; t < ntiles; t++) {
    mx = fmaxf(mx, part_max[(size_t)r * ntiles + t]);
    acc += part_A[(size_t)r * ntiles + t];
  }
  float sc = 1.0f;
  if (mx > 0.0f && mx < 3.0e38f) {
    int ex;
    frexpf(mx, &ex);  // mx < 2^ex
    sc = ldexpf(1.0f, max(-120, min(120, 14 - ex)));
  }
  const float pv = have_prov ? scale[r] : 0.0f;
  const float sc0 = pv > 0.0f ? 1.0f / pv : 1.0f;
  A[r] = acc;
  scale[r] = 1.0f / sc;
  redo[r] = sc != sc0;
}

// Rows whose provisional scale was not theirs: the pair again from the FP32
// slice with the row's own scale.
__global__ __launch_bounds__(256) void k_pack_fix(const float* __restrict__ slices, const int* __restrict__ pix,
                                                  const int* __restrict__ redo, int R, int N2d, int n_m, int n_mp,
                                                  const float* __restrict__ scale_inv, _Float16* __restrict__ S16,
                                                  _Float16* __restrict__ S16lo, int poisson) {
  const int r = blockIdx.x;
  if (r >= R || !redo[r]) return;
  const float sc = 1.0f / scale_inv[r];  // exact: a power of two
  for (int k = threadIdx.x; k < n_m; k += 256) {
    const float sx = slice_operand(slices[(size_t)r * N2d + pix[k]], poisson) * sc;
    const _Float16 hi = (_Float16)sx;
    S16[(size_t)r * n_mp + k] = hi;
    S16lo[(size_t)r * n_mp + k] = (_Float16)(sx - (float)hi);
  }
}

static bool expand_pack_lds(int S, int n_mp) {
  return (size_t)S * S * 4 <= 65536 && n_mp <= kPackThreads * kPackPer;
}

bool expand_pack_fits(int S, int n_mp) {
  (void)S;
  (void)n_mp;
  return true;  // the LDS form, else the provisional-scale form
}

void launch_expand_pack_f16(const float* model, int S, const float* rot, int R, int R_pad, const float* x,
                            const float* y, const float* z, float* slices, const int* pix, const int* kidx, int n_m,
                            int n_mp, _Float16* S16, _Float16* S16lo, float* scale, double* A, int interp,
                            bool poisson, hipStream_t st, bool bricked) {
  const size_t lds = (size_t)S * S * sizeof(float);
  const dim3 g(R_pad), b(kPackThreads);
  const int po = (int)poisson;
  const int* ord = rotation_order(rot, R, st);
  if (failed()) return;
  if (!expand_pack_lds(S, n_mp) || engine().expand_tiled) {
    Engine& e = engine();
    const int ntx = (S + kXT - 1) / kXT, ntiles = ntx * ntx;
    if (!e.d_redo.ensure(sizeof(int) * (size_t)R_pad) ||
        !e.d_xpart_max.ensure(sizeof(float) * (size_t)R_pad * ntiles) ||
        !e.d_xpart_A.ensure(sizeof(double) * (size_t)R_pad * ntiles))
      return;
    // the previous pack's row scales, if they were for these rotations
    const float* prov = e.sscale_R == R ? scale : nullptr;
    const bool tri = interp == INTERP_TRILINEAR;
    const int xr = e.expand_rotations;
    const dim3 gt((R_pad + xr - 1) / xr, ntiles), gb(256);
    int* redo = e.d_redo.as<int>();
    float* pmx = e.d_xpart_max.as<float>();
    double* pA = e.d_xpart_A.as<double>();
    if (tri && bricked)
      hipLaunchKernelGGL((k_expand_pack_tile<true, true>), gt, gb, 0, st, model, S, rot, x, y, z, slices, kidx, R,
                         R_pad, n_mp, prov, S16, S16lo, pmx, pA, po, xr, ord);
    else if (tri)
      hipLaunchKernelGGL((k_expand_pack_tile<true, false>), gt, gb, 0, st, model, S, rot, x, y, z, slices, kidx, R,
                         R_pad, n_mp, prov, S16, S16lo, pmx, pA, po, xr, ord);
    else if (bricked)
      hipLaunchKernelGGL((k_expand_pack_tile<false, true>), gt, gb, 0, st, model, S, rot, x, y, z, slices, kidx, R,
                         R_pad, n_mp, prov, S16, S16lo, pmx, pA, po, xr, ord);
    else
      hipLaunchKernelGGL((k_expand_pack_tile<false, false>), gt, gb, 0, st, model, S, rot, x, y, z, slices, kidx, R,
                         R_pad, n_mp, prov, S16, S16lo, pmx, pA, po, xr, ord);
    EMC_LAUNCHED();
    hipLaunchKernelGGL(k_pack_tile_finish, dim3(R_pad), dim3(64), 0, st, pmx, pA, ntiles, R, n_m, n_mp, scale, A,
                       redo, (int)(prov != nullptr), S16, S16lo);
    EMC_LAUNCHED();
    hipLaunchKernelGGL(k_pack_fix, dim3(R), gb, 0, st, slices, pix, redo, R, S * S, n_m, n_mp, scale, S16, S16lo, po);
    EMC_LAUNCHED();
    e.sscale_R = R;
    return;
  }
  if (interp == INTERP_TRILINEAR) {
    if (bricked)
      hipLaunchKernelGGL((k_expand_pack_f16<true, true>), g, b, lds, st, model, S, rot, x, y, z, slices, pix, R, n_mp,
                         S16, S16lo, scale, A, po, ord, (int)engine().expand_rows);
    else
      hipLaunchKernelGGL((k_expand_pack_f16<true, false>), g, b, lds, st, model, S, rot, x, y, z, slices, pix, R, n_mp,
                         S16, S16lo, scale, A, po, ord, (int)engine().expand_rows);
  } else {
    if (bricked)
      hipLaunchKernelGGL((k_expand_pack_f16<false, true>), g, b, lds, st, model, S, rot, x, y, z, slices, pix, R, n_mp,
                         S16, S16lo, scale, A, po, ord, (int)engine().expand_rows);
    else
      hipLaunchKernelGGL((k_expand_pack_f16<false, false>), g, b, lds, st, model, S, rot, x, y, z, slices, pix, R, n_mp,
                         S16, S16lo, scale, A, po, ord, (int)engine().expand_rows);
  }
  EMC_LAUNCHED();
  engine().sscale_R = R;
}

void launch_pack_slices_f32(const float* slices, const int* pix, int R, int N2d, int n_mp,
                            int R_pad, float* S32, double* A, hipStream_t st, bool poisson) {
  hipLaunchKernelGGL(k_pack_slices_f32, dim3(R_pad), dim3(256), 0, st, slices, pix, R, N2d, n_mp, S32, A,
                     (int)poisson);
  EMC_LAUNCHED();
}

void launch_pack_slices_f16(const float* slices, const int* pix, int R, int N2d, int n_mp,
                            int R_pad, _Float16* S16, _Float16* S16lo, float* scale, double* A,
                            hipStream_t st, bool poisson) {
  hipLaunchKernelGGL(k_pack_slices_f16, dim3(R_pad), dim3(kPackThreads), 0, st, slices, pix, R, N2d, n_mp, S16,
                     S16lo, scale, A, (int)poisson);
  EMC_LAUNCHED();
  engine().sscale_R = R;
}

}  // namespace emc
