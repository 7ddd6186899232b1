// Compress: per-rotation totals, insertion of the updated slices into the
// model and weight accumulators, model normalisation and maximum.
#include <hip/hip_runtime.h>

#include <algorithm>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include "engine.hpp"
#include "geom.hpp"
#include "reduce.hpp"

namespace emc {

// T_r = sum_i P_ri (emc_atomic.cu:99-101, FP64 here) and, for the FP16
// M-step, a power-of-two row scale with max_i P_ri/phi_i * scale <= 2^14.
__global__ __launch_bounds__(256) void k_rowsum(const float* __restrict__ P, int R, int N,
                                                const float* __restrict__ invphi, double* __restrict__ T,
                                                float* __restrict__ Tf, float* __restrict__ rowscale, int want,
                                                const float* __restrict__ phi, float* __restrict__ Tphi,
                                                int* __restrict__ nnz) {
  const int r = blockIdx.x;
  const float* row = P + (size_t)r * N;
  double t = 0.0, tp = 0.0;
  float mx = 0.0f;
  int nz = 0;
  auto take = [&](float p, int i) {
    t += p;
    nz += p != 0.0f;
    if (phi) tp += (double)p * (double)phi[i];
    if (want) mx = fmaxf(mx, fabsf(p * invphi[i]));
  };
  // 16-byte loads when every row starts 16-byte aligned (N % 4 == 0)
  const int nv = ((N & 3) == 0 && ((uintptr_t)P & 15) == 0) ? N / 4 : 0;
  for (int v = threadIdx.x; v < nv; v += 256) {
    const float4 q = reinterpret_cast<const float4*>(row)[v];
    if (q.x == 0.0f && q.y == 0.0f && q.z == 0.0f && q.w == 0.0f) continue;
    take(q.x, 4 * v);
    take(q.y, 4 * v + 1);
    take(q.z, 4 * v + 2);
    take(q.w, 4 * v + 3);
  }
  for (int i = 4 * nv + threadIdx.x; i < N; i += 256) take(row[i], i);
  t = block_sum_f64(t);
  if (phi) tp = block_sum_f64(tp);
  if (want) mx = block_max_f32(mx);
  if (nnz) nz = (int)block_sum_f64((double)nz);
  if (threadIdx.x == 0) {
    T[r] = t;
    Tf[r] = (float)t;
    if (phi) Tphi[r] = (float)tp;
    if (nnz) nnz[r] = nz;
    if (want) {
      float sc = 1.0f;
      if (mx > 0.0f && mx < 3.0e38f) {
        int e;
        frexpf(mx, &e);
        sc = ldexpf(1.0f, max(-120, min(120, 14 - e)));
      }
      rowscale[r] = sc;
    }
  }
}

void launch_rowsum(const float* P, int R, int N, const float* invphi, double* T, float* Tf, float* rowscale,
                   bool want_scale, hipStream_t st, const float* phi, float* Tphi, int* nnz) {
  hipLaunchKernelGGL(k_rowsum, dim3(R), dim3(256), 0, st, P, R, N, invphi, T, Tf, rowscale, (int)want_scale, phi,
                     Tphi, nnz);
  EMC_LAUNCHED();
}

// den[i] = sum_r P[r][i] A[r] in FP64 (lanes along patterns: coalesced rows).
__global__ __launch_bounds__(256) void k_p_dot_a(const float* __restrict__ P, int R, int N,
                                                 const double* __restrict__ A, double* __restrict__ den) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= N) return;
  double s = 0.0;
  for (int r = 0; r < R; r++) s += (double)P[(size_t)r * N + i] * A[r];
  den[i] = s;
}

void launch_p_dot_a(const float* P, int R, int N, const double* A, double* den, hipStream_t st) {
  hipLaunchKernelGGL(k_p_dot_a, dim3((N + 255) / 256), dim3(256), 0, st, P, R, N, A, den);
  EMC_LAUNCHED();
}

// w_r T_r as the reference forms it: weights[i_slice] * total_respons
// (emc_atomic.cu:129), float.
__global__ void k_set_wT(const float* __restrict__ w, const float* __restrict__ Tf, int R, float* __restrict__ wT) {
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r < R) wT[r] = w[r] * Tf[r];
}

void launch_set_wT(const float* wts, const float* Tf, int R, float* wT, hipStream_t st) {
  hipLaunchKernelGGL(k_set_wT, dim3((R + 255) / 256), dim3(256), 0, st, wts, Tf, R, wT);
  EMC_LAUNCHED();
}

// Atomic insertion (emc_atomic.cu:17-73,115-133): for rotations with
// T_r > 1e-10 and pixels whose mask is 1, model[v] += w_r T_r S'_rp and
// weight[v] += w_r T_r at the nearest voxel (bounds 0 <= v < S), or spread
// over the trilinear footprint.  The order of the float additions is not
// fixed.  The inserted pixels (compact indices ins_k) come in 8 x 8 detector
// tiles, so the atomics of one wave land on a small patch of the rotated
// plane (few cache lines).
__global__ __launch_bounds__(256) void k_insert_atomic(const float* __restrict__ Sc, int n_mp, int n_ins,
                                                       const int* __restrict__ pix, const int* __restrict__ ins_k,
                                                       const float* __restrict__ wT, const float* __restrict__ Tf,
                                                       const float* __restrict__ rot, int S,
                                                       const float* __restrict__ x, const float* __restrict__ y,
                                                       const float* __restrict__ z, float* __restrict__ model,
                                                       float* __restrict__ weight, int interp) {
#pragma clang fp contract(off)
  const int r = blockIdx.x;
  if (!(Tf[r] > 1e-10f)) return;
  const float w = wT[r];
  float e[9];
  rot_matrix(&rot[4 * (size_t)r], e);
  const float half = (float)S / 2.0f + 0.5f;
  for (int j = threadIdx.x; j < n_ins; j += 256) {
    const int k = ins_k[j];
    const int p = pix[k];
    float ux, uy, uz;
    voxel_coord(e, half, x[p], y[p], z[p], ux, uy, uz);
    const float val = w * Sc[(size_t)r * n_mp + k];
    if (interp == INTERP_NEAREST) {
      const int vx = (int)roundf(ux), vy = (int)roundf(uy), vz = (int)roundf(uz);
      if (vx < 0 || vx >= S || vy < 0 || vy >= S || vz < 0 || vz >= S) continue;
      const size_t v = ((size_t)vz * S + vy) * S + vx;
      atomicAdd(&model[v], val);
      atomicAdd(&weight[v], w);
    } else {
#pragma unroll
      for (int c = 0; c < 8; c++) {
        int v;
        float cw;
        if (!tri_corner(c, ux, uy, uz, S, v, cw)) continue;
        atomicAdd(&model[v], cw * val);
        atomicAdd(&weight[v], cw * w);
      }
    }
  }
}

void launch_insert_atomic(const float* Sc, int n_mp, const int* pix, const int* ins_k, const float* wT,
                          const float* rot, int R, int S, const float* x, const float* y, const float* z,
                          float* model, float* weight, int interp, hipStream_t st) {
  Engine& eng = engine();
  if (eng.n_ins <= 0) return;
  hipLaunchKernelGGL(k_insert_atomic, dim3(R), dim3(256), 0, st, Sc, n_mp, eng.n_ins, pix, ins_k, wT,
                     eng.d_Tf.as<float>(), rot, S, x, y, z, model, weight, interp);
  EMC_LAUNCHED();
}

// ---------------------------------------------------------------------------
// Brick-owned insertion.  One workgroup owns an IB^3 brick of the model and a
// share (group g of `groups`) of the active rotations (T_r > 1e-10): it
// accumulates in LDS every contribution of its rotations whose voxel lies in
// the brick, then adds the brick to the model and weight.  The pixels that
// can reach the brick under rotation r are found from the brick's box mapped
// back to detector coordinates, p = R^T (u - h): the plane test (the box's
// range of p_z against the detector's z range) drops most rotations for
// bricks away from the centre, and the box's range of (p_x, p_y) bounds the
// detector pixels to test; each is mapped forward exactly as the reference
// maps it (rot_matrix, voxel_coord, roundf) and added when its voxel falls
// in the brick.
//
// "atomic" (8^3 bricks, 256 threads): LDS float atomics, so the order of the
// additions is open, as in the reference's insertion (emc_atomic.cu:56-66);
// with one group per brick the flush is a plain read-modify-write (no other
// workgroup touches those voxels), with several (small models: too few
// bricks to fill the chip) a global float atomic per voxel.
//
// "deterministic" (4^3 bricks, one wave, k_insert_brick_det below): the
// same culling, additions in the sequential insertion's order.
// ---------------------------------------------------------------------------

// Active rotations (T_r > 1e-10) into act[0 .. *n), in no particular order.
__global__ __launch_bounds__(256) void k_active_rotations(const float* __restrict__ Tf, int R, int* __restrict__ act,
                                                          int* __restrict__ n) {
  const int r = blockIdx.x * 256 + threadIdx.x;
  const bool on = r < R && Tf[r] > 1e-10f;
  const unsigned long long m = __ballot(on);
  const int lane = threadIdx.x & 63;
  int base = 0;
  if (lane == 0 && m) base = atomicAdd(n, __popcll(m));
  base = __shfl(base, 0, 64);
  if (on) act[base + __popcll(m & ((1ull << lane) - 1ull))] = r;
}

// The same list in increasing order: one workgroup, chunks of 1,024
// rotations, wave ballots and a prefix over the 16 waves.
__global__ __launch_bounds__(1024) void k_active_ordered(const float* __restrict__ Tf, int R, int* __restrict__ act,
                                                         int* __restrict__ n) {
  __shared__ int wcnt[16];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  int base = 0;
  for (int c0 = 0; c0 < R; c0 += 1024) {
    const int r = c0 + t;
    const bool on = r < R && Tf[r] > 1e-10f;
    const unsigned long long m = __ballot(on);
    if (lane == 0) wcnt[wave] = __popcll(m);
    __syncthreads();
    int before = 0, total = 0;
    for (int w = 0; w < 16; w++) {
      before += w < wave ? wcnt[w] : 0;
      total += wcnt[w];
    }
    if (on) act[base + before + __popcll(m & ((1ull << lane) - 1ull))] = r;
    base += total;
    __syncthreads();
  }
  if (t == 0) *n = base;
}

// zrange = (min, max) of z over the pixels with mask == 1
__global__ __launch_bounds__(256) void k_z_range(const float* __restrict__ z, const int* __restrict__ mask, int N2d,
                                                 float* __restrict__ zrange) {
  float lo = 3.0e38f, hi = -3.0e38f;
  for (int p = threadIdx.x; p < N2d; p += 256)
    if (mask[p] == 1) {
      lo = fminf(lo, z[p]);
      hi = fmaxf(hi, z[p]);
    }
  lo = -block_max_f32(-lo);
  hi = block_max_f32(hi);
  if (threadIdx.x == 0) {
    zrange[0] = lo;
    zrange[1] = hi;
  }
}

template <int IB, int NT>
__global__ __launch_bounds__(NT) void k_insert_brick(const float* __restrict__ Sc, int n_mp,
                                                     const int* __restrict__ kidx, const int* __restrict__ mask,
                                                     const int* __restrict__ act, const int* __restrict__ n_act_p,
                                                     int groups, const float* __restrict__ wT,
                                                     const float* __restrict__ rot, int S,
                                                     const float* __restrict__ x, const float* __restrict__ y,
                                                     const float* __restrict__ z, const float* __restrict__ zrange,
                                                     float* __restrict__ model, float* __restrict__ weight,
                                                     int interp) {
#pragma clang fp contract(off)
  constexpr int NV = IB * IB * IB;
  __shared__ float acc_m[NV], acc_w[NV];
  __shared__ int list[NT];
  __shared__ int nlist;
  const int t = threadIdx.x;
  const int nb = S / IB, nbricks = nb * nb * nb;
  const int brick = blockIdx.x % nbricks, g = blockIdx.x / nbricks;
  const int bx = brick % nb, by = (brick / nb) % nb, bz = brick / (nb * nb);
  const int x0 = bx * IB, y0 = by * IB, z0 = bz * IB;
  for (int l = t; l < NV; l += NT) {
    acc_m[l] = 0.0f;
    acc_w[l] = 0.0f;
  }
  const float half = (float)S / 2.0f + 0.5f;
  const bool tri = interp != INTERP_NEAREST;
  // continuous coordinates that land in the brick: round to [x0, x0 + IB)
  // (nearest) or have a trilinear corner there; kept 0.05 voxel wide
  const float lo = tri ? -1.0f : -0.5f, hi = tri ? (float)IB : (float)IB - 0.5f;
  const float hs = 0.5f * (hi - lo) + 0.05f;  // half-size of the box
  const float cx = (float)x0 + 0.5f * (lo + hi) - half, cy = (float)y0 + 0.5f * (lo + hi) - half,
              cz = (float)z0 + 0.5f * (lo + hi) - half;  // box centre relative to the origin
  const float zlo = zrange[0] - 0.05f, zhi = zrange[1] + 0.05f;
  const float dlo = 0.5f - 0.5f * (float)S - 0.05f, dhi = 0.5f * (float)S - 0.5f + 0.05f;  // detector extent
  const int n_act = *n_act_p;
  const int a0 = (int)((long long)n_act * g / groups), a1 = (int)((long long)n_act * (g + 1) / groups);
  for (int base = a0; base < a1; base += NT) {
    if (t == 0) nlist = 0;
    __syncthreads();
    bool pass = false;
    int r = 0;
    if (base + t < a1) {
      r = act[base + t];
      float e[9];
      rot_matrix(&rot[4 * (size_t)r], e);
      // p = R^T (u - h): column j of R gives detector axis j
      const float pz = e[2] * cx + e[5] * cy + e[8] * cz, rz = hs * (fabsf(e[2]) + fabsf(e[5]) + fabsf(e[8]));
      const float px = e[0] * cx + e[3] * cy + e[6] * cz, rx = hs * (fabsf(e[0]) + fabsf(e[3]) + fabsf(e[6]));
      const float py = e[1] * cx + e[4] * cy + e[7] * cz, ry = hs * (fabsf(e[1]) + fabsf(e[4]) + fabsf(e[7]));
      pass = pz - rz <= zhi && pz + rz >= zlo && px - rx <= dhi && px + rx >= dlo && py - ry <= dhi && py + ry >= dlo;
    }
    if (pass) list[atomicAdd(&nlist, 1)] = r;
    __syncthreads();
    const int n = nlist;
    for (int j = 0; j < n; j++) {
      const int r = list[j];
      float e[9];
      rot_matrix(&rot[4 * (size_t)r], e);
      const float px = e[0] * cx + e[3] * cy + e[6] * cz, rx = hs * (fabsf(e[0]) + fabsf(e[3]) + fabsf(e[6]));
      const float py = e[1] * cx + e[4] * cy + e[7] * cz, ry = hs * (fabsf(e[1]) + fabsf(e[4]) + fabsf(e[7]));
      // pixel index = detector coordinate + S/2 - 0.5
      const float off = 0.5f * (float)S - 0.5f;
      const int ix0 = max(0, (int)ceilf(px - rx + off)), ix1 = min(S - 1, (int)floorf(px + rx + off));
      const int iy0 = max(0, (int)ceilf(py - ry + off)), iy1 = min(S - 1, (int)floorf(py + ry + off));
      if (ix1 < ix0 || iy1 < iy0) continue;  // uniform over the workgroup
      const int w_ = ix1 - ix0 + 1, h_ = iy1 - iy0 + 1, npx = w_ * h_;
      const float w = wT[r];
      for (int k0 = 0; k0 < npx; k0 += NT) {  // uniform over the workgroup
        const int k = k0 + t;
        // this lane's contributions, one slot per trilinear corner (slot 0
        // for the nearest voxel): voxel in the brick (-1: none), model and
        // weight terms
        int lv[8];
        float vm[8], vw[8];
#pragma unroll
        for (int c = 0; c < 8; c++) {
          lv[c] = -1;
          vm[c] = 0.0f;
          vw[c] = 0.0f;
        }
        if (k < npx) {
          const int p = (iy0 + k / w_) * S + ix0 + k % w_;
          if (mask[p] == 1) {
            float ux, uy, uz;
            voxel_coord(e, half, x[p], y[p], z[p], ux, uy, uz);
            const float val = w * Sc[(size_t)r * n_mp + kidx[p]];
            if (!tri) {
              const int vx = (int)roundf(ux) - x0, vy = (int)roundf(uy) - y0, vz = (int)roundf(uz) - z0;
              if ((unsigned)vx < (unsigned)IB && (unsigned)vy < (unsigned)IB && (unsigned)vz < (unsigned)IB) {
                lv[0] = (vz * IB + vy) * IB + vx;
                vm[0] = val;
                vw[0] = w;
              }
            } else {
#pragma unroll
              for (int c = 0; c < 8; c++) {
                int gx, gy, gz;
                float cw;
                if (!tri_corner_xyz(c, ux, uy, uz, S, gx, gy, gz, cw)) continue;
                const int vx = gx - x0, vy = gy - y0, vz = gz - z0;
                if ((unsigned)vx < (unsigned)IB && (unsigned)vy < (unsigned)IB && (unsigned)vz < (unsigned)IB) {
                  lv[c] = (vz * IB + vy) * IB + vx;
                  vm[c] = cw * val;
                  vw[c] = cw * w;
                }
              }
            }
          }
        }
#pragma unroll
        for (int q = 0; q < 8; q++)
          if (lv[q] >= 0) {
            atomicAdd(&acc_m[lv[q]], vm[q]);
            atomicAdd(&acc_w[lv[q]], vw[q]);
          }
      }
    }
    __syncthreads();
  }
  for (int l = t; l < NV; l += NT) {
    const int vx = l % IB, vy = (l / IB) % IB, vz = l / (IB * IB);
    const size_t v = ((size_t)(z0 + vz) * S + (y0 + vy)) * S + (x0 + vx);
    if (acc_w[l] != 0.0f || acc_m[l] != 0.0f) {
      if (groups == 1) {  // this workgroup owns the voxel
        model[v] += acc_m[l];
        weight[v] += acc_w[l];
      } else {
        atomicAdd(&model[v], acc_m[l]);
        atomicAdd(&weight[v], acc_w[l]);
      }
    }
  }
}


// Deterministic form: one wave per IB^3 brick, starting from the model's
// values.  The rotations that can reach the brick are listed 64 at a time in
// increasing order, with their matrices, pixel boxes and weights in LDS.
// Work items are (listed rotation, chunk of 64 of its box's pixels, taken
// x-major); NB items at a time have their pixels' contributions gathered,
// every item's loads issued before any is waited for, so a brick that many
// rotations reach pays one memory latency per NB items.  The items are then
// added in order, and within one item lanes whose voxels collide add one at
// a time in pixel order (nearest: the lowest pixel of each voxel, found with
// one LDS atomicMax of stamped keys, adds, then the rest lane by lane;
// trilinear: rounds in which the lowest pixel index holding a voxel, found
// with an LDS atomicMin, adds).  Every voxel receives the float
// additions of the sequential insertion (emc.c:491-539 per rotation) in the
// same order.
template <int IB, int NB, bool TRI>
__global__ __launch_bounds__(64) void k_insert_brick_det(const float* __restrict__ Sc, int n_mp,
                                                         const int* __restrict__ kidx, const int* __restrict__ mask,
                                                         const int* __restrict__ act, const int* __restrict__ n_act_p,
                                                         const float* __restrict__ wT,
                                                         const float* __restrict__ rot, int S,
                                                         const float* __restrict__ x, const float* __restrict__ y,
                                                         const float* __restrict__ z,
                                                         const float* __restrict__ zrange,
                                                         float* __restrict__ model, float* __restrict__ weight) {
#pragma clang fp contract(off)
  constexpr int NV = IB * IB * IB, NC = TRI ? 8 : 1;
  __shared__ float acc_m[NV], acc_w[NV];
  __shared__ int own[NV];
  __shared__ float le[9][64];  // listed rotations: matrix entries
  __shared__ int4 lbox[64];    // first pixel column, first row, width, pixel count
  __shared__ int lr[64];
  __shared__ float lw[64];
  const int t = threadIdx.x;
  const int nb = S / IB;
  const int brick = blockIdx.x;
  const int x0 = (brick % nb) * IB, y0 = ((brick / nb) % nb) * IB, z0 = (brick / (nb * nb)) * IB;
  for (int l = t; l < NV; l += 64) {
    const int vx = l % IB, vy = (l / IB) % IB, vz = l / (IB * IB);
    const size_t v = ((size_t)(z0 + vz) * S + (y0 + vy)) * S + (x0 + vx);
    acc_m[l] = model[v];
    acc_w[l] = weight[v];
    own[l] = TRI ? 0x7fffffff : -1;
  }
  int seq = 0;  // items added so far (nearest: key stamps)
  const float half = (float)S / 2.0f + 0.5f;
  const float lo = TRI ? -1.0f : -0.5f, hi = TRI ? (float)IB : (float)IB - 0.5f;
  const float hs = 0.5f * (hi - lo) + 0.05f;
  const float cx = (float)x0 + 0.5f * (lo + hi) - half, cy = (float)y0 + 0.5f * (lo + hi) - half,
              cz = (float)z0 + 0.5f * (lo + hi) - half;
  const float zlo = zrange[0] - 0.05f, zhi = zrange[1] + 0.05f;
  const float dlo = 0.5f - 0.5f * (float)S - 0.05f, dhi = 0.5f * (float)S - 0.5f + 0.05f;
  const float off = 0.5f * (float)S - 0.5f;
  const int n_act = *n_act_p;
  for (int base = 0; base < n_act; base += 64) {
    bool pass = false;
    int r = 0, ix0 = 0, iy0 = 0, w_ = 0, npx = 0;
    float e[9];
    if (base + t < n_act) {
      r = act[base + t];
      rot_matrix(&rot[4 * (size_t)r], e);
      const float pz = e[2] * cx + e[5] * cy + e[8] * cz, rz = hs * (fabsf(e[2]) + fabsf(e[5]) + fabsf(e[8]));
      const float px = e[0] * cx + e[3] * cy + e[6] * cz, rx = hs * (fabsf(e[0]) + fabsf(e[3]) + fabsf(e[6]));
      const float py = e[1] * cx + e[4] * cy + e[7] * cz, ry = hs * (fabsf(e[1]) + fabsf(e[4]) + fabsf(e[7]));
      if (pz - rz <= zhi && pz + rz >= zlo && px - rx <= dhi && px + rx >= dlo && py - ry <= dhi &&
          py + ry >= dlo) {
        ix0 = max(0, (int)ceilf(px - rx + off));
        iy0 = max(0, (int)ceilf(py - ry + off));
        const int ix1 = min(S - 1, (int)floorf(px + rx + off)), iy1 = min(S - 1, (int)floorf(py + ry + off));
        pass = ix1 >= ix0 && iy1 >= iy0;
        w_ = ix1 - ix0 + 1;
        npx = w_ * (iy1 - iy0 + 1);
      }
    }
    const unsigned long long m = __ballot(pass);
    const int n = __popcll(m);
    __syncthreads();  // the previous list is no longer read
    if (pass) {
      const int i = __popcll(m & ((1ull << t) - 1ull));
#pragma unroll
      for (int c = 0; c < 9; c++) le[c][i] = e[c];
      lbox[i] = make_int4(ix0, iy0, w_, npx);
      lr[i] = r;
      lw[i] = wT[r];
    }
    __syncthreads();
    int j = 0, k0 = 0;  // next item: listed rotation j, its pixels k0 .. k0 + 63 (uniform)
    while (j < n) {
      int ij[NB], ik[NB];
#pragma unroll
      for (int b = 0; b < NB; b++) {
        ij[b] = j;
        ik[b] = k0 + t;
        if (j < n) {
          k0 += 64;
          if (k0 >= lbox[j].w) {
            j++;
            k0 = 0;
          }
        }
      }
      // gather: pixel, then its coordinates and column, then the slice value,
      // each level issued for all NB items before it is waited for
      int p[NB], msk[NB], col[NB];
      float px_[NB], py_[NB], pz_[NB], sv[NB];
#pragma unroll
      for (int b = 0; b < NB; b++) {
        const int4 bx = lbox[ij[b] < n ? ij[b] : 0];
        const bool in = ij[b] < n && ik[b] < bx.w;
        const int h_ = bx.w / bx.z;  // rows of the box (npx = width x rows)
        p[b] = in ? (bx.y + ik[b] % h_) * S + bx.x + ik[b] / h_ : -1;
      }
#pragma unroll
      for (int b = 0; b < NB; b++) {
        const int q = p[b] < 0 ? 0 : p[b];
        msk[b] = mask[q];
        col[b] = kidx[q];
        px_[b] = x[q];
        py_[b] = y[q];
        pz_[b] = z[q];
      }
#pragma unroll
      for (int b = 0; b < NB; b++) {
        if (p[b] < 0 || msk[b] != 1) p[b] = -1;
        const int jj = ij[b] < n ? ij[b] : 0;
        sv[b] = Sc[(size_t)lr[jj] * n_mp + (p[b] < 0 ? 0 : col[b])];
      }
      // add: item by item in order
#pragma unroll
      for (int b = 0; b < NB; b++) {
        if (ij[b] >= n) break;  // uniform
        const int jj = ij[b];
        float em[9];
#pragma unroll
        for (int c = 0; c < 9; c++) em[c] = le[c][jj];
        const float w = lw[jj];
        int lv[NC];
        float vm[NC], vw[NC];
#pragma unroll
        for (int c = 0; c < NC; c++) {
          lv[c] = -1;
          vm[c] = 0.0f;
          vw[c] = 0.0f;
        }
        if (p[b] >= 0) {
          float ux, uy, uz;
          voxel_coord(em, half, px_[b], py_[b], pz_[b], ux, uy, uz);
          const float val = w * sv[b];
          if (!TRI) {
            const int vx = (int)roundf(ux) - x0, vy = (int)roundf(uy) - y0, vz = (int)roundf(uz) - z0;
            if ((unsigned)vx < (unsigned)IB && (unsigned)vy < (unsigned)IB && (unsigned)vz < (unsigned)IB) {
              lv[0] = (vz * IB + vy) * IB + vx;
              vm[0] = val;
              vw[0] = w;
            }
          } else {
#pragma unroll
            for (int c = 0; c < NC; c++) {
              int gx, gy, gz;
              float cw;
              if (!tri_corner_xyz(c, ux, uy, uz, S, gx, gy, gz, cw)) continue;
              const int vx = gx - x0, vy = gy - y0, vz = gz - z0;
              if ((unsigned)vx < (unsigned)IB && (unsigned)vy < (unsigned)IB && (unsigned)vz < (unsigned)IB) {
                lv[c] = (vz * IB + vy) * IB + vx;
                vm[c] = cw * val;
                vw[c] = cw * w;
              }
            }
          }
        }
        const int k = ik[b];
        if (!TRI) {
          // the lowest pixel of each voxel adds first (stamped keys: the
          // largest key of this item is its lowest pixel, and keys of
          // earlier items are smaller), then the others one lane at a time
          // in pixel order
          seq++;
          const int key = (seq << 6) | (63 - t);
          const int v = lv[0];
          if (v >= 0) atomicMax(&own[v], key);
          __syncthreads();
          const bool win = v >= 0 && own[v] == key;
          if (win) {
            acc_m[v] += vm[0];
            acc_w[v] += vw[0];
          }
          unsigned long long lose = __ballot(v >= 0 && !win);
          while (lose) {  // uniform
            __syncthreads();
            const int l = __ffsll((long long)lose) - 1;
            lose &= lose - 1ull;
            if (t == l) {
              acc_m[v] += vm[0];
              acc_w[v] += vw[0];
            }
          }
          __syncthreads();
          continue;
        }
        // rounds: the lowest pixel index holding a voxel adds to it (a
        // pixel's corners are distinct voxels)
        unsigned pend = 0;
#pragma unroll
        for (int q = 0; q < NC; q++) pend |= (lv[q] >= 0 ? 1u : 0u) << q;
        for (;;) {
#pragma unroll
          for (int q = 0; q < NC; q++)
            if ((pend >> q) & 1u) atomicMin(&own[lv[q]], k);
          __syncthreads();
          unsigned mine = 0;
#pragma unroll
          for (int q = 0; q < NC; q++)
            if (((pend >> q) & 1u) && own[lv[q]] == k) {
              acc_m[lv[q]] += vm[q];
              acc_w[lv[q]] += vw[q];
              mine |= 1u << q;
            }
          pend &= ~mine;
          const int more = __syncthreads_or(pend != 0u);
#pragma unroll
          for (int q = 0; q < NC; q++)
            if ((mine >> q) & 1u) own[lv[q]] = 0x7fffffff;
          __syncthreads();
          if (!more) break;
        }
      }
    }
  }
  __syncthreads();
  for (int l = t; l < NV; l += 64) {
    const int vx = l % IB, vy = (l / IB) % IB, vz = l / (IB * IB);
    const size_t v = ((size_t)(z0 + vz) * S + (y0 + vy)) * S + (x0 + vx);
    model[v] = acc_m[l];
    weight[v] = acc_w[l];
  }
}

constexpr int kIB = 8;   // atomic form: brick edge (voxels), 256 threads
constexpr int kIBd = 4;  // deterministic form: brick edge, one wave

void launch_insert_brick(const float* Sc, int n_mp, const int* kidx, const int* mask, const float* wT,
                         const float* rot, int R, int S, const float* x, const float* y, const float* z,
                         float* model, float* weight, int interp, bool deterministic, hipStream_t st) {
  Engine& eng = engine();
  if (!eng.d_act.ensure(sizeof(int) * ((size_t)R + 4)) || !eng.d_zrange.ensure(16)) return;
  int* act = eng.d_act.as<int>();
  int* n_act = act + R;
  if (deterministic) {
    hipLaunchKernelGGL(k_active_ordered, dim3(1), dim3(1024), 0, st, eng.d_Tf.as<float>(), R, act, n_act);
  } else {
    launch_fill(reinterpret_cast<float*>(n_act), 1, 0.0f, st);
    hipLaunchKernelGGL(k_active_rotations, dim3((R + 255) / 256), dim3(256), 0, st, eng.d_Tf.as<float>(), R, act,
                       n_act);
  }
  EMC_LAUNCHED();
  auto& zk = eng.zrange_key;  // the detector's z range, once per (coordinates, mask)
  if (!(zk.z == z && zk.mask == mask && zk.S == S && zk.stamp == eng.invalidate_stamp)) {
    hipLaunchKernelGGL(k_z_range, dim3(1), dim3(256), 0, st, z, mask, S * S, eng.d_zrange.as<float>());
    EMC_LAUNCHED();
    zk.z = z;
    zk.mask = mask;
    zk.S = S;
    zk.stamp = eng.invalidate_stamp;
  }
  if (deterministic) {
    const int nb = S / kIBd;
    if (interp == INTERP_NEAREST)
      hipLaunchKernelGGL((k_insert_brick_det<kIBd, 8, false>), dim3((unsigned)nb * nb * nb), dim3(64), 0, st, Sc,
                         n_mp, kidx, mask, act, n_act, wT, rot, S, x, y, z, eng.d_zrange.as<float>(), model, weight);
    else
      hipLaunchKernelGGL((k_insert_brick_det<kIBd, 4, true>), dim3((unsigned)nb * nb * nb), dim3(64), 0, st, Sc,
                         n_mp, kidx, mask, act, n_act, wT, rot, S, x, y, z, eng.d_zrange.as<float>(), model, weight);
  } else {
    // rotation groups per brick: enough workgroups to fill the chip
    const int nb = S / kIB, nbricks = nb * nb * nb;
    const int groups = nbricks >= 4096 ? 1 : std::min(64, (4096 + nbricks - 1) / nbricks);
    hipLaunchKernelGGL((k_insert_brick<kIB, 256>), dim3((unsigned)(nbricks * groups)), dim3(256), 0, st, Sc,
                       n_mp, kidx, mask, act, n_act, groups, wT, rot, S, x, y, z, eng.d_zrange.as<float>(), model,
                       weight, interp);
  }
  EMC_LAUNCHED();
}

// The deterministic nearest form stamps its keys with the item count of a
// brick (at most one item per rotation, 64 pixels): R < 2^24 keeps them in
// an int.
bool insert_brick_fits(int S, int R, bool deterministic) {
  return deterministic ? S % kIBd == 0 && R < (1 << 24) : S % kIB == 0;
}

// ---------------------------------------------------------------------------
// Deterministic insertion.  The voxel hit by (rotation r, pixel t) depends
// only on geometry, so the inverse map voxel -> entries is built once: entry
// e = r * n_ins + t (t enumerates inserted pixels x-major, the reference's
// loop order), keyed by its voxel and stably sorted, which leaves every
// voxel's entries in increasing e.  Each iteration then sums every voxel's
// contributions sequentially in that order -- no atomics, the same float
// additions in the same order on every run (and the order of the reference's
// sequential CPU insertion, emc.c:491-539 called per rotation).
// Trilinear entries carry (e, corner) as 8 e + c.
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_map_keys(const int* __restrict__ ins_pix, int n_ins, int R,
                                                  const float* __restrict__ rot, int S, const float* __restrict__ x,
                                                  const float* __restrict__ y, const float* __restrict__ z, int interp,
                                                  uint32_t sentinel, uint32_t* __restrict__ keys) {
  const size_t e_id = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (e_id >= (size_t)R * n_ins) return;
  const int r = (int)(e_id / n_ins);
  const int t = (int)(e_id % n_ins);
  float e[9];
  rot_matrix(&rot[4 * (size_t)r], e);
  const float half = (float)S / 2.0f + 0.5f;
  const int p = ins_pix[t];
  float ux, uy, uz;
  voxel_coord(e, half, x[p], y[p], z[p], ux, uy, uz);
  const size_t ent = (size_t)r * n_ins + t;
  if (interp == INTERP_NEAREST) {
    const int vx = (int)roundf(ux), vy = (int)roundf(uy), vz = (int)roundf(uz);
    uint32_t key = sentinel;
    if (vx >= 0 && vx < S && vy >= 0 && vy < S && vz >= 0 && vz < S) key = (uint32_t)((vz * S + vy) * S + vx);
    keys[ent] = key;
  } else {
#pragma unroll
    for (int c = 0; c < 8; c++) {
      int v;
      float cw;
      keys[ent * 8 + c] = tri_corner(c, ux, uy, uz, S, v, cw) ? (uint32_t)v : sentinel;
    }
  }
}

__global__ void k_map_count(const uint32_t* __restrict__ sorted_keys, size_t n, uint32_t nvox, uint32_t* __restrict__ off) {
  // off[v] = first position with key >= v, for v in [0, nvox]; keys sorted.
  const size_t j = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (j >= n) return;
  const uint32_t k = sorted_keys[j];
  const uint32_t prev = j == 0 ? 0u : min(sorted_keys[j - 1], nvox) + 1u;
  const uint32_t hi = min(k, nvox);
  if (j == 0)
    for (uint32_t v = 0; v <= hi; v++) off[v] = 0;
  else
    for (uint32_t v = prev; v <= hi; v++) off[v] = (uint32_t)j;
}

__global__ void k_map_tail(const uint32_t* __restrict__ sorted_keys, size_t n, uint32_t nvox, uint32_t* __restrict__ off) {
  // voxels above the largest key: off = n (or the sentinel start)
  const uint32_t last = n ? min(sorted_keys[n - 1], nvox) : 0u;
  const uint32_t from = n ? last + 1u : 0u;
  for (uint32_t v = from + threadIdx.x; v <= nvox; v += 256) off[v] = (uint32_t)n;
}

__global__ void k_map_weights(const uint32_t* __restrict__ ent, size_t n, const int* __restrict__ ins_pix, int n_ins,
                              const float* __restrict__ rot, int S, const float* __restrict__ x,
                              const float* __restrict__ y, const float* __restrict__ z, float* __restrict__ wt) {
  const size_t j = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (j >= n) return;
  const uint32_t code = ent[j];
  const uint32_t eidx = code >> 3;
  const int c = (int)(code & 7u);
  const int r = (int)(eidx / (uint32_t)n_ins), t = (int)(eidx % (uint32_t)n_ins);
  float e[9];
  rot_matrix(&rot[4 * (size_t)r], e);
  const float half = (float)S / 2.0f + 0.5f;
  const int p = ins_pix[t];
  float ux, uy, uz;
  voxel_coord(e, half, x[p], y[p], z[p], ux, uy, uz);
  int v;
  float cw = 0.0f;
  tri_corner(c, ux, uy, uz, S, v, cw);
  wt[j] = cw;
}

// Voxels whose entry list holds at least kHeavy entries, in no particular
// order (each is summed by one wave; the order of the voxels is immaterial).
constexpr uint32_t kHeavy = 64;
__global__ __launch_bounds__(256) void k_map_heavy(const uint32_t* __restrict__ off, uint32_t nvox,
                                                   uint32_t* __restrict__ heavy, uint32_t* __restrict__ n_heavy) {
  const uint32_t v = blockIdx.x * 256 + threadIdx.x;
  const bool h = v < nvox && off[v + 1] - off[v] >= kHeavy;
  const unsigned long long m = __ballot(h);
  const int lane = threadIdx.x & 63;
  uint32_t base = 0;
  if (lane == 0 && m) base = atomicAdd(n_heavy, (uint32_t)__popcll(m));
  base = __shfl(base, 0, 64);
  if (h) heavy[base + __popcll(m & ((1ull << lane) - 1ull))] = v;
}

void launch_build_map(const int* ins_pix, int n_ins, const float* rot, int R, int S, const float* x,
                      const float* y, const float* z, int interp, Engine& eng, hipStream_t st) {
  const size_t per = interp == INTERP_NEAREST ? 1 : 8;
  const size_t n = (size_t)R * n_ins * per;
  const uint32_t nvox = (uint32_t)S * S * S;
  eng.map_entries = 0;
  if (n == 0 || n >= 0xffffffffull) {
    if (n) set_error("insertion map: %zu entries exceed 32-bit entry ids", n);
    return;
  }
  DBuf keys, keys_sorted, temp;
  if (!keys.ensure(n * 4) || !keys_sorted.ensure(n * 4) || !eng.d_map_ent.ensure(n * 4) ||
      !eng.d_map_off.ensure(((size_t)nvox + 1) * 4)) {
    keys.release();
    keys_sorted.release();
    return;
  }
  hipLaunchKernelGGL(k_map_keys, dim3((unsigned)(((size_t)R * n_ins + 255) / 256)), dim3(256), 0, st, ins_pix, n_ins, R,
                     rot, S, x, y, z, interp, nvox, keys.as<uint32_t>());
  EMC_LAUNCHED();
  unsigned int bits = 1;
  while ((1ull << bits) <= nvox) bits++;
  rocprim::counting_iterator<uint32_t> ids(0u);
  size_t tbytes = 0;
  EMC_CHECK(rocprim::radix_sort_pairs(nullptr, tbytes, keys.as<uint32_t>(), keys_sorted.as<uint32_t>(), ids,
                                      eng.d_map_ent.as<uint32_t>(), n, 0, bits, st));
  if (!temp.ensure(tbytes)) {
    keys.release();
    keys_sorted.release();
    return;
  }
  EMC_CHECK(rocprim::radix_sort_pairs(temp.p, tbytes, keys.as<uint32_t>(), keys_sorted.as<uint32_t>(), ids,
                                      eng.d_map_ent.as<uint32_t>(), n, 0, bits, st));
  hipLaunchKernelGGL(k_map_count, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, keys_sorted.as<uint32_t>(), n,
                     nvox, eng.d_map_off.as<uint32_t>());
  EMC_LAUNCHED();
  hipLaunchKernelGGL(k_map_tail, dim3(1), dim3(256), 0, st, keys_sorted.as<uint32_t>(), n, nvox,
                     eng.d_map_off.as<uint32_t>());
  EMC_LAUNCHED();
  if (interp != INTERP_NEAREST) {
    if (eng.d_map_wt.ensure(n * 4)) {
      hipLaunchKernelGGL(k_map_weights, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                         eng.d_map_ent.as<uint32_t>(), n, ins_pix, n_ins, rot, S, x, y, z, eng.d_map_wt.as<float>());
      EMC_LAUNCHED();
    }
  }
  // voxels with at least kHeavy entries, gathered by a wave each
  if (!eng.d_map_heavy.ensure(sizeof(uint32_t) * ((size_t)nvox + 1))) {
    keys.release();
    keys_sorted.release();
    temp.release();
    return;
  }
  uint32_t* heavy = eng.d_map_heavy.as<uint32_t>();
  launch_fill(reinterpret_cast<float*>(heavy + nvox), 1, 0.0f, st);
  hipLaunchKernelGGL(k_map_heavy, dim3((nvox + 255) / 256), dim3(256), 0, st, eng.d_map_off.as<uint32_t>(), nvox,
                     heavy, heavy + nvox);
  EMC_LAUNCHED();
  uint32_t nh = 0;
  EMC_CHECK(hipMemcpyAsync(&nh, heavy + nvox, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  EMC_CHECK(hipStreamSynchronize(st));
  temp.release();
  keys.release();
  keys_sorted.release();
  eng.map_heavy = nh;
  eng.map_entries = n;
}

// One thread per voxel with fewer than kHeavy entries, contributions added
// sequentially in entry order, starting from the accumulator's current value
// (the caller may not have reset it).
__global__ __launch_bounds__(256) void k_insert_gather(const float* __restrict__ Sc, int n_mp,
                                                       const int* __restrict__ ins_col, int n_ins,
                                                       const float* __restrict__ wT, const float* __restrict__ Tf,
                                                       const uint32_t* __restrict__ off, const uint32_t* __restrict__ ent,
                                                       const float* __restrict__ ent_wt, int nvox, int tri,
                                                       float* __restrict__ model, float* __restrict__ weight) {
#pragma clang fp contract(off)
  const int v = blockIdx.x * 256 + threadIdx.x;
  if (v >= nvox) return;
  const uint32_t b = off[v], en = off[v + 1];
  if (b == en || en - b >= kHeavy) return;
  float m = model[v], wsum = weight[v];
  for (uint32_t j = b; j < en; j++) {
    const uint32_t code = ent[j];
    const uint32_t eidx = tri ? code >> 3 : code;
    const int r = (int)(eidx / (uint32_t)n_ins), t = (int)(eidx % (uint32_t)n_ins);
    if (!(Tf[r] > 1e-10f)) continue;
    const float w = wT[r];
    const float val = w * Sc[(size_t)r * n_mp + ins_col[t]];
    if (tri) {
      const float cw = ent_wt[j];
      m += cw * val;
      wsum += cw * w;
    } else {
      m += val;
      wsum += w;
    }
  }
  model[v] = m;
  weight[v] = wsum;
}

// One wave per voxel with at least kHeavy entries (near the centre every
// rotation's slice passes through a voxel, so its list is as long as the
// table): the wave loads 64 consecutive entries at once, then every lane
// adds the 64 terms in entry order, read lane by lane -- the same float
// additions in the same order as the single thread above, without a memory
// round trip per entry.
__global__ __launch_bounds__(256) void k_insert_gather_wave(const float* __restrict__ Sc, int n_mp,
                                                            const int* __restrict__ ins_col, int n_ins,
                                                            const float* __restrict__ wT,
                                                            const float* __restrict__ Tf,
                                                            const uint32_t* __restrict__ off,
                                                            const uint32_t* __restrict__ ent,
                                                            const float* __restrict__ ent_wt,
                                                            const uint32_t* __restrict__ heavy, uint32_t n_heavy,
                                                            int tri, float* __restrict__ model,
                                                            float* __restrict__ weight) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x & 63;
  const uint32_t h = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (h >= n_heavy) return;  // uniform over the wave
  const uint32_t v = heavy[h];
  const uint32_t b = off[v], en = off[v + 1];
  float m = model[v], wsum = weight[v];
  for (uint32_t j0 = b; j0 < en; j0 += 64) {
    const uint32_t j = j0 + lane;
    int on = 0;
    float dm = 0.0f, dw = 0.0f;
    if (j < en) {
      const uint32_t code = ent[j];
      const uint32_t eidx = tri ? code >> 3 : code;
      const int r = (int)(eidx / (uint32_t)n_ins), t = (int)(eidx % (uint32_t)n_ins);
      if (Tf[r] > 1e-10f) {
        const float w = wT[r];
        const float val = w * Sc[(size_t)r * n_mp + ins_col[t]];
        on = 1;
        if (tri) {
          const float cw = ent_wt[j];
          dm = cw * val;
          dw = cw * w;
        } else {
          dm = val;
          dw = w;
        }
      }
    }
    // uniform: every lane adds the same terms, the active entries only, in
    // increasing lane order (most rotations are inactive once the posterior
    // is sharp, so the serial chain follows the active entries)
    unsigned long long act = __ballot(on);
    while (act) {
      const int l = __ffsll((long long)act) - 1;
      act &= act - 1ull;
      m += __int_as_float(__builtin_amdgcn_readlane(__float_as_int(dm), l));
      wsum += __int_as_float(__builtin_amdgcn_readlane(__float_as_int(dw), l));
    }
  }
  if (lane == 0) {
    model[v] = m;
    weight[v] = wsum;
  }
}

void launch_insert_gather(const float* Sc, int n_mp, const int* ins_col, int n_ins, const float* wT,
                          const uint32_t* off, const uint32_t* ent, const float* ent_wt, int S, int interp,
                          float* model, float* weight, hipStream_t st) {
  const int nvox = S * S * S;
  Engine& eng = engine();
  hipLaunchKernelGGL(k_insert_gather, dim3((nvox + 255) / 256), dim3(256), 0, st, Sc, n_mp, ins_col, n_ins, wT,
                     eng.d_Tf.as<float>(), off, ent, ent_wt, nvox, (int)(interp != INTERP_NEAREST), model, weight);
  EMC_LAUNCHED();
  if (eng.map_heavy > 0) {
    hipLaunchKernelGGL(k_insert_gather_wave, dim3((eng.map_heavy + 3) / 4), dim3(256), 0, st, Sc, n_mp, ins_col,
                       n_ins, wT, eng.d_Tf.as<float>(), off, ent, ent_wt, eng.d_map_heavy.as<uint32_t>(),
                       (uint32_t)eng.map_heavy, (int)(interp != INTERP_NEAREST), model, weight);
    EMC_LAUNCHED();
  }
}

// ---------------------------------------------------------------------------
// Normalisation (emc_cuda.cu:414-433): model = weight > 0 ? model/weight : 0,
// then model *= 1 / (sum(model) / n).  The sum is FP64 in a fixed order.
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_norm_divide(float* __restrict__ model, const float* __restrict__ weight, int n,
                                                     double* __restrict__ red) {
  double s = 0.0;
  for (int k = blockIdx.x * 256 + threadIdx.x; k < n; k += gridDim.x * 256) {
    const float w = weight[k];
    const float m = w > 0.0f ? model[k] / w : 0.0f;
    model[k] = m;
    s += m;
  }
  s = block_sum_f64(s);
  if (threadIdx.x == 0) red[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_norm_factor(const double* __restrict__ red, int nb, int n, float* __restrict__ fac) {
  double s = 0.0;
  for (int k = threadIdx.x; k < nb; k += 256) s += red[k];
  s = block_sum_f64(s);
  if (threadIdx.x == 0) {
    float msum = (float)s;
    msum /= (float)n;
    fac[0] = 1.0f / msum;
  }
}

__global__ __launch_bounds__(256) void k_norm_scale(float* __restrict__ model, int n, const float* __restrict__ fac) {
  const float f = fac[0];
  for (int k = blockIdx.x * 256 + threadIdx.x; k < n; k += gridDim.x * 256) model[k] *= f;
}

// red must hold >= kNormMaxBlocks + 2 doubles.
void launch_normalize_model(float* model, const float* weight, int n, double* red, hipStream_t st) {
  const int nb = (int)std::min<long>(kNormMaxBlocks, ((long)n + 255) / 256);
  float* fac = reinterpret_cast<float*>(red + kNormMaxBlocks);
  hipLaunchKernelGGL(k_norm_divide, dim3(nb), dim3(256), 0, st, model, weight, n, red);
  EMC_LAUNCHED();
  hipLaunchKernelGGL(k_norm_factor, dim3(1), dim3(256), 0, st, red, nb, n, fac);
  EMC_LAUNCHED();
  hipLaunchKernelGGL(k_norm_scale, dim3(nb), dim3(256), 0, st, model, n, fac);
  EMC_LAUNCHED();
}

// max(0, max_v model[v]) (emc_cuda.cu:364-368).  out[0] must be 0 on entry
// (as a bit pattern); non-negative floats order like their bit patterns.
__global__ __launch_bounds__(256) void k_model_max(const float* __restrict__ model, int n, unsigned int* __restrict__ out) {
  float m = 0.0f;
  for (int k = blockIdx.x * 256 + threadIdx.x; k < n; k += gridDim.x * 256) m = fmaxf(m, model[k]);
  m = block_max_f32(m);
  if (threadIdx.x == 0) atomicMax(out, __float_as_uint(m));
}

void launch_model_max(const float* model, int n, float* out, hipStream_t st) {
  const int nb = std::min(1024, (n + 255) / 256);
  hipLaunchKernelGGL(k_model_max, dim3(nb), dim3(256), 0, st, model, n, reinterpret_cast<unsigned int*>(out));
  EMC_LAUNCHED();
}

__global__ void k_fill(float* __restrict__ p, size_t n, float v) {
  for (size_t k = (size_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (size_t)gridDim.x * 256) p[k] = v;
}

void launch_fill(float* p, size_t n, float v, hipStream_t st) {
  if (!n) return;
  const unsigned nb = (unsigned)std::min<size_t>(4096, (n + 255) / 256);
  hipLaunchKernelGGL(k_fill, dim3(nb), dim3(256), 0, st, p, n, v);
  EMC_LAUNCHED();
}

}  // namespace emc
