// Patterns stored as CSR (compact pixel index int32, photon count uint16) and
// the sparse-count x log-slice contraction of the Poisson E-step:
//   C_ir = sum over the nonzero counts K_ik of pattern i of K_ik LS[r][k],
// with LS the FP16 hi/lo log-slice operand (per-rotation power-of-two scale),
// read transposed ([k][r], rotations contiguous) so that one nonzero feeds
// 256 rotations from one coalesced 1 KB row segment.
#include <hip/hip_runtime.h>

#include "engine.hpp"
#include "gemm_common.hpp"
#include "reduce.hpp"

namespace emc {

typedef _Float16 half2v __attribute__((ext_vector_type(2)));

// Nonzero counts of each pattern over the compact pixels.
__global__ __launch_bounds__(256) void k_csr_row_nnz(const float* __restrict__ images, const int* __restrict__ pix,
                                                     int N2d, int n_mp, int* __restrict__ nnz) {
  const int i = blockIdx.x;
  int c = 0;
  for (int k = threadIdx.x; k < n_mp; k += 256) {
    const int p = pix[k];
    if (p >= 0 && images[(size_t)i * N2d + p] != 0.0f) c++;
  }
  c = (int)block_sum_f32((float)c);
  if (threadIdx.x == 0) nnz[i] = c;
}

// One wave per pattern: its nonzeros in increasing compact pixel order.
__global__ __launch_bounds__(256) void k_csr_fill(const float* __restrict__ images, const int* __restrict__ pix,
                                                  int N, int N2d, int n_mp, const int* __restrict__ off,
                                                  int* __restrict__ col, unsigned short* __restrict__ cnt) {
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= N) return;
  int at = off[i];
  for (int k0 = 0; k0 < n_mp; k0 += 64) {
    const int k = k0 + lane;
    const int p = pix[k];
    const float v = p >= 0 ? images[(size_t)i * N2d + p] : 0.0f;
    const bool nz = v != 0.0f;
    const unsigned long long m = __ballot(nz);
    if (nz) {
      const int pos = at + __popcll(m & ((1ull << lane) - 1ull));
      col[pos] = k;
      cnt[pos] = (unsigned short)v;
    }
    at += __popcll(m);
  }
}

// LSt[k][r] = (hi, lo) of S16 / S16lo [R_pad][n_mp]; 64 x 64 tiles.
__global__ __launch_bounds__(256) void k_transpose_hilo(const _Float16* __restrict__ hi, const _Float16* __restrict__ lo,
                                                        int R_pad, int n_mp, half2v* __restrict__ out) {
  __shared__ half2v tile[64][65];
  const int r0 = blockIdx.y * 64, k0 = blockIdx.x * 64;
  for (int idx = threadIdx.x; idx < 64 * 64; idx += 256) {
    const int rr = idx >> 6, kk = idx & 63;
    const size_t src = (size_t)(r0 + rr) * n_mp + k0 + kk;
    tile[rr][kk] = half2v{hi[src], lo[src]};
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < 64 * 64; idx += 256) {
    const int kk = idx >> 6, rr = idx & 63;
    out[(size_t)(k0 + kk) * R_pad + r0 + rr] = tile[rr][kk];
  }
}

// One workgroup = 16 patterns x 256 rotations; thread = rotation.  The
// patterns' nonzeros are wave-uniform (scalar loads); each feeds one FMA per
// thread from a coalesced row of LSt.  Epilogue: the likelihood of
// like_store (Poisson form).
constexpr int kCsrPatterns = 16;

__global__ __launch_bounds__(256) void k_estep_csr(const int* __restrict__ off, const int* __restrict__ col,
                                                   const unsigned short* __restrict__ cnt,
                                                   const half2v* __restrict__ LSt, int R_pad, LikeEpilogue ep) {
  const int r = blockIdx.y * 256 + threadIdx.x;
  const int i0 = blockIdx.x * kCsrPatterns;
  float acc[kCsrPatterns];
#pragma unroll
  for (int b = 0; b < kCsrPatterns; b++) {
    acc[b] = 0.0f;
    const int i = i0 + b;
    if (i >= ep.N) continue;
    const int j1 = off[i + 1];
    float s = 0.0f;
    for (int j = off[i]; j < j1; j++) {
      const half2v v = LSt[(size_t)col[j] * R_pad + r];
      s += (float)cnt[j] * ((float)v.x + (float)v.y);
    }
    acc[b] = s;
  }
  if (r >= ep.R) return;
  const double cs = (double)ep.cscale[r];
  const double a = ep.A[r];
#pragma unroll
  for (int b = 0; b < kCsrPatterns; b++) {
    const int i = i0 + b;
    if (i >= ep.N) break;
    const double c = (double)acc[b] * cs;
    const double phi = (double)ep.phi[i];
    ep.L[(size_t)r * ep.N + i] = (float)(c + ep.B[i] * log(phi) - phi * a);
  }
}

// ---------------------------------------------------------------------------
// M-step over the nonzero responsibilities.  Concentrated posteriors leave
// almost every P_ri exactly zero, so U_r = sum_i P_ri w_i K_i (w = 1/phi, or
// 1 for Poisson) is a short sum per rotation: lists of each rotation's
// nonzero (pattern, P_ri w_i), patterns ascending, then one FP32 FMA per
// (nonzero, pixel) against the packed count rows.  The dropped terms are
// exact zeros, so S' is the dense M-step's sum in another order (and with P
// in FP32 instead of the FP16 hi/lo operand).
// ---------------------------------------------------------------------------

// One wave per rotation: the nonzeros of row r of P [R][N] in pattern order.
__global__ __launch_bounds__(256) void k_p_row_fill(const float* __restrict__ P, int R, int N,
                                                    const float* __restrict__ opscale, const int* __restrict__ off,
                                                    int* __restrict__ col, float* __restrict__ val) {
  const int lane = threadIdx.x & 63;
  const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= R) return;
  const float* row = P + (size_t)r * N;
  int at = off[r];
  // 16-byte loads (4 patterns per lane, 256 per wave step) when rows are
  // 16-byte aligned; positions from a wave prefix sum of the lanes' counts
  const int nv = ((N & 3) == 0 && ((uintptr_t)P & 15) == 0) ? N / 4 : 0;
  for (int v0 = 0; v0 < nv; v0 += 64) {
    const int v = v0 + lane;
    float4 q = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (v < nv) q = reinterpret_cast<const float4*>(row)[v];
    const int c = (q.x != 0.0f) + (q.y != 0.0f) + (q.z != 0.0f) + (q.w != 0.0f);
    if (__ballot(c != 0) == 0ull) continue;
    int incl = c;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int u = __shfl_up(incl, o, 64);
      if (lane >= o) incl += u;
    }
    int pos = at + incl - c;
    const float e[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
    for (int t = 0; t < 4; t++)
      if (e[t] != 0.0f) {
        const int i = 4 * v + t;
        col[pos] = i;
        val[pos] = e[t] * opscale[i];
        pos++;
      }
    at += __shfl(incl, 63, 64);
  }
  for (int i0 = 4 * nv; i0 < N; i0 += 64) {
    const int i = i0 + lane;
    const float p = i < N ? row[i] : 0.0f;
    const bool nz = p != 0.0f;
    const unsigned long long m = __ballot(nz);
    if (nz) {
      const int pos = at + __popcll(m & ((1ull << lane) - 1ull));
      col[pos] = i;
      val[pos] = p * opscale[i];
    }
    at += __popcll(m);
  }
}

// One workgroup = one rotation x 2,048 compact pixels (8 per thread, one
// 16-byte load of FP16 counts per nonzero); the epilogue of mstep_store.
// Rotations with more than kSpSeg nonzeros are split into segments of at
// most kSpSeg (partial sums, then a fixed-order finish), so no workgroup
// walks a long list alone.
constexpr int kSpPix = 8;
constexpr int kSpSeg = 128;
typedef _Float16 half8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ void sparse_accumulate(const int* __restrict__ col, const float* __restrict__ val, int j,
                                                  int j1, const _Float16* __restrict__ K,
                                                  const _Float16* __restrict__ Klo, int n_mp, int k0,
                                                  float (&acc)[kSpPix]) {
  // eight nonzeros per step: their count rows are loaded together (the loop
  // is latency-bound otherwise); accumulation stays in list order
  constexpr int U = 8;
  for (; j + U <= j1; j += U) {
    half8 h[U], l[U];
    float w[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const size_t base = (size_t)col[j + u] * n_mp + k0;
      w[u] = val[j + u];
      h[u] = *reinterpret_cast<const half8*>(K + base);
      if (Klo) l[u] = *reinterpret_cast<const half8*>(Klo + base);
    }
#pragma unroll
    for (int u = 0; u < U; u++)
#pragma unroll
      for (int t = 0; t < kSpPix; t++) acc[t] += w[u] * (Klo ? (float)h[u][t] + (float)l[u][t] : (float)h[u][t]);
  }
  for (; j < j1; j++) {
    const size_t base = (size_t)col[j] * n_mp + k0;
    const float w = val[j];
    const half8 h = *reinterpret_cast<const half8*>(K + base);
    if (Klo) {
      const half8 l = *reinterpret_cast<const half8*>(Klo + base);
#pragma unroll
      for (int t = 0; t < kSpPix; t++) acc[t] += w * ((float)h[t] + (float)l[t]);
    } else {
#pragma unroll
      for (int t = 0; t < kSpPix; t++) acc[t] += w * (float)h[t];
    }
  }
}

__device__ __forceinline__ void sparse_store(const MstepArgs& a, int r, int k0, const float (&acc)[kSpPix]) {
  const float T = a.Tf[r];
#pragma unroll
  for (int t = 0; t < kSpPix; t++) {
    const int k = k0 + t;
    if (k >= a.n_m) break;
    const float u = acc[t] * a.kscale_inv;
    const float s = a.poisson ? (T > 1e-30f ? u / T : 0.0f) : (T > 1e-10f ? u / T : u);
    a.slices[(size_t)r * a.N2d + a.pix[k]] = s;
    a.Sc[(size_t)r * a.n_mp + k] = s;
  }
}

// rotations with 1 .. kSpSeg nonzeros (rows[])
__global__ __launch_bounds__(256) void k_mstep_sparse(const int* __restrict__ rows, const int* __restrict__ off,
                                                      const int* __restrict__ col, const float* __restrict__ val,
                                                      const _Float16* __restrict__ K,
                                                      const _Float16* __restrict__ Klo, MstepArgs a) {
  const int r = rows[blockIdx.x];
  const int k0 = (blockIdx.y * 256 + threadIdx.x) * kSpPix;
  const int j0 = off[r], j1 = off[r + 1];
  if (k0 >= a.n_mp) return;
  float acc[kSpPix];
#pragma unroll
  for (int t = 0; t < kSpPix; t++) acc[t] = 0.0f;
  sparse_accumulate(col, val, j0, j1, K, Klo, a.n_mp, k0, acc);
  sparse_store(a, r, k0, acc);
}

// rotations without a nonzero responsibility: T_r = 0 and S' = 0 on the
// masked pixels of the slice (the compact copy is not written: insertion
// skips rotations with T_r <= 1e-10); lanes along pixels
__global__ __launch_bounds__(256) void k_mstep_zero(const int* __restrict__ rows, MstepArgs a) {
  float* out = a.slices + (size_t)rows[blockIdx.x] * a.N2d;
  for (int k = threadIdx.x; k < a.n_m; k += 256) __builtin_nontemporal_store(0.0f, out + a.pix[k]);
}

// the same with 16-byte stores: four pixels per lane, all four written at
// once when all are masked in (m4: bit q for pixel 4v + q), else one by
// one.  A workgroup takes one chunk of 1,024 groups of kZeroRows rows, which
// share the mask words, loaded once.  With qfast the chunk is blockIdx.x and
// the row group blockIdx.y, so that consecutive workgroups fill consecutive
// chunks of the same rows; else the other way round (row groups beyond the
// grid's y limit).
constexpr int kZeroRows = 4;

__global__ __launch_bounds__(256) void k_mstep_zero4(const int* __restrict__ rows, int nrows,
                                                     const int* __restrict__ m4, MstepArgs a, int qfast) {
  const int chunk = qfast ? blockIdx.x : blockIdx.y, grp = qfast ? blockIdx.y : blockIdx.x;
  typedef float f4 __attribute__((ext_vector_type(4)));
  const int nv = a.N2d / 4;
  const f4 z = {0.0f, 0.0f, 0.0f, 0.0f};
  {
    const int v0 = chunk * 256 * 4;  // this workgroup's chunk of the row
    int m[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int v = v0 + u * 256 + threadIdx.x;
      m[u] = v < nv ? m4[v] : 0;
    }
    for (int h = 0; h < kZeroRows; h++) {
      const int j = grp * kZeroRows + h;
      if (j >= nrows) break;
      float* out = a.slices + (size_t)rows[j] * a.N2d;
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int v = v0 + u * 256 + threadIdx.x;
        if (m[u] == 15) {
          reinterpret_cast<f4*>(out)[v] = z;
        } else if (m[u] != 0) {
#pragma unroll
          for (int q = 0; q < 4; q++)
            if (m[u] & (1 << q)) out[4 * v + q] = 0.0f;
        }
      }
    }
  }
}

__global__ __launch_bounds__(256) void k_mask4(const int* __restrict__ pix, int n_m, int* __restrict__ m4) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k < n_m) atomicOr(&m4[pix[k] >> 2], 1 << (pix[k] & 3));
}

// one segment (nonzeros seg.x .. seg.y - 1 of a long row) -> part[s][n_mp]
__global__ __launch_bounds__(256) void k_mstep_sparse_part(const int2* __restrict__ seg, const int* __restrict__ col,
                                                           const float* __restrict__ val,
                                                           const _Float16* __restrict__ K,
                                                           const _Float16* __restrict__ Klo, int n_mp,
                                                           float* __restrict__ part) {
  const int sg = blockIdx.x;
  const int k0 = (blockIdx.y * 256 + threadIdx.x) * kSpPix;
  if (k0 >= n_mp) return;
  float acc[kSpPix];
#pragma unroll
  for (int t = 0; t < kSpPix; t++) acc[t] = 0.0f;
  const int2 s = seg[sg];
  sparse_accumulate(col, val, s.x, s.y, K, Klo, n_mp, k0, acc);
  float* out = part + (size_t)sg * n_mp + k0;
#pragma unroll
  for (int t = 0; t < kSpPix; t++) out[t] = acc[t];
}

// long rows: their segments' partials summed in segment order, then stored
__global__ __launch_bounds__(256) void k_mstep_sparse_fin(const int* __restrict__ rows,
                                                          const int* __restrict__ row_seg, const float* __restrict__ part,
                                                          MstepArgs a) {
  const int h = blockIdx.x;
  const int k0 = (blockIdx.y * 256 + threadIdx.x) * kSpPix;
  if (k0 >= a.n_mp) return;
  float acc[kSpPix];
#pragma unroll
  for (int t = 0; t < kSpPix; t++) acc[t] = 0.0f;
  for (int sg = row_seg[h]; sg < row_seg[h + 1]; sg++) {
    const float* in = part + (size_t)sg * a.n_mp + k0;
#pragma unroll
    for (int t = 0; t < kSpPix; t++) acc[t] += in[t];
  }
  sparse_store(a, rows[h], k0, acc);
}

void launch_p_row_fill(const float* P, int R, int N, const float* opscale, const int* off, int* col, float* val,
                       hipStream_t st) {
  hipLaunchKernelGGL(k_p_row_fill, dim3((R + 3) / 4), dim3(256), 0, st, P, R, N, opscale, off, col, val);
  EMC_LAUNCHED();
}

bool launch_mstep_sparse(const std::vector<int>& h_off, const int* off, const int* col, const float* val,
                         const _Float16* K, const _Float16* Klo, const MstepArgs& a, hipStream_t st) {
  Engine& e = engine();
  const int chunks = (a.n_mp + 256 * kSpPix - 1) / (256 * kSpPix);
  // rows by nonzero count (host: the offsets are already here): none,
  // up to kSpSeg, more (split into segments)
  e.h_sp_seg.clear();
  e.h_sp_rows.clear();
  e.h_sp_rowseg.assign(1, 0);
  e.h_sp_light.clear();
  e.h_sp_zero.clear();
  for (int r = 0; r < a.R; r++) {
    const int j0 = h_off[r], j1 = h_off[r + 1];
    if (j1 == j0) {
      e.h_sp_zero.push_back(r);
      continue;
    }
    if (j1 - j0 <= kSpSeg) {
      e.h_sp_light.push_back(r);
      continue;
    }
    for (int j = j0; j < j1; j += kSpSeg) e.h_sp_seg.push_back(make_int2(j, std::min(j + kSpSeg, j1)));
    e.h_sp_rows.push_back(r);
    e.h_sp_rowseg.push_back((int)e.h_sp_seg.size());
  }
  const int nl = (int)e.h_sp_light.size(), nz = (int)e.h_sp_zero.size();
  if (!e.d_sp_light.ensure(sizeof(int) * std::max(nl, 1)) || !e.d_sp_zero.ensure(sizeof(int) * std::max(nz, 1)))
    return false;
  if (nl > 0) {
    EMC_CHECK(hipMemcpyAsync(e.d_sp_light.p, e.h_sp_light.data(), sizeof(int) * nl, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_mstep_sparse, dim3(nl, chunks), dim3(256), 0, st, e.d_sp_light.as<int>(), off, col, val, K,
                       Klo, a);
    EMC_LAUNCHED();
  }
  if (nz > 0) {
    EMC_CHECK(hipMemcpyAsync(e.d_sp_zero.p, e.h_sp_zero.data(), sizeof(int) * nz, hipMemcpyHostToDevice, st));
    const bool vec = a.N2d % 4 == 0 && (uintptr_t)a.slices % 16 == 0;
    if (vec && e.mask4_gen != e.pix_gen) {
      if (!e.d_mask4.ensure(sizeof(int) * (a.N2d / 4))) return false;
      launch_fill(e.d_mask4.as<float>(), (size_t)(a.N2d / 4), 0.0f, st);
      hipLaunchKernelGGL(k_mask4, dim3((a.n_m + 255) / 256), dim3(256), 0, st, a.pix, a.n_m, e.d_mask4.as<int>());
      EMC_LAUNCHED();
      e.mask4_gen = e.pix_gen;
    }
    const int ngrp = (nz + kZeroRows - 1) / kZeroRows, nchunk = (a.N2d / 4 + 1023) / 1024;
    if (vec && ngrp <= 65535)
      hipLaunchKernelGGL(k_mstep_zero4, dim3(nchunk, ngrp), dim3(256), 0, st, e.d_sp_zero.as<int>(), nz,
                         e.d_mask4.as<int>(), a, 1);
    else if (vec)
      hipLaunchKernelGGL(k_mstep_zero4, dim3(ngrp, nchunk), dim3(256), 0, st, e.d_sp_zero.as<int>(), nz,
                         e.d_mask4.as<int>(), a, 0);
    else
      hipLaunchKernelGGL(k_mstep_zero, dim3(nz), dim3(256), 0, st, e.d_sp_zero.as<int>(), a);
    EMC_LAUNCHED();
  }
  const int ns = (int)e.h_sp_seg.size(), nh = (int)e.h_sp_rows.size();
  if (nh == 0) return !failed();
  if (!e.d_sp_seg.ensure(sizeof(int2) * ns) || !e.d_sp_rows.ensure(sizeof(int) * nh) ||
      !e.d_sp_rowseg.ensure(sizeof(int) * (nh + 1)) || !e.d_sp_part.ensure(sizeof(float) * (size_t)ns * a.n_mp))
    return false;
  EMC_CHECK(hipMemcpyAsync(e.d_sp_seg.p, e.h_sp_seg.data(), sizeof(int2) * ns, hipMemcpyHostToDevice, st));
  EMC_CHECK(hipMemcpyAsync(e.d_sp_rows.p, e.h_sp_rows.data(), sizeof(int) * nh, hipMemcpyHostToDevice, st));
  EMC_CHECK(hipMemcpyAsync(e.d_sp_rowseg.p, e.h_sp_rowseg.data(), sizeof(int) * (nh + 1), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_mstep_sparse_part, dim3(ns, chunks), dim3(256), 0, st, e.d_sp_seg.as<int2>(), col, val, K,
                     Klo, a.n_mp, e.d_sp_part.as<float>());
  EMC_LAUNCHED();
  hipLaunchKernelGGL(k_mstep_sparse_fin, dim3(nh, chunks), dim3(256), 0, st, e.d_sp_rows.as<int>(),
                     e.d_sp_rowseg.as<int>(), e.d_sp_part.as<float>(), a);
  EMC_LAUNCHED();
  return !failed();
}

// Dense images from CSR: images[i][pix] = count (zero elsewhere, filled first).
__global__ __launch_bounds__(256) void k_images_from_csr(const int* __restrict__ off, const int* __restrict__ pix,
                                                         const unsigned short* __restrict__ cnt, int N, int N2d,
                                                         float* __restrict__ images) {
  const int i = blockIdx.x;
  for (int j = off[i] + threadIdx.x; j < off[i + 1]; j += 256) {
    const int p = pix[j];
    if (p >= 0 && p < N2d) images[(size_t)i * N2d + p] = (float)cnt[j];
  }
}

void launch_csr_row_nnz(const float* images, const int* pix, int N, int N2d, int n_mp, int* nnz, hipStream_t st) {
  hipLaunchKernelGGL(k_csr_row_nnz, dim3(N), dim3(256), 0, st, images, pix, N2d, n_mp, nnz);
  EMC_LAUNCHED();
}

void launch_csr_fill(const float* images, const int* pix, int N, int N2d, int n_mp, const int* off, int* col,
                     unsigned short* cnt, hipStream_t st) {
  hipLaunchKernelGGL(k_csr_fill, dim3((N + 3) / 4), dim3(256), 0, st, images, pix, N, N2d, n_mp, off, col, cnt);
  EMC_LAUNCHED();
}

void launch_transpose_hilo(const _Float16* hi, const _Float16* lo, int R_pad, int n_mp, void* out, hipStream_t st) {
  hipLaunchKernelGGL(k_transpose_hilo, dim3(n_mp / 64, R_pad / 64), dim3(256), 0, st, hi, lo, R_pad, n_mp,
                     (half2v*)out);
  EMC_LAUNCHED();
}

void launch_estep_csr(const int* off, const int* col, const unsigned short* cnt, const void* LSt, int R_pad,
                      const LikeEpilogue& ep, hipStream_t st) {
  const dim3 grid((ep.N + kCsrPatterns - 1) / kCsrPatterns, R_pad / 256);
  hipLaunchKernelGGL(k_estep_csr, grid, dim3(256), 0, st, off, col, cnt, (const half2v*)LSt, R_pad, ep);
  EMC_LAUNCHED();
}

void launch_images_from_csr(const int* off, const int* pix, const unsigned short* cnt, int N, int N2d, float* images,
                            hipStream_t st) {
  launch_fill(images, (size_t)N * N2d, 0.0f, st);
  hipLaunchKernelGGL(k_images_from_csr, dim3(N), dim3(256), 0, st, off, pix, cnt, N, N2d, images);
  EMC_LAUNCHED();
}

}  // namespace emc
