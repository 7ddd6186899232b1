// Kernels of the rotation-proposal network (no counterpart in the reference).
//
// Network: x = log1p(counts) on the S x S detector (unmasked pixels 0);
// four 3x3 / stride-2 / pad-1 convolutions with ReLU (channels 16, 32, 64,
// 64), flatten, dense 256 + ReLU, dense C (coarse rotation cells), softmax.
// Activations are NHWC FP32; convolutions run as GEMMs that gather their
// im2col operand on the fly (the first layer, one channel, from an explicit
// im2col buffer), input gradients as the transposed convolution's GEMM with
// the ReLU' gate in its epilogue; every GEMM rounds its operands to BF16 and
// accumulates in FP32 on v_mfma_f32_16x16x32_bf16 (BF16 compute, FP32
// master weights).
#include <hip/hip_runtime.h>

#include <climits>

#include <algorithm>

#include "engine.hpp"
#include "cnn.hpp"
#include "reduce.hpp"

namespace emc {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// ---------------------------------------------------------------------------
// Generic strided GEMM: C(m, n) = sum_k A(m, k) B(n, k), operands FP32 in
// memory (any strides), rounded to BF16 on the way to LDS.  64 x 64 tiles,
// 4 waves of 32 x 32, BK = 32, two LDS stages: the next k-step's operand
// values are loaded into registers while the current stage's products run.
// Each thread moves 8 values of each operand per k-step: 8 consecutive k of
// one row (load mode kKFast: 16-byte loads when the row's k run is
// contiguous and aligned), 8 consecutive rows at one k (kRowFast: the same
// along rows), or 8 strided scalars (kStrided).  Split-K over gridDim.z
// writes partial tiles to `partial` ([z][M][N]); gemm_finish adds them in a
// fixed order.
// ---------------------------------------------------------------------------
constexpr int GT = 64, GK = 32, GLD = GK + 8;
enum { kStrided = 0, kKFast = 1, kRowFast = 2, kConvMK = 3, kConvKM = 4, kConvT = 5 };

struct OpView {
  const float* p;
  long sr, sk;  // element (row, k) at p[row * sr + k * sk]
  int rows;
  int H, W, C;   // convolution input (kConvMK / kConvKM)
  int tcls = -1;  // kConvT: parity class of the rows (-1: all pixels, all 9 taps)
};

// Tap (ky, kx) of local tap j of parity class (py, px): ky = 1 when py = 0,
// else 0 or 2 (likewise kx), ky-major.
__device__ __forceinline__ void class_tap(int py, int px, int j, int& ky, int& kx) {
  const int nkx = px ? 2 : 1;
  const int jy = j / nkx, jx = j - jy * nkx;
  ky = py ? 2 * jy : 1;
  kx = px ? 2 * jx : 1;
}

// Element (m, kk) of the implicit im2col matrix of a 3x3 / stride-2 / pad-1
// convolution over NHWC [b][H][W][C]: its address, or nullptr for a tap in
// the padding.
__device__ __forceinline__ const float* conv_at(const OpView& v, int m, int kk) {
  const int OH = v.H >> 1, OW = v.W >> 1;
  const int ox = m % OW, t = m / OW, oy = t % OH, b = t / OH;
  const int c = kk % v.C, t2 = kk / v.C, kx = t2 % 3, ky = t2 / 3;
  const int y = 2 * oy + ky - 1, x = 2 * ox + kx - 1;
  if (y < 0 || y >= v.H || x < 0 || x >= v.W) return nullptr;
  return v.p + (((size_t)b * v.H + y) * v.W + x) * v.C + c;
}

// Element (m, k) of the transposed convolution's operand: m a pixel
// (b, y, x) of the convolution's H x W input, k = tap * C + oc; the output
// gradient [b][H/2][W/2][C] at the output pixel whose tap (ky, kx) covers
// (y, x), or nullptr where none does.
__device__ __forceinline__ const float* convt_at(const OpView& v, int m, int k) {
  const int OH = v.H >> 1, OW = v.W >> 1;
  if (v.tcls >= 0) {  // pixel (2yy + py, 2xx + px), the class's taps only
    const int py = v.tcls >> 1, px = v.tcls & 1;
    const int xx = m % OW, t = m / OW, yy = t % OH, b = t / OH;
    const int oc = k % v.C;
    int ky, kx;
    class_tap(py, px, k / v.C, ky, kx);
    const int oy = yy + ((py + 1 - ky) >> 1), ox = xx + ((px + 1 - kx) >> 1);
    if (oy >= OH || ox >= OW) return nullptr;
    return v.p + (((size_t)b * OH + oy) * OW + ox) * v.C + oc;
  }
  const int x = m % v.W, t = m / v.W, y = t % v.H, b = t / v.H;
  const int oc = k % v.C, tap = k / v.C, kx = tap % 3, ky = tap / 3;
  const int ty = y + 1 - ky, tx = x + 1 - kx;
  if (ty < 0 || tx < 0 || (ty & 1) || (tx & 1)) return nullptr;
  const int oy = ty >> 1, ox = tx >> 1;
  if (oy >= OH || ox >= OW) return nullptr;
  return v.p + (((size_t)b * OH + oy) * OW + ox) * v.C + oc;
}

// 8 consecutive columns c .. c + 7 of one (m, tap): contiguous, 16-byte aligned
__device__ __forceinline__ void conv8(const float* q, float (&x)[8]) {
  if (q) {
    const float4 a = reinterpret_cast<const float4*>(q)[0], b = reinterpret_cast<const float4*>(q)[1];
    x[0] = a.x; x[1] = a.y; x[2] = a.z; x[3] = a.w; x[4] = b.x; x[5] = b.y; x[6] = b.z; x[7] = b.w;
  } else {
#pragma unroll
    for (int e = 0; e < 8; e++) x[e] = 0.0f;
  }
}

// the 8 values thread t moves for rows r0.. of the tile and k0.. of the
// k-step (bounds: rows < v.rows, k < k_end), zero outside
template <int MODE>
__device__ __forceinline__ void load8(const OpView& v, int row0, int k0, int k_end, int t, float (&x)[8]) {
  if (MODE == kConvT) {  // rows = input pixels, k = taps x output channels: as kKFast
    const int r = row0 + (t >> 2), k = k0 + (t & 3) * 8;
    if (r < v.rows && k + 8 <= k_end) {
      conv8(convt_at(v, r, k), x);
    } else {
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const float* q = (r < v.rows && k + e < k_end) ? convt_at(v, r, k + e) : nullptr;
        x[e] = q ? *q : 0.0f;
      }
    }
  } else if (MODE == kConvMK) {  // rows m, k = taps x channels: as kKFast
    const int r = row0 + (t >> 2), k = k0 + (t & 3) * 8;
    if (r < v.rows && k + 8 <= k_end) {
      conv8(conv_at(v, r, k), x);
    } else {
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const float* q = (r < v.rows && k + e < k_end) ? conv_at(v, r, k + e) : nullptr;
        x[e] = q ? *q : 0.0f;
      }
    }
  } else if (MODE == kConvKM) {  // rows = taps x channels, k = m: as kRowFast
    const int k = k0 + (t >> 3), r = row0 + (t & 7) * 8;
    if (k < k_end && r + 8 <= v.rows) {
      conv8(conv_at(v, k, r), x);
    } else {
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const float* q = (k < k_end && r + e < v.rows) ? conv_at(v, k, r + e) : nullptr;
        x[e] = q ? *q : 0.0f;
      }
    }
  } else if (MODE == kRowFast) {
    const int k = k0 + (t >> 3), r = row0 + (t & 7) * 8;
    if (k < k_end && r + 8 <= v.rows) {
      const float4* q = reinterpret_cast<const float4*>(v.p + (long)r + (long)k * v.sk);
      const float4 a = q[0], b = q[1];
      x[0] = a.x; x[1] = a.y; x[2] = a.z; x[3] = a.w; x[4] = b.x; x[5] = b.y; x[6] = b.z; x[7] = b.w;
    } else {
#pragma unroll
      for (int e = 0; e < 8; e++) x[e] = (k < k_end && r + e < v.rows) ? v.p[(long)(r + e) + (long)k * v.sk] : 0.0f;
    }
  } else if (MODE == kKFast) {
    const int r = row0 + (t >> 2), k = k0 + (t & 3) * 8;
    if (r < v.rows && k + 8 <= k_end) {
      const float4* q = reinterpret_cast<const float4*>(v.p + (long)r * v.sr + (long)k);
      const float4 a = q[0], b = q[1];
      x[0] = a.x; x[1] = a.y; x[2] = a.z; x[3] = a.w; x[4] = b.x; x[5] = b.y; x[6] = b.z; x[7] = b.w;
    } else {
#pragma unroll
      for (int e = 0; e < 8; e++) x[e] = (r < v.rows && k + e < k_end) ? v.p[(long)r * v.sr + (long)(k + e)] : 0.0f;
    }
  } else {
    const bool kf = v.sk == 1;
#pragma unroll
    for (int e = 0; e < 8; e++) {
      int r, k;
      if (kf) { r = t >> 2; k = (t & 3) * 8 + e; }
      else { k = t >> 3; r = (t & 7) * 8 + e; }
      r += row0;
      k += k0;
      x[e] = (r < v.rows && k < k_end) ? v.p[(long)r * v.sr + (long)k * v.sk] : 0.0f;
    }
  }
}

// the same 8 values into the LDS tile [row][k] as BF16
template <int MODE>
__device__ __forceinline__ void store8(__bf16 (*s)[GLD], int t, bool kf_strided, const float (&x)[8]) {
  if (MODE == kKFast || MODE == kConvMK || MODE == kConvT || (MODE == kStrided && kf_strided)) {
    bf16x8 v;
#pragma unroll
    for (int e = 0; e < 8; e++) v[e] = (__bf16)x[e];
    *reinterpret_cast<bf16x8*>(&s[t >> 2][(t & 3) * 8]) = v;
  } else {
#pragma unroll
    for (int e = 0; e < 8; e++) s[(t & 7) * 8 + e][t >> 3] = (__bf16)x[e];
  }
}

// C's row of output row m (the pixel's row for a parity class)
__device__ __forceinline__ long c_row(const GemmDesc& d, int m) {
  if (!d.conv_t || d.tcls < 0) return m;
  const int OH = d.cH >> 1, OW = d.cW >> 1;
  const int xx = m % OW, t = m / OW, yy = t % OH, b = t / OH;
  return ((long)b * d.cH + 2 * yy + (d.tcls >> 1)) * d.cW + 2 * xx + (d.tcls & 1);
}

template <int AM, int BM>
__global__ __launch_bounds__(256) void k_gemm_bf16(GemmDesc d, float* __restrict__ partial, int ksplit_len) {
  __shared__ __attribute__((aligned(16))) __bf16 sA[2][GT][GLD];
  __shared__ __attribute__((aligned(16))) __bf16 sB[2][GT][GLD];
  const int m0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
  const int kz0 = blockIdx.z * ksplit_len, kz1 = min(d.K, kz0 + ksplit_len);
  const int t = threadIdx.x, lane = t & 63, w = t >> 6, wm = w >> 1, wn = w & 1;
  OpView va{d.A, d.sam, d.sak, d.M, d.cH, d.cW, d.cC}, vb{d.B, d.sbn, d.sbk, d.N, d.cH, d.cW, d.cC};
  va.tcls = d.conv_t ? d.tcls : -1;
  const bool a_kf = d.sak == 1, b_kf = d.sbk == 1;
  floatx4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) acc[a][b] = floatx4{0, 0, 0, 0};
  float xa[8], xb[8];
  int cur = 0;
  if (kz0 < kz1) {
    load8<AM>(va, m0, kz0, kz1, t, xa);
    load8<BM>(vb, n0, kz0, kz1, t, xb);
    store8<AM>(sA[0], t, a_kf, xa);
    store8<BM>(sB[0], t, b_kf, xb);
  }
  __syncthreads();
  const int fr = lane & 15, fk = (lane >> 4) * 8;
  for (int k0 = kz0; k0 < kz1; k0 += GK) {
    const bool more = k0 + GK < kz1;
    if (more) {
      load8<AM>(va, m0, k0 + GK, kz1, t, xa);
      load8<BM>(vb, n0, k0 + GK, kz1, t, xb);
    }
    bf16x8 fa[2], fb[2];
#pragma unroll
    for (int a = 0; a < 2; a++) fa[a] = *reinterpret_cast<const bf16x8*>(&sA[cur][wm * 32 + a * 16 + fr][fk]);
#pragma unroll
    for (int b = 0; b < 2; b++) fb[b] = *reinterpret_cast<const bf16x8*>(&sB[cur][wn * 32 + b * 16 + fr][fk]);
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 2; b++) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[a], fb[b], acc[a][b], 0, 0, 0);
    if (more) {
      store8<AM>(sA[cur ^ 1], t, a_kf, xa);
      store8<BM>(sB[cur ^ 1], t, b_kf, xb);
    }
    __syncthreads();
    cur ^= 1;
  }
  const int g = lane >> 4;
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int m = m0 + wm * 32 + a * 16 + 4 * g + j, n = n0 + wn * 32 + b * 16 + fr;
        if (m >= d.M || n >= d.N) continue;
        if (partial) {
          partial[((long)blockIdx.z * d.M + m) * d.N + n] = acc[a][b][j];
        } else {
          float v = acc[a][b][j];
          if (d.bias_n) v += d.bias_n[n];
          if (d.bias_m) v += d.bias_m[m];
          if (d.relu) v = fmaxf(v, 0.0f);
          const long cm = c_row(d, m);
          if (d.gate && !(d.gate[cm * d.scm + (long)n * d.scn] > 0.0f)) v = 0.0f;
          float* c = d.C + cm * d.scm + (long)n * d.scn;
          *c = d.accumulate ? *c + v : v;
        }
      }
}

__global__ void k_gemm_finish(GemmDesc d, const float* __restrict__ partial, int nz) {
  const long idx = (long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (long)d.M * d.N) return;
  const int m = (int)(idx / d.N), n = (int)(idx % d.N);
  float v = 0.0f;
  for (int z = 0; z < nz; z++) v += partial[(long)z * d.M * d.N + idx];
  if (d.bias_n) v += d.bias_n[n];
  if (d.bias_m) v += d.bias_m[m];
  if (d.relu) v = fmaxf(v, 0.0f);
  const long cm = c_row(d, m);
  if (d.gate && !(d.gate[cm * d.scm + (long)n * d.scn] > 0.0f)) v = 0.0f;
  float* c = d.C + cm * d.scm + (long)n * d.scn;
  *c = d.accumulate ? *c + v : v;
}

// Load mode of an operand view: 16-byte loads along k (k contiguous, rows
// 16-byte aligned) or along rows (rows contiguous, every k 16-byte aligned),
// else strided scalars.
static int load_mode(const float* p, long sr, long sk) {
  const bool al = ((uintptr_t)p & 15) == 0;
  if (al && sk == 1 && sr % 4 == 0) return kKFast;
  if (al && sr == 1 && sk % 4 == 0) return kRowFast;
  return kStrided;
}

// Highest element index a strided 2-D view (rows x cols, strides sr, sc)
// touches, or -1 for an empty view.
static long last_index(long rows, long cols, long sr, long sc) {
  return rows <= 0 || cols <= 0 ? -1 : (rows - 1) * sr + (cols - 1) * sc;
}

// Highest element a convolution view reads (the whole NHWC input of
// `positions` output pixels' batch), or -2 when the view's shape is invalid.
static long conv_last(const GemmDesc& d, long positions, long taps) {
  if (d.cH < 2 || d.cW < 2 || d.cH % 2 || d.cW % 2 || d.cC < 8 || d.cC % 8 || taps != 9L * d.cC ||
      ((uintptr_t)(d.conv_a ? d.A : d.B) & 15))
    return -2;
  const long per = (long)(d.cH / 2) * (d.cW / 2);
  if (positions % per) return -2;
  return positions / per * d.cH * d.cW * d.cC - 1;
}

// Highest element the transposed-convolution view reads (the output
// gradient of the batch whose input pixels are the M rows), or -2 when the
// shape is invalid.
static long convt_last(const GemmDesc& d) {
  const int taps = d.tcls < 0 ? 9 : d.tcls < 4 ? kTapClassN[d.tcls] : 0;
  if (d.cH < 2 || d.cW < 2 || d.cH % 2 || d.cW % 2 || d.cC < 8 || d.cC % 8 || taps == 0 || d.K != taps * d.cC ||
      ((uintptr_t)d.A & 15))
    return -2;
  const long per = d.tcls < 0 ? (long)d.cH * d.cW : (long)(d.cH / 2) * (d.cW / 2);
  if (d.M % per) return -2;
  return d.M / per * (d.cH / 2) * (d.cW / 2) * d.cC - 1;
}

// Highest element of C written: the last row's, or for a parity class the
// batch's last pixel row.
static long c_last(const GemmDesc& d) {
  if (d.conv_t && d.tcls >= 0) {
    const long per = (long)(d.cH / 2) * (d.cW / 2);
    if (per <= 0 || d.M % per) return LONG_MAX;
    return last_index(d.M / per * d.cH * d.cW, d.N, d.scm, d.scn);
  }
  return last_index(d.M, d.N, d.scm, d.scn);
}

bool gemm_bf16(const GemmDesc& d, float* scratch, size_t scratch_floats, hipStream_t st) {
  if (d.M <= 0 || d.N <= 0) return true;
  const long la = d.conv_t ? convt_last(d) : d.conv_a ? conv_last(d, d.M, d.K) : last_index(d.M, d.K, d.sam, d.sak);
  const long lb = d.conv_b ? conv_last(d, d.K, d.N) : last_index(d.N, d.K, d.sbn, d.sbk);
  if (d.K < 0 || d.sam < 0 || d.sak < 0 || d.sbn < 0 || d.sbk < 0 || d.scm < 0 || d.scn < 0 || la < -1 ||
      lb < -1 || (d.conv_a + d.conv_b + d.conv_t > 1) || la >= d.a_len || lb >= d.b_len ||
      c_last(d) >= d.c_len || !d.A || !d.B || !d.C) {
    set_error("gemm_bf16: descriptor out of bounds (M %d N %d K %d; A %ld/%ld, B %ld/%ld, C %ld/%ld)", d.M, d.N, d.K,
              la, d.a_len, lb, d.b_len, c_last(d), d.c_len);
    return false;
  }
  const int gm = (d.M + GT - 1) / GT, gn = (d.N + GT - 1) / GT;
  // split K so that a long reduction still spreads over the chip
  int nz = 1;
  const long tiles = (long)gm * gn;
  while (tiles * nz < 512 && d.K / (nz * 2) >= 4 * GK && (size_t)(nz * 2) * d.M * d.N <= scratch_floats) nz *= 2;
  const int klen = ((d.K + nz - 1) / nz + GK - 1) / GK * GK;
  nz = (d.K + klen - 1) / klen;
  if (nz > 1 && (!scratch || (size_t)nz * d.M * d.N > scratch_floats)) {
    set_error("gemm_bf16: split-k partials (%d x %d x %d) exceed the scratch (%zu floats)", nz, d.M, d.N,
              scratch_floats);
    return false;
  }
  const int am = d.conv_t ? kConvT : d.conv_a ? kConvMK : load_mode(d.A, d.sam, d.sak);
  const int bm = d.conv_b ? kConvKM : load_mode(d.B, d.sbn, d.sbk);
  const dim3 grid(gn, gm, nz > 1 ? nz : 1);
  float* part = nz > 1 ? scratch : nullptr;
  const int kl = nz > 1 ? klen : d.K;
#define EMC_GEMM_CASE(A_, B_)                                                                  \
  if (am == A_ && bm == B_) hipLaunchKernelGGL((k_gemm_bf16<A_, B_>), grid, dim3(256), 0, st, d, part, kl);
  EMC_GEMM_CASE(kStrided, kStrided) EMC_GEMM_CASE(kStrided, kKFast) EMC_GEMM_CASE(kStrided, kRowFast)
  EMC_GEMM_CASE(kKFast, kStrided) EMC_GEMM_CASE(kKFast, kKFast) EMC_GEMM_CASE(kKFast, kRowFast)
  EMC_GEMM_CASE(kRowFast, kStrided) EMC_GEMM_CASE(kRowFast, kKFast) EMC_GEMM_CASE(kRowFast, kRowFast)
  EMC_GEMM_CASE(kConvMK, kStrided) EMC_GEMM_CASE(kConvMK, kKFast) EMC_GEMM_CASE(kConvMK, kRowFast)
  EMC_GEMM_CASE(kStrided, kConvKM) EMC_GEMM_CASE(kKFast, kConvKM) EMC_GEMM_CASE(kRowFast, kConvKM)
  EMC_GEMM_CASE(kConvT, kStrided) EMC_GEMM_CASE(kConvT, kKFast) EMC_GEMM_CASE(kConvT, kRowFast)
#undef EMC_GEMM_CASE
  if (nz > 1) {
    EMC_LAUNCHED();
    const long n = (long)d.M * d.N;
    hipLaunchKernelGGL(k_gemm_finish, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, d, scratch, nz);
  }
  EMC_LAUNCHED();
  return true;
}

// ---------------------------------------------------------------------------
// Layers
// ---------------------------------------------------------------------------

// x[b][y][x] = log1p(max(K, 0)) on masked-in pixels, 0 elsewhere.
__global__ void k_cnn_input(const float* __restrict__ images, const int* __restrict__ mask, int N2d,
                            const int* __restrict__ idx, int B, float* __restrict__ x) {
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t >= (long)B * N2d) return;
  const int b = (int)(t / N2d), p = (int)(t % N2d);
  const float k = images[(long)idx[b] * N2d + p];
  x[t] = mask[p] != 0 ? log1pf(fmaxf(k, 0.0f)) : 0.0f;
}

// col[(b*OH + oy)*OW + ox][(ky*3 + kx)*C + c] = in[b][2oy+ky-1][2ox+kx-1][c]
__global__ void k_im2col(const float* __restrict__ in, int B, int H, int W, int C, float* __restrict__ col) {
  const int OH = H / 2, OW = W / 2, K = 9 * C;
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t >= (long)B * OH * OW * K) return;
  const long m = t / K;
  const int kk = (int)(t % K);
  const int c = kk % C, kx = (kk / C) % 3, ky = kk / (3 * C);
  const int ox = (int)(m % OW), oy = (int)((m / OW) % OH), b = (int)(m / ((long)OW * OH));
  const int y = 2 * oy + ky - 1, x = 2 * ox + kx - 1;
  col[t] = (y >= 0 && y < H && x >= 0 && x < W) ? in[(((long)b * H + y) * W + x) * C + c] : 0.0f;
}

// First convolution (one input channel, kC0 output channels, 3 x 3, stride
// 2, padding 1) computed directly: its 9-long sums are too short for a
// matrix-core tile (a 64 x 64 tile with k padded to 32 wasted 3/4 of the
// columns and 2/3 of k).  Forward: a thread per output pixel gathers its 9
// inputs and writes its kC0 channels (bias, ReLU) as one 64-byte row.  The
// operands are rounded to BF16 as the GEMM rounds them (FP32 products and
// sums), so every layer computes the same function.
constexpr int kC0 = 16;

__device__ __forceinline__ float bf16_round(float x) { return (float)(__bf16)x; }

__global__ __launch_bounds__(256) void k_conv0_fwd(const float* __restrict__ in, int B, int H,
                                                   const float* __restrict__ w, const float* __restrict__ bias,
                                                   float* __restrict__ out) {
  const int OH = H / 2;
  const int m = blockIdx.x * 256 + threadIdx.x;  // B * OH * OH < 2^31 (host check)
  if (m >= B * OH * OH) return;
  const int ox = m % OH, oy = (m / OH) % OH;
  const long b = m / (OH * OH);
  // the 9 taps loaded without branches (a tap in the padding reads the
  // pixel's own value and is then zeroed), so all are in flight together
  float v[9];
#pragma unroll
  for (int t = 0; t < 9; t++) {
    const int y = 2 * oy + t / 3 - 1, x = 2 * ox + t % 3 - 1;
    const bool ok = y >= 0 && y < H && x >= 0 && x < H;
    const float u = in[(b * H + (ok ? y : 2 * oy)) * H + (ok ? x : 2 * ox)];
    v[t] = ok ? bf16_round(u) : 0.0f;
  }
  float o[kC0];
#pragma unroll
  for (int c = 0; c < kC0; c++) {
    float s = 0.0f;
#pragma unroll
    for (int t = 0; t < 9; t++) s += v[t] * bf16_round(w[c * 9 + t]);
    o[c] = fmaxf(s + bias[c], 0.0f);
  }
  float4* dst = reinterpret_cast<float4*>(out + m * kC0);
#pragma unroll
  for (int q = 0; q < kC0 / 4; q++) dst[q] = make_float4(o[4 * q], o[4 * q + 1], o[4 * q + 2], o[4 * q + 3]);
}

// Its weight and bias gradients: dW[c][t] = sum_m dz[m][c] v_t(m) and
// dB[c] = sum_m dz[m][c] over the output pixels m.  A workgroup takes `per`
// consecutive pixels; thread (pixel lane t / 16, channel t % 16) takes
// every 16th of them (a wave reads 4 pixels x 16 channels, 256 contiguous
// bytes of dz), the 16 lanes of a channel are added in order into the
// workgroup's partial sums, and k_conv0_wgrad_fin adds the workgroups' in
// order.
__global__ __launch_bounds__(256) void k_conv0_wgrad(const float* __restrict__ dz, const float* __restrict__ in,
                                                     int B, int H, long per, double* __restrict__ part) {
  static_assert(kC0 == 16, "one channel per lane of 16");
  __shared__ float red[10][256];
  const int OH = H / 2;
  const int M = B * OH * OH;  // < 2^31 (host check)
  const int c = threadIdx.x & 15, pl = threadIdx.x >> 4;
  const int m0 = blockIdx.x * (int)per, m1 = min(M, m0 + (int)per);
  float acc[10];
#pragma unroll
  for (int k = 0; k < 10; k++) acc[k] = 0.0f;
  // pixel (b, oy, ox) of m, advanced by 16 pixels per step without
  // divisions; 4 pixels (m, m + 16, m + 32, m + 48) per iteration, all their
  // loads issued before any is used
  int ox = (m0 + pl) % OH, oy = ((m0 + pl) / OH) % OH, b = (m0 + pl) / (OH * OH);
  auto advance = [&](int& x_, int& y_, int& b_) {
    x_ += 16;
    while (x_ >= OH) {
      x_ -= OH;
      if (++y_ == OH) {
        y_ = 0;
        b_++;
      }
    }
  };
  for (int m = m0 + pl; m < m1; m += 64) {
    float gf[4], u[4][9];
    bool ok[4][9];
    int px = ox, py = oy, pb = b;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const bool live = m + 16 * q < m1;
      gf[q] = live ? dz[(long)(m + 16 * q) * kC0 + c] : 0.0f;
#pragma unroll
      for (int t = 0; t < 9; t++) {  // branch-free loads (see k_conv0_fwd)
        const int y = 2 * py + t / 3 - 1, x = 2 * px + t % 3 - 1;
        ok[q][t] = live && y >= 0 && y < H && x >= 0 && x < H;
        const long pix = live ? ((long)pb * H + (ok[q][t] ? y : 2 * py)) * H + (ok[q][t] ? x : 2 * px) : 0;
        u[q][t] = in[pix];
      }
      advance(px, py, pb);
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const float g = bf16_round(gf[q]);
#pragma unroll
      for (int t = 0; t < 9; t++) acc[t] += g * (ok[q][t] ? bf16_round(u[q][t]) : 0.0f);
      acc[9] += gf[q];  // the bias gradient sums dz itself (FP32, as the column sums do)
    }
    ox = px;
    oy = py;
    b = pb;
  }
#pragma unroll
  for (int k = 0; k < 10; k++) red[k][threadIdx.x] = acc[k];
  __syncthreads();
  if (threadIdx.x < 10 * kC0) {
    const int k = threadIdx.x / kC0, cc = threadIdx.x % kC0;
    double s = 0.0;
    for (int p = 0; p < 16; p++) s += (double)red[k][p * 16 + cc];
    part[(long)blockIdx.x * 10 * kC0 + cc * 10 + k] = s;
  }
}

__global__ __launch_bounds__(256) void k_conv0_wgrad_fin(const double* __restrict__ part, int nblk,
                                                         float* __restrict__ dW, float* __restrict__ dB) {
  const int j = blockIdx.x;  // channel j / 10; taps 0..8, then the bias
  double s = 0.0;
  for (int sp = threadIdx.x; sp < nblk; sp += 256) s += part[(long)sp * 10 * kC0 + j];
  s = block_sum_f64(s);
  if (threadIdx.x == 0) {
    const int c = j / 10, k = j % 10;
    if (k < 9)
      dW[c * 9 + k] = (float)s;
    else
      dB[c] = (float)s;
  }
}

// wt[(t*OC + o)*IC + c] = w[(o*9 + t)*IC + c]
__global__ void k_conv_weights_t(const float* __restrict__ w, int OC, int IC, float* __restrict__ wt) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= 9 * OC * IC) return;
  const int c = i % IC, o = (i / IC) % OC, t = i / (IC * OC);
  wt[i] = w[((long)o * 9 + t) * IC + c];
}

// out[j] = sum_i G[i*si + j*sj], i < rows (fixed order per column).
// Column sums in two passes: partial sums of row ranges (grid cols x
// splits), then the splits added in order -- fixed order, and enough
// workgroups for the million-row activations of the first layers.
__global__ void k_reduce_rows_part(const float* __restrict__ G, long rows, int cols, long si, long sj, long per,
                                   double* __restrict__ part) {
  const int j = blockIdx.x, sp = blockIdx.y;
  const long r0 = (long)sp * per, r1 = min(rows, r0 + per);
  double s = 0.0;
  for (long i = r0 + threadIdx.x; i < r1; i += 256) s += G[i * si + (long)j * sj];
  s = block_sum_f64(s);
  if (threadIdx.x == 0) part[(long)sp * cols + j] = s;
}

// The same for a row-major [rows][cols] matrix, cols a multiple of 4
// dividing 1,024 (the bias gradients of the convolutions and of the first
// dense layer): a workgroup takes whole rows, its threads read consecutive
// 16-byte groups (4 KB contiguous per step), and element i of thread t's
// group always falls in column (4t + i) % cols; the 1,024 slots are then
// added per column in slot order.
__global__ __launch_bounds__(256) void k_colsum_part(const float* __restrict__ G, long rows, int cols, long per,
                                                     double* __restrict__ part) {
  __shared__ double red[1024];
  const int sp = blockIdx.x;
  const long e0 = (long)sp * per * cols / 4, e1 = min(rows, (long)(sp + 1) * per) * cols / 4;
  const float4* G4 = reinterpret_cast<const float4*>(G);
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  for (long e = e0 + threadIdx.x; e < e1; e += 256) {
    const float4 v = G4[e];
    s0 += v.x;
    s1 += v.y;
    s2 += v.z;
    s3 += v.w;
  }
  red[4 * threadIdx.x] = s0;
  red[4 * threadIdx.x + 1] = s1;
  red[4 * threadIdx.x + 2] = s2;
  red[4 * threadIdx.x + 3] = s3;
  __syncthreads();
  if ((int)threadIdx.x < cols) {
    double t = 0.0;
    for (int k = threadIdx.x; k < 1024; k += cols) t += red[k];
    part[(long)sp * cols + threadIdx.x] = t;
  }
}

// The partial sums of one column added by one workgroup: thread t takes
// splits t, t + 256, ... in order, then the threads are added in a fixed
// tree.
__global__ __launch_bounds__(256) void k_colsum_fin(const double* __restrict__ part, int nsplit, int cols,
                                                    float* __restrict__ out) {
  const int j = blockIdx.x;
  double s = 0.0;
  for (int sp = threadIdx.x; sp < nsplit; sp += 256) s += part[(long)sp * cols + j];
  s = block_sum_f64(s);
  if (threadIdx.x == 0) out[j] = (float)s;
}

__global__ void k_reduce_rows_fin(const double* __restrict__ part, int nsplit, int cols, float* __restrict__ out) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= cols) return;
  double s = 0.0;
  for (int sp = 0; sp < nsplit; sp++) s += part[(long)sp * cols + j];
  out[j] = (float)s;
}

// Softmax cross-entropy over logits stored [C][B], one wave per pattern
// (lanes over classes).  targets: pattern-major [Nt][C], row tidx[b] (one
// contiguous row per pattern); grad = (p - t) * scale; loss accumulates
// -sum t log p.
__global__ __launch_bounds__(256) void k_softmax_xent(const float* __restrict__ logits, int C, int B,
                                                      const float* __restrict__ targets, long Nt,
                                                      const int* __restrict__ tidx, float scale,
                                                      float* __restrict__ grad, double* __restrict__ loss) {
  const int lane = threadIdx.x & 63;
  const int b = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (b >= B) return;
  float m = -3.0e38f;
  for (int c = lane; c < C; c += 64) m = fmaxf(m, logits[(long)c * B + b]);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
  double s = 0.0;
  for (int c = lane; c < C; c += 64) s += exp((double)logits[(long)c * B + b] - m);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  const double ls = log(s);
  const float* trow = targets + (long)tidx[b] * C;
  double l = 0.0;
  for (int c = lane; c < C; c += 64) {
    const double lp = (double)logits[(long)c * B + b] - m - ls;
    const float tc = trow[c];
    if (tc > 0.0f) l -= tc * lp;
    if (grad) grad[(long)c * B + b] = (float)((exp(lp) - tc) * scale);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o, 64);
  if (lane == 0 && loss) atomicAdd(loss, l);
}

// Adam on FP32 master weights (bias-corrected).
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       long n, float lr, float b1, float b2, float eps, float c1, float c2) {
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  const float gt = g[t];
  const float mt = b1 * m[t] + (1.0f - b1) * gt;
  const float vt = b2 * v[t] + (1.0f - b2) * gt * gt;
  m[t] = mt;
  v[t] = vt;
  p[t] -= lr * (mt / c1) / (sqrtf(vt / c2) + eps);
}

// The k largest logits of each pattern (logits [C][B]), descending, ties by
// lower index: one wave per pattern, k rounds of a wave-wide (value, index)
// maximum below the previous pick.  With mass != nullptr, mass[b] = softmax
// probability held by the k picks.
__global__ __launch_bounds__(256) void k_topk(const float* __restrict__ logits, int C, int B, int k,
                                              const int* __restrict__ out_rows, int* __restrict__ out,
                                              float* __restrict__ mass) {
  const int lane = threadIdx.x & 63;
  const int b = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (b >= B) return;
  const long row = out_rows ? out_rows[b] : b;
  float last_v = 3.0e38f;
  int last_i = -1;
  float top = 0.0f;
  double picked = 0.0;
  for (int j = 0; j < k; j++) {
    float bv = -3.0e38f;
    int bi = 0x7fffffff;
    for (int c = lane; c < C; c += 64) {
      const float v = logits[(long)c * B + b];
      const bool below = v < last_v || (v == last_v && c > last_i);
      if (below && (v > bv || (v == bv && c < bi))) {
        bv = v;
        bi = c;
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const float ov = __shfl_xor(bv, o, 64);
      const int oi = __shfl_xor(bi, o, 64);
      if (ov > bv || (ov == bv && oi < bi)) {
        bv = ov;
        bi = oi;
      }
    }
    if (bi == 0x7fffffff) bi = -1;
    if (lane == 0) out[row * k + j] = bi;
    if (j == 0) top = bv;
    if (bi >= 0) picked += exp((double)bv - (double)top);
    last_v = bv;
    last_i = bi;
  }
  if (mass) {
    double z = 0.0;
    for (int c = lane; c < C; c += 64) z += exp((double)logits[(long)c * B + b] - (double)top);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) z += __shfl_xor(z, o, 64);
    if (lane == 0) mass[row] = (float)(picked / z);
  }
}

// Coarse targets, pattern-major: T[i][c] = sum over the fine rotations r of
// cell c of P[r][i] (P in the ABI's [R][N] layout; cells as CSR lists of
// fine rotations).  A workgroup takes 64 patterns x 64 cells: lanes along
// patterns read the rows of P coalesced, the 64 x 64 sums are transposed in
// LDS and written along cells.
__global__ __launch_bounds__(256) void k_coarse_targets(const float* __restrict__ P, int N,
                                                        const int* __restrict__ cell_off,
                                                        const int* __restrict__ cell_rot, int C,
                                                        float* __restrict__ T) {
  __shared__ float tile[64][65];
  const int i0 = blockIdx.x * 64, c0 = blockIdx.y * 64;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int i = i0 + lane;
  for (int cc = w; cc < 64; cc += 4) {
    const int c = c0 + cc;
    double s = 0.0;
    if (c < C && i < N) {
      // 8 rotations' indices, then their responsibilities, loaded together;
      // added in rotation order
      const int j1 = cell_off[c + 1];
      int j = cell_off[c];
      for (; j + 8 <= j1; j += 8) {
        int r[8];
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) r[u] = cell_rot[j + u];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = P[(long)r[u] * N + i];
#pragma unroll
        for (int u = 0; u < 8; u++) s += v[u];
      }
      for (; j < j1; j++) s += P[(long)cell_rot[j] * N + i];
    }
    tile[lane][cc] = (float)s;
  }
  __syncthreads();
  for (int ii = w; ii < 64; ii += 4) {
    const int c = c0 + lane;
    if (i0 + ii < N && c < C) T[(long)(i0 + ii) * C + c] = tile[ii][lane];
  }
}

static inline unsigned nblk(long n) { return (unsigned)((n + 255) / 256); }

void cnn_input(const float* images, const int* mask, int N2d, const int* idx, int B, float* x, hipStream_t st) {
  hipLaunchKernelGGL(k_cnn_input, dim3(nblk((long)B * N2d)), dim3(256), 0, st, images, mask, N2d, idx, B, x);
  EMC_LAUNCHED();
}
void im2col(const float* in, int B, int H, int W, int C, float* col, hipStream_t st) {
  hipLaunchKernelGGL(k_im2col, dim3(nblk((long)B * (H / 2) * (W / 2) * 9 * C)), dim3(256), 0, st, in, B, H, W, C, col);
  EMC_LAUNCHED();
}
void conv0_forward(const float* in, int B, int H, const float* w, const float* bias, float* out, hipStream_t st) {
  if ((long)B * (H / 2) * (H / 2) >= (1L << 31)) {
    set_error("conv0_forward: %d x %d x %d output pixels exceed 32-bit indices", B, H / 2, H / 2);
    return;
  }
  hipLaunchKernelGGL(k_conv0_fwd, dim3(nblk((long)B * (H / 2) * (H / 2))), dim3(256), 0, st, in, B, H, w, bias, out);
  EMC_LAUNCHED();
}
bool conv0_wgrad(const float* dz, const float* in, int B, int H, float* dW, float* dB, double* scratch,
                 size_t scratch_doubles, hipStream_t st) {
  const long M = (long)B * (H / 2) * (H / 2);
  const long maxblk = (long)(scratch_doubles / (10 * kC0));
  if (M <= 0 || M >= (1L << 31) || maxblk < 1) {
    set_error("conv0_wgrad: %ld output pixels (32-bit indices) or no scratch", M);
    return false;
  }
  long per = 1024;  // 64 pixels per thread, thousands of workgroups
  if ((M + per - 1) / per > maxblk) per = (M + maxblk - 1) / maxblk;
  const long nb = (M + per - 1) / per;
  hipLaunchKernelGGL(k_conv0_wgrad, dim3((unsigned)nb), dim3(256), 0, st, dz, in, B, H, per, scratch);
  EMC_LAUNCHED();
  hipLaunchKernelGGL(k_conv0_wgrad_fin, dim3(10 * kC0), dim3(256), 0, st, scratch, (int)nb, dW, dB);
  EMC_LAUNCHED();
  return !failed();
}
// wtc by parity class: slot s (0..8) is class q's local tap j (slots 0 | 1 2
// | 3 4 | 5 6 7 8 for q = 0 .. 3), element (s*OC + o)*IC + c
__global__ void k_conv_weights_tc(const float* __restrict__ w, int OC, int IC, float* __restrict__ wtc) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= 9 * OC * IC) return;
  const int c = i % IC, o = (i / IC) % OC, s = i / (IC * OC);
  const int q = s == 0 ? 0 : s < 3 ? 1 : s < 5 ? 2 : 3;
  const int j = s - (q == 0 ? 0 : q == 1 ? 1 : q == 2 ? 3 : 5);
  int ky, kx;
  class_tap(q >> 1, q & 1, j, ky, kx);
  wtc[i] = w[((long)o * 9 + ky * 3 + kx) * IC + c];
}
void conv_weights_tc(const float* w, int OC, int IC, float* wtc, hipStream_t st) {
  hipLaunchKernelGGL(k_conv_weights_tc, dim3(nblk(9L * OC * IC)), dim3(256), 0, st, w, OC, IC, wtc);
  EMC_LAUNCHED();
}
void conv_weights_t(const float* w, int OC, int IC, float* wt, hipStream_t st) {
  hipLaunchKernelGGL(k_conv_weights_t, dim3(nblk(9L * OC * IC)), dim3(256), 0, st, w, OC, IC, wt);
  EMC_LAUNCHED();
}
void reduce_rows(const float* G, long rows, int cols, long si, long sj, float* out, double* scratch,
                 size_t scratch_doubles, hipStream_t st) {
  if (sj == 1 && si == cols && cols >= 4 && cols % 4 == 0 && 1024 % cols == 0 && rows > 0 &&
      ((uintptr_t)G & 15) == 0) {
    // row-major with rows of a length dividing 1,024: contiguous reads
    long per = std::max<long>(1, 65536 / cols);  // ~64K elements per workgroup
    long nsplit = (rows + per - 1) / per;
    if (nsplit * cols > (long)scratch_doubles) {
      nsplit = std::max<long>(1, (long)scratch_doubles / cols);
      per = (rows + nsplit - 1) / nsplit;
      nsplit = (rows + per - 1) / per;
    }
    hipLaunchKernelGGL(k_colsum_part, dim3((unsigned)nsplit), dim3(256), 0, st, G, rows, cols, per, scratch);
    EMC_LAUNCHED();
    hipLaunchKernelGGL(k_colsum_fin, dim3(cols), dim3(256), 0, st, scratch, (int)nsplit, cols, out);
    EMC_LAUNCHED();
    return;
  }
  long nsplit = std::max<long>(1, std::min<long>((rows + 8191) / 8192, (long)(scratch_doubles / std::max(cols, 1))));
  nsplit = std::min<long>(nsplit, 65535);
  const long per = (rows + nsplit - 1) / nsplit;
  hipLaunchKernelGGL(k_reduce_rows_part, dim3(cols, nsplit), dim3(256), 0, st, G, rows, cols, si, sj, per, scratch);
  EMC_LAUNCHED();
  hipLaunchKernelGGL(k_reduce_rows_fin, dim3((cols + 255) / 256), dim3(256), 0, st, scratch, (int)nsplit, cols, out);
  EMC_LAUNCHED();
}
void softmax_xent(const float* logits, int C, int B, const float* targets, long Nt, const int* tidx, float scale,
                  float* grad, double* loss, hipStream_t st) {
  hipLaunchKernelGGL(k_softmax_xent, dim3((B + 3) / 4), dim3(256), 0, st, logits, C, B, targets, Nt, tidx, scale, grad,
                     loss);
  EMC_LAUNCHED();
}
void adam(float* p, const float* g, float* m, float* v, long n, float lr, float b1, float b2, float eps, int step,
          hipStream_t st) {
  const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(nblk(n)), dim3(256), 0, st, p, g, m, v, n, lr, b1, b2, eps, c1, c2);
  EMC_LAUNCHED();
}
void topk(const float* logits, int C, int B, int k, const int* out_rows, int* out, float* mass, hipStream_t st) {
  hipLaunchKernelGGL(k_topk, dim3((B + 3) / 4), dim3(256), 0, st, logits, C, B, k, out_rows, out, mass);
  EMC_LAUNCHED();
}
void coarse_targets(const float* P, int N, const int* cell_off, const int* cell_rot, int C, float* T, hipStream_t st) {
  hipLaunchKernelGGL(k_coarse_targets, dim3((N + 63) / 64, (C + 63) / 64), dim3(256), 0, st, P, N, cell_off, cell_rot,
                     C, T);
  EMC_LAUNCHED();
}

}  // namespace emc
