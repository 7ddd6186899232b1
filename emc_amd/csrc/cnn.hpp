// Rotation-proposal network: GEMM descriptor, kernel launchers and the
// network object (weights, Adam state, activations).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "engine.hpp"

namespace emc {

// C(m, n) = sum_k A(m, k) B(n, k) [+ bias_n[n] + bias_m[m]] [ReLU] [+= C]
// with A(m, k) = A[m*sam + k*sak], B(n, k) = B[n*sbn + k*sbk],
// C(m, n) = C[m*scm + n*scn].  a_len, b_len, c_len: floats available at A,
// B and C (and bias_n >= N, bias_m >= M floats); gemm_bf16 checks every
// descriptor against them on the host and launches nothing out of bounds.
struct GemmDesc {
  const float* A = nullptr;
  long sam = 0, sak = 0;
  const float* B = nullptr;
  long sbn = 0, sbk = 0;
  float* C = nullptr;
  long scm = 0, scn = 0;
  const float* bias_n = nullptr;
  const float* bias_m = nullptr;
  int M = 0, N = 0, K = 0;
  int relu = 0;
  int accumulate = 0;
  long a_len = 0, b_len = 0, c_len = 0;
  // Implicit im2col of a 3x3 / stride-2 / pad-1 convolution over an NHWC
  // input [batch][cH][cW][cC] (cC a multiple of 8): conv_a reads A(m, k) =
  // in[b][2oy+ky-1][2ox+kx-1][c] with m = (b*OH + oy)*OW + ox and
  // k = (ky*3 + kx)*cC + c (sam, sak unused); conv_b reads B(n, k) the same
  // way with n = (ky*3 + kx)*cC + c and k = m.  Taps outside the input are 0.
  // conv_t: the transposed convolution's operand (input gradients): A(m, k)
  // with m = (b*cH + y)*cW + x a pixel of the convolution's input and
  // k = (ky*3 + kx)*cC + oc reads the output gradient dz[b][oy][ox][oc]
  // ([batch][cH/2][cW/2][cC]) at 2oy = y + 1 - ky, 2ox = x + 1 - kx, 0 where
  // those are odd or outside.
  int conv_a = 0, conv_b = 0, conv_t = 0, cH = 0, cW = 0, cC = 0;
  // tcls >= 0 (with conv_t): one parity class (py, px) = (tcls >> 1,
  // tcls & 1) of the input pixels, whose valid taps are ky = 1 (py = 0) or
  // 0, 2 (py = 1) and likewise kx: rows m = (b*cH/2 + yy)*cW/2 + xx are the
  // pixels (2yy + py, 2xx + px), k = j*cC + oc runs over the class's taps j
  // (ky-major), and C's row of m is that pixel's (b*cH + y)*cW + x.
  int tcls = -1;
  // gate (optional, C's layout): outputs where gate <= 0 are 0 (the ReLU's
  // backward of the activation gate holds)
  const float* gate = nullptr;
};

// False (error set, nothing launched) when the descriptor reaches outside
// a_len / b_len / c_len or has negative strides.
bool gemm_bf16(const GemmDesc& d, float* scratch, size_t scratch_floats, hipStream_t st);
void cnn_input(const float* images, const int* mask, int N2d, const int* idx, int B, float* x, hipStream_t st);
void im2col(const float* in, int B, int H, int W, int C, float* col, hipStream_t st);
// the first convolution when it has 1 input and 16 output channels: direct
// forward (bias, ReLU) into out [B][H/2][H/2][16], and its weight and bias
// gradients from dz [B][H/2][H/2][16] (scratch: partial sums)
void conv0_forward(const float* in, int B, int H, const float* w, const float* bias, float* out, hipStream_t st);
bool conv0_wgrad(const float* dz, const float* in, int B, int H, float* dW, float* dB, double* scratch,
                 size_t scratch_doubles, hipStream_t st);
// wt[(t*OC + o)*IC + c] = w[(o*9 + t)*IC + c]: conv weights [OC][3][3][IC]
// as [3][3][OC][IC], the B operand of the transposed convolution
void conv_weights_t(const float* w, int OC, int IC, float* wt, hipStream_t st);
// the same grouped by the parity classes of the input pixels: class q's
// taps (kTapClassN[q] of them, ky-major) from kTapClassOff[q] * OC * IC on
constexpr int kTapClassN[4] = {1, 2, 2, 4};
constexpr int kTapClassOff[4] = {0, 1, 3, 5};
void conv_weights_tc(const float* w, int OC, int IC, float* wtc, hipStream_t st);
// out[j] = sum over rows i of G[i*si + j*sj]; scratch holds the per-split
// partial sums (cols * splits doubles)
void reduce_rows(const float* G, long rows, int cols, long si, long sj, float* out, double* scratch,
                 size_t scratch_doubles, hipStream_t st);
// targets pattern-major [Nt][C]; pattern b of the batch takes row tidx[b]
void softmax_xent(const float* logits, int C, int B, const float* targets, long Nt, const int* tidx, float scale,
                  float* grad, double* loss, hipStream_t st);
void adam(float* p, const float* g, float* m, float* v, long n, float lr, float b1, float b2, float eps, int step,
          hipStream_t st);
void topk(const float* logits, int C, int B, int k, const int* out_rows, int* out, float* mass, hipStream_t st);
// T [N][C] (pattern-major) = per-cell sums of P [R][N]
void coarse_targets(const float* P, int N, const int* cell_off, const int* cell_rot, int C, float* T, hipStream_t st);

// The network.  Parameters in one FP32 buffer:
//   conv l (l = 0..3): W [OC][3][3][IC], b [OC]; dense1 W [H][F], b [H];
//   dense2 W [C][H], b [C]  (F = 64 * (S/16)^2 flattened NHWC features).
struct Cnn {
  int side = 0, ncell = 0, hidden = 256, batch = 0;
  static constexpr int kConvs = 4;
  int ic[kConvs] = {1, 16, 32, 64};
  int oc[kConvs] = {16, 32, 64, 64};
  int hin[kConvs] = {0, 0, 0, 0};  // input height of conv l
  long off_w[kConvs + 2] = {0}, off_b[kConvs + 2] = {0};
  long nparam = 0;
  int feat = 0;
  int step = 0;
  DBuf params, grads, adam_m, adam_v;
  // activations for one batch
  DBuf x, col[kConvs], act[kConvs], h, logits;
  DBuf dact, dh, dlogits, scratch, idx, loss, red, wt;
  size_t scratch_floats = 0;

  bool init(int side, int ncell, int batch, uint64_t seed);
  void release();
  // forward for patterns idx_host[0..B) (B <= batch) into logits [C][B]
  void forward(const float* images, const int* mask, int N2d, const int* d_idx, int B, hipStream_t st);
  // backward from dlogits [C][B] and Adam step
  void backward_and_step(int B, float lr, hipStream_t st);
  float* W(int l) const { return params.as<float>() + off_w[l]; }
  float* Bv(int l) const { return params.as<float>() + off_b[l]; }
  float* dW(int l) const { return grads.as<float>() + off_w[l]; }
  float* dB(int l) const { return grads.as<float>() + off_b[l]; }
  // conv l reads its input through the GEMM's implicit im2col (input
  // channels a multiple of 8: every layer but the first)
  bool implicit(int l) const { return l > 0 && ic[l] % 8 == 0; }
  bool direct0() const { return ic[0] == 1 && oc[0] == 16; }  // first convolution by conv0_forward / conv0_wgrad
};

}  // namespace emc
