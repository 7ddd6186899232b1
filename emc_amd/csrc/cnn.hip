// The rotation-proposal network: weights, Adam state, forward, backward.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "cnn.hpp"
#include "emc_host.h"

namespace emc {

static constexpr size_t kRedDoubles = 1 << 16;  // partial sums of the bias gradients

static long floats(const DBuf& b) { return (long)(b.bytes / sizeof(float)); }

bool Cnn::init(int side_, int ncell_, int batch_, uint64_t seed) {
  release();
  if (side_ < 16 || side_ % 16 || ncell_ <= 0 || batch_ <= 0) {
    set_error("proposal network: side %d must be a multiple of 16, %d cells, batch %d", side_, ncell_, batch_);
    return false;
  }
  side = side_;
  ncell = ncell_;
  batch = batch_;
  step = 0;
  for (int l = 0; l < kConvs; l++) hin[l] = side >> l;
  feat = oc[kConvs - 1] * (side / 16) * (side / 16);
  long off = 0;
  for (int l = 0; l < kConvs; l++) {
    off_w[l] = off;
    off += (long)oc[l] * 9 * ic[l];
    off_b[l] = off;
    off += oc[l];
  }
  off_w[4] = off;
  off += (long)hidden * feat;
  off_b[4] = off;
  off += hidden;
  off_w[5] = off;
  off += (long)ncell * hidden;
  off_b[5] = off;
  off += ncell;
  nparam = off;
  // He-normal weights (std sqrt(2 / fan_in); the output layer sqrt(1 /
  // fan_in)), zero biases, from the counter-based generator
  std::vector<float> p(nparam, 0.0f);
  auto normal = [&](uint64_t k) {
    const double u1 = emc_uniform(seed, 77, k, 0), u2 = emc_uniform(seed, 77, k, 1);
    return std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(2.0 * M_PI * u2);
  };
  for (int l = 0; l < kConvs + 2; l++) {
    const long n = off_b[l] - off_w[l];
    const long fan_in = l < kConvs ? 9L * ic[l] : (l == 4 ? feat : hidden);
    const double sd = std::sqrt((l == 5 ? 1.0 : 2.0) / (double)fan_in);
    for (long k = 0; k < n; k++) p[off_w[l] + k] = (float)(sd * normal((uint64_t)(off_w[l] + k)));
  }
  if (!params.ensure(nparam * 4) || !grads.ensure(nparam * 4) || !adam_m.ensure(nparam * 4) ||
      !adam_v.ensure(nparam * 4))
    return false;
  EMC_CHECK(hipMemcpy(params.p, p.data(), nparam * 4, hipMemcpyHostToDevice));
  EMC_CHECK(hipMemset(adam_m.p, 0, nparam * 4));
  EMC_CHECK(hipMemset(adam_v.p, 0, nparam * 4));
  EMC_CHECK(hipMemset(grads.p, 0, nparam * 4));
  // activations for one batch
  const long B = batch;
  long maxact = 0, maxw = 0;
  if (!x.ensure(4 * B * side * side)) return false;
  for (int l = 0; l < kConvs; l++) {
    const long oh = hin[l] / 2;
    const long m = B * oh * oh;
    // layers whose input has a multiple of 8 channels read it through the
    // GEMM's implicit im2col; only the others keep an im2col buffer
    const bool need_col = !implicit(l) && !(l == 0 && direct0());
    if ((need_col && !col[l].ensure(4 * m * 9 * ic[l])) || !act[l].ensure(4 * m * oc[l])) return false;
    maxact = std::max(maxact, m * oc[l]);
    maxw = std::max(maxw, 9L * oc[l] * ic[l]);
  }
  scratch_floats = std::max<size_t>((size_t)16 << 20, (size_t)4 * hidden * feat);
  if (!h.ensure(4 * B * hidden) || !logits.ensure(4 * (long)ncell * B) || !dact.ensure(4 * 2 * maxact) ||
      !wt.ensure(4 * maxw) || !dh.ensure(4 * B * hidden) || !dlogits.ensure(4 * (long)ncell * B) ||
      !scratch.ensure(4 * scratch_floats) || !idx.ensure(4 * B) || !loss.ensure(8) ||
      !red.ensure(8 * kRedDoubles))
    return false;
  return !failed();
}

void Cnn::release() {
  DBuf* bufs[] = {&params, &grads, &adam_m, &adam_v, &x, &h, &logits, &dact, &wt, &dh, &dlogits, &scratch, &idx, &loss,
                  &red};
  for (DBuf* b : bufs) b->release();
  for (int l = 0; l < kConvs; l++) {
    col[l].release();
    act[l].release();
  }
}

void Cnn::forward(const float* images, const int* mask, int N2d, const int* d_idx, int B, hipStream_t st) {
  float* sc = scratch.as<float>();
  cnn_input(images, mask, N2d, d_idx, B, x.as<float>(), st);
  const float* in = x.as<float>();
  for (int l = 0; l < kConvs; l++) {
    const int H = hin[l], oh = H / 2, K = 9 * ic[l];
    if (l == 0 && direct0()) {
      conv0_forward(in, B, H, W(0), Bv(0), act[0].as<float>(), st);
      in = act[0].as<float>();
      continue;
    }
    GemmDesc d;
    if (implicit(l)) {
      d.A = in; d.conv_a = 1; d.cH = H; d.cW = H; d.cC = ic[l]; d.a_len = floats(act[l - 1]);
    } else {
      im2col(in, B, H, H, ic[l], col[l].as<float>(), st);
      d.A = col[l].as<float>(); d.sam = K; d.sak = 1; d.a_len = floats(col[l]);
    }
    d.B = W(l); d.sbn = K; d.sbk = 1; d.b_len = off_b[l] - off_w[l];
    d.C = act[l].as<float>(); d.scm = oc[l]; d.scn = 1; d.c_len = floats(act[l]);
    d.bias_n = Bv(l);
    d.M = B * oh * oh; d.N = oc[l]; d.K = K;
    d.relu = 1;
    gemm_bf16(d, sc, scratch_floats, st);
    in = act[l].as<float>();
  }
  GemmDesc d1;  // dense 1: h[b][j] = relu(sum_k a3[b][k] W4[j][k] + b4[j])
  d1.A = act[kConvs - 1].as<float>(); d1.sam = feat; d1.sak = 1; d1.a_len = floats(act[kConvs - 1]);
  d1.B = W(4); d1.sbn = feat; d1.sbk = 1; d1.b_len = off_b[4] - off_w[4];
  d1.C = h.as<float>(); d1.scm = hidden; d1.scn = 1; d1.c_len = floats(h);
  d1.bias_n = Bv(4);
  d1.M = B; d1.N = hidden; d1.K = feat; d1.relu = 1;
  gemm_bf16(d1, sc, scratch_floats, st);
  GemmDesc d2;  // dense 2, stored [C][B]: logits[c][b] = sum_k W5[c][k] h[b][k] + b5[c]
  d2.A = W(5); d2.sam = hidden; d2.sak = 1; d2.a_len = off_b[5] - off_w[5];
  d2.B = h.as<float>(); d2.sbn = hidden; d2.sbk = 1; d2.b_len = floats(h);
  d2.C = logits.as<float>(); d2.scm = B; d2.scn = 1; d2.c_len = floats(logits);
  d2.bias_m = Bv(5);
  d2.M = ncell; d2.N = B; d2.K = hidden;
  gemm_bf16(d2, sc, scratch_floats, st);
}

void Cnn::backward_and_step(int B, float lr, hipStream_t st) {
  float* sc = scratch.as<float>();
  const float* dl = dlogits.as<float>();
  {  // dW5[c][k] = sum_b dlog[c][b] h[b][k]
    GemmDesc d;
    d.A = dl; d.sam = B; d.sak = 1; d.a_len = floats(dlogits);
    d.B = h.as<float>(); d.sbn = 1; d.sbk = hidden; d.b_len = floats(h);
    d.C = dW(5); d.scm = hidden; d.scn = 1; d.c_len = off_b[5] - off_w[5];
    d.M = ncell; d.N = hidden; d.K = B;
    gemm_bf16(d, sc, scratch_floats, st);
    reduce_rows(dl, B, ncell, 1, B, dB(5), red.as<double>(), kRedDoubles, st);
  }
  {  // dh[b][k] = sum_c dlog[c][b] W5[c][k], gated by ReLU' (h > 0)
    GemmDesc d;
    d.A = dl; d.sam = 1; d.sak = B; d.a_len = floats(dlogits);
    d.B = W(5); d.sbn = 1; d.sbk = hidden; d.b_len = off_b[5] - off_w[5];
    d.C = dh.as<float>(); d.scm = hidden; d.scn = 1; d.c_len = floats(dh);
    d.gate = h.as<float>();
    d.M = B; d.N = hidden; d.K = ncell;
    gemm_bf16(d, sc, scratch_floats, st);
  }
  float* da = dact.as<float>();
  const long half = dact.bytes / 8;  // two ping-pong halves
  float* dbufs[2] = {da, da + half};
  int cur = 0;
  {  // dW4[j][k] = sum_b dh[b][j] a3[b][k]; da3[b][k] = sum_j dh[b][j] W4[j][k] gated by a3 > 0
    GemmDesc d;
    d.A = dh.as<float>(); d.sam = 1; d.sak = hidden; d.a_len = floats(dh);
    d.B = act[kConvs - 1].as<float>(); d.sbn = 1; d.sbk = feat; d.b_len = floats(act[kConvs - 1]);
    d.C = dW(4); d.scm = feat; d.scn = 1; d.c_len = off_b[4] - off_w[4];
    d.M = hidden; d.N = feat; d.K = B;
    gemm_bf16(d, sc, scratch_floats, st);
    reduce_rows(dh.as<float>(), B, hidden, hidden, 1, dB(4), red.as<double>(), kRedDoubles, st);
    GemmDesc e;
    e.A = dh.as<float>(); e.sam = hidden; e.sak = 1; e.a_len = floats(dh);
    e.B = W(4); e.sbn = 1; e.sbk = feat; e.b_len = off_b[4] - off_w[4];
    e.C = dbufs[cur]; e.scm = feat; e.scn = 1; e.c_len = half;
    e.gate = act[kConvs - 1].as<float>();
    e.M = B; e.N = feat; e.K = hidden;
    gemm_bf16(e, sc, scratch_floats, st);
  }
  for (int l = kConvs - 1; l >= 0; l--) {
    const int H = hin[l], oh = H / 2, K = 9 * ic[l], OC = oc[l];
    const long M = (long)B * oh * oh;
    float* dz = dbufs[cur];  // already gated by act[l] > 0
    if (l == 0 && direct0()) {
      // partial sums in the GEMMs' split-k scratch (free here: stream order)
      conv0_wgrad(dz, x.as<float>(), B, H, dW(0), dB(0), reinterpret_cast<double*>(sc), scratch_floats / 2, st);
      continue;
    }
    GemmDesc d;  // dW[oc][kk] = sum_m dz[m][oc] col[m][kk]
    d.A = dz; d.sam = 1; d.sak = OC; d.a_len = half;
    if (implicit(l)) {
      d.B = act[l - 1].as<float>(); d.conv_b = 1; d.cH = H; d.cW = H; d.cC = ic[l]; d.b_len = floats(act[l - 1]);
    } else {
      d.B = col[l].as<float>(); d.sbn = 1; d.sbk = K; d.b_len = floats(col[l]);
    }
    d.C = dW(l); d.scm = K; d.scn = 1; d.c_len = off_b[l] - off_w[l];
    d.M = OC; d.N = K; d.K = (int)M;
    gemm_bf16(d, sc, scratch_floats, st);
    reduce_rows(dz, M, OC, OC, 1, dB(l), red.as<double>(), kRedDoubles, st);
    if (l > 0) {
      // input gradient as the transposed convolution:
      // dx[b][y][x][c] = sum over taps and oc of dz at the output pixel the
      // tap covers x W[oc][tap][c], gated by act[l - 1] > 0 (the previous
      // layer's ReLU')
      // by the parity classes of the input pixels: each class has 1, 2 or 4
      // valid taps of the 9 (the rest land between output pixels)
      conv_weights_tc(W(l), OC, ic[l], wt.as<float>(), st);
      for (int q = 0; q < 4; q++) {
        GemmDesc e;
        e.A = dz; e.conv_t = 1; e.tcls = q; e.cH = H; e.cW = H; e.cC = OC; e.a_len = half;
        e.B = wt.as<float>() + (long)kTapClassOff[q] * OC * ic[l]; e.sbn = 1; e.sbk = ic[l];
        e.b_len = (long)kTapClassN[q] * OC * ic[l];
        e.C = dbufs[cur ^ 1]; e.scm = ic[l]; e.scn = 1; e.c_len = half;
        e.gate = act[l - 1].as<float>();
        e.M = B * oh * oh; e.N = ic[l]; e.K = kTapClassN[q] * OC;
        gemm_bf16(e, sc, scratch_floats, st);
      }
      cur ^= 1;
    }
  }
  step++;
  adam(params.as<float>(), grads.as<float>(), adam_m.as<float>(), adam_v.as<float>(), nparam, lr, 0.9f, 0.999f, 1e-8f,
       step, st);
}

}  // namespace emc
