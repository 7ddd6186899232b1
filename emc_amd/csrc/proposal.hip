// Learned rotation proposal: the network's C-ABI, training on EMC's own
// responsibilities, prediction of candidate coarse cells, and the
// candidate-restricted iteration (E-step, softmax, audit, scaling, M-step).
// None of this has a counterpart in the reference.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "cnn.hpp"
#include "emc_host.h"
#include "engine.hpp"
#include "proposal.hpp"

namespace emc {

Proposal& proposal() {
  static Proposal* p = new Proposal();
  return *p;
}

}  // namespace emc

using namespace emc;

extern "C" {

// ---------------------------------------------------------------------------
// network C-ABI (tests and tools)
// ---------------------------------------------------------------------------
int emc_hip_cnn_init(int side, int ncell, int batch, unsigned long long seed) {
  if (!ready()) return 1;
  return proposal().cnn.init(side, ncell, batch, seed) ? 0 : 1;
}

long emc_hip_cnn_param_count(void) { return proposal().cnn.nparam; }

int emc_hip_cnn_get(float* params_host, float* grads_host) {
  if (!ready()) return 1;
  Cnn& n = proposal().cnn;
  if (!n.nparam) return 1;
  if (params_host) EMC_CHECK(hipMemcpy(params_host, n.params.p, n.nparam * 4, hipMemcpyDeviceToHost));
  if (grads_host) EMC_CHECK(hipMemcpy(grads_host, n.grads.p, n.nparam * 4, hipMemcpyDeviceToHost));
  return failed() ? 1 : 0;
}

int emc_hip_cnn_set(const float* params_host) {
  if (!ready()) return 1;
  Cnn& n = proposal().cnn;
  if (!n.nparam || !params_host) return 1;
  EMC_CHECK(hipMemcpy(n.params.p, params_host, n.nparam * 4, hipMemcpyHostToDevice));
  return failed() ? 1 : 0;
}

int emc_hip_cnn_forward(const real* d_images, const int* d_mask, int N2d, const int* idx_host, int B,
                        float* logits_host) {
  if (!ready()) return 1;
  Cnn& n = proposal().cnn;
  Engine& e = engine();
  if (!n.nparam || B <= 0 || B > n.batch || N2d != n.side * n.side) {
    set_error("emc_hip_cnn_forward: invalid argument");
    return 1;
  }
  EMC_CHECK(hipMemcpyAsync(n.idx.p, idx_host, sizeof(int) * B, hipMemcpyHostToDevice, e.stream));
  n.forward(d_images, d_mask, N2d, n.idx.as<int>(), B, e.stream);
  if (logits_host)
    EMC_CHECK(hipMemcpyAsync(logits_host, n.logits.p, sizeof(float) * (size_t)n.ncell * B, hipMemcpyDeviceToHost,
                             e.stream));
  EMC_CHECK(hipStreamSynchronize(e.stream));
  return failed() ? 1 : 0;
}

// One training step on B patterns: targets_host is [C][B]; lr = 0 leaves the
// weights unchanged (gradients are still computed and readable).
int emc_hip_cnn_step(const real* d_images, const int* d_mask, int N2d, const int* idx_host, int B,
                     const float* targets_host, float lr, double* loss) {
  if (!ready()) return 1;
  Cnn& n = proposal().cnn;
  Engine& e = engine();
  if (!n.nparam || B <= 0 || B > n.batch || N2d != n.side * n.side || !targets_host) {
    set_error("emc_hip_cnn_step: invalid argument");
    return 1;
  }
  DBuf t, ti;
  if (!t.ensure(sizeof(float) * (size_t)n.ncell * B) || !ti.ensure(sizeof(int) * B)) return 1;
  std::vector<int> iota(B);
  std::iota(iota.begin(), iota.end(), 0);
  EMC_CHECK(hipMemcpyAsync(t.p, targets_host, sizeof(float) * (size_t)n.ncell * B, hipMemcpyHostToDevice, e.stream));
  EMC_CHECK(hipMemcpyAsync(ti.p, iota.data(), sizeof(int) * B, hipMemcpyHostToDevice, e.stream));
  EMC_CHECK(hipMemcpyAsync(n.idx.p, idx_host, sizeof(int) * B, hipMemcpyHostToDevice, e.stream));
  n.forward(d_images, d_mask, N2d, n.idx.as<int>(), B, e.stream);
  launch_fill(n.loss.as<float>(), 2, 0.0f, e.stream);
  softmax_xent(n.logits.as<float>(), n.ncell, B, t.as<float>(), B, ti.as<int>(), 1.0f / B, n.dlogits.as<float>(),
               n.loss.as<double>(), e.stream);
  n.backward_and_step(B, lr, e.stream);
  double l = 0;
  EMC_CHECK(hipMemcpyAsync(&l, n.loss.p, 8, hipMemcpyDeviceToHost, e.stream));
  EMC_CHECK(hipStreamSynchronize(e.stream));
  if (loss) *loss = l / B;
  t.release();
  ti.release();
  return failed() ? 1 : 0;
}

}  // extern "C"
