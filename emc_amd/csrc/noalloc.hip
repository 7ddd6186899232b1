// Model normalisation that allocates nothing: the kernels of
// cuda_normalize_model (per-block FP64 partial sums of model/weight, the
// mean factor, the scaling pass) with their scratch in a file-scope device
// array instead of a buffer the library allocates on first use.  For callers
// that cannot call hipMalloc (code under stream capture, tight serving
// loops).  No counterpart in the reference.
#include <hip/hip_runtime.h>

#include <climits>

#include "engine.hpp"

namespace emc {

// Scratch of the normalisation: kNormMaxBlocks per-block FP64 partial sums
// followed by the float factor (two doubles of room), 8,208 bytes per
// device.  The partial-sum grid is capped at kNormMaxBlocks blocks, so every
// model size an int can state fits; every call rewrites all of it that it
// reads (partials of its own blocks, then the factor), so nothing needs to
// start at zero.
constexpr int kNoallocScratch = kNormMaxBlocks + 2;
static_assert(kNoallocScratch >= kNormMaxBlocks + 2, "normalisation scratch too small");
static __device__ __align__(256) double scratch[kNoallocScratch];

// Device address of `scratch` on each device, resolved once per device with
// that device current.  The array's name in host code is only a stand-in
// for the symbol: kernels and copies get these addresses.
constexpr int kMaxDevices = 64;
static double* g_scratch_ptr[kMaxDevices];

double* noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) {
    set_error("normalize (no allocation): device %d outside 0..%d", device, kMaxDevices - 1);
    return nullptr;
  }
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (err != hipSuccess || !p) {
      (void)hipGetLastError();
      set_error("normalize (no allocation): hipGetSymbolAddress of the scratch failed on device %d: %s", device,
                hipGetErrorString(err));
      return nullptr;
    }
    g_scratch_ptr[device] = static_cast<double*>(p);
  }
  return g_scratch_ptr[device];
}

}  // namespace emc

using namespace emc;

extern "C" int emc_hip_normalize_model_noalloc(float* d_model, const float* d_weight, int model_size) {
  if (!ready()) return -1;
  Engine& e = engine();
  if (!d_model || !d_weight || model_size < 0) {
    set_error("emc_hip_normalize_model_noalloc: invalid argument (model_size %d; the scratch takes 0 to %d voxels)",
              model_size, INT_MAX);
    return -1;
  }
  if (model_size == 0) return 0;
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) {
    (void)hipGetLastError();
    set_error("emc_hip_normalize_model_noalloc: no current device");
    return -1;
  }
  double* red = noalloc_scratch(dev);
  if (!red) return -1;
  launch_normalize_model(d_model, d_weight, model_size, red, e.stream);
  EMC_CHECK(hipStreamSynchronize(e.stream));
  return failed() ? -1 : 0;
}
