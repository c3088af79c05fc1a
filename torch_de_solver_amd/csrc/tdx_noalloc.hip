// No-allocation entry point of the 'NN' fused training step.
//
// tdx_nn_train_step takes its workspace from the caller.  Callers that
// cannot allocate device memory at all (code under HIP graph capture that
// must not call hipMalloc, tight serving loops, harnesses that forbid it)
// use tdx_nn_train_step_noalloc instead: the same kernels, with the
// workspace in the file-scope device array `scratch` below.
//
// `scratch` is a host-side stand-in in host code, not a device address: its
// device address is looked up once per device with hipGetSymbolAddress
// (each GPU holds its own copy of the array until the process ends), and
// only that address is handed to the kernels.  Every call on one GPU shares
// that GPU's array, so calls that overlap on one GPU (several streams or
// threads at once) race; calls in stream order do not.  The kernels write
// every workspace word they read within the call, so nothing needs clearing
// between calls.
#include "tdx_common.h"

#include <mutex>

constexpr size_t kScratchBytes = size_t(64) << 20;  // 64 MiB per GPU

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_ptr[kMaxDevices] = {nullptr};

// Device address of `scratch` on the current device (looked up on the first
// call there), or nullptr with the error message set.
void *scratch_address() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    tdx::set_error("no current HIP device for the no-allocation entry point");
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_ptr[dev]) {
    void *ptr = nullptr;
    const hipError_t rc = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (rc != hipSuccess || !ptr) {
      tdx::set_error("hipGetSymbolAddress of the scratch array failed: %s",
                     hipGetErrorString(rc));
      return nullptr;
    }
    g_ptr[dev] = ptr;
  }
  return g_ptr[dev];
}

}  // namespace

extern "C" size_t tdx_noalloc_scratch_bytes(void) { return kScratchBytes; }

extern "C" int tdx_nn_train_step_noalloc(const tdx_nn_plan *p, void *params, void *exp_avg,
                                         void *exp_avg_sq, int64_t *step, double lr,
                                         double beta1, double beta2, double eps, void *loss,
                                         tdx_stream_t stream) {
  TDX_REQUIRE(p && params && exp_avg && exp_avg_sq && step, "null argument");
  const size_t need = tdx_nn_workspace_bytes(p);
  TDX_REQUIRE(need <= kScratchBytes,
              "the plan needs %zu bytes of workspace; the no-allocation entry point holds "
              "%zu (use tdx_nn_train_step with a caller workspace)",
              need, kScratchBytes);
  void *ws = scratch_address();
  if (!ws) return TDX_ELAUNCH;
  return tdx_nn_train_step(p, params, exp_avg, exp_avg_sq, step, lr, beta1, beta2, eps, loss,
                           ws, stream);
}
