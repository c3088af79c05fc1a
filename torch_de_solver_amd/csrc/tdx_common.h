// Internal helpers shared by the tdx translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <mutex>
#include <set>
#include <string>
#include <utility>
#include <vector>

#include "tdx.h"
#include "tdx_dev.h"

namespace tdx {

void set_error(const char *fmt, ...);
void clear_error();

#define TDX_TRY_HIP(expr)                                                    \
  do {                                                                       \
    hipError_t _e = (expr);                                                  \
    if (_e != hipSuccess) {                                                  \
      ::tdx::set_error("%s failed: %s", #expr, hipGetErrorString(_e));       \
      return TDX_ELAUNCH;                                                    \
    }                                                                        \
  } while (0)

#define TDX_REQUIRE(cond, ...)                                               \
  do {                                                                       \
    if (!(cond)) {                                                           \
      ::tdx::set_error(__VA_ARGS__);                                         \
      return TDX_EINVAL;                                                     \
    }                                                                        \
  } while (0)

// Device copy of an immutable host table, owned by a plan.
struct DevTable {
  void *ptr = nullptr;
  size_t bytes = 0;
  int upload(const void *host, size_t n) {
    bytes = n;
    if (n == 0) return TDX_OK;
    if (hipMalloc(&ptr, n) != hipSuccess) {
      set_error("hipMalloc of %zu bytes for a plan table failed", n);
      return TDX_ENOMEM;
    }
    if (hipMemcpy(ptr, host, n, hipMemcpyHostToDevice) != hipSuccess) {
      set_error("hipMemcpy of a plan table failed");
      return TDX_ELAUNCH;
    }
    return TDX_OK;
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
  }
};

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Raises the dynamic LDS limit of each given kernel to 160 KB, once per
// (kernel, device) and process; safe to call from several threads.
inline void raise_lds_limit(std::initializer_list<const void *> kernels) {
  static std::mutex mu;
  static std::set<std::pair<const void *, int>> done;
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lock(mu);
  for (const void *k : kernels)
    if (done.insert({k, dev}).second)
      (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
}

template <typename T> inline int crow_host(int q, int i) {
  return sizeof(T) == 4 ? 4 * q + i : q + 4 * i;
}

struct AdamHyper {
  double lr, b1, b2, eps;
};

// torch.optim.Adam (single-tensor form): m = b1 m + (1-b1) g;
// v = b2 v + (1-b2) g^2; p -= lr/(1-b1^t) * m / (sqrt(v)/sqrt(1-b2^t) + eps)
// The per-step scalars (bias corrections in double, as torch computes them
// in Python floats) depend only on the step t.
template <typename T>
struct AdamScalars {
  T w1, w2, b2, step_size, bc2s, eps;
};
template <typename T>
__device__ __forceinline__ AdamScalars<T> adam_scalars(int64_t t, const AdamHyper &h) {
  AdamScalars<T> s;
  s.b2 = T(h.b2);
  s.w1 = T(1.0 - h.b1);
  s.w2 = T(1.0 - h.b2);
  const double bc1 = 1.0 - pow(h.b1, (double)t);
  const double bc2 = 1.0 - pow(h.b2, (double)t);
  s.step_size = T(h.lr / bc1);
  s.bc2s = T(sqrt(bc2));
  s.eps = T(h.eps);
  return s;
}
template <typename T>
__device__ __forceinline__ void adam_apply(T &p, T g, T &m, T &v, const AdamScalars<T> &s) {
  m = m + s.w1 * (g - m);
  v = v * s.b2 + s.w2 * g * g;
  const T denom = sqrt(v) / s.bc2s + s.eps;
  p = p - s.step_size * (m / denom);
}

}  // namespace tdx
